"""Gauss smearing (lib/src/qcd_smearing.c:25-199) and vector BLAS
(lib/src/qcd_blas.c:237-491) on the GPU against the CPU oracle, through the
device API and the drop-in host-struct API."""
import ctypes

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

L = (8, 6, 6, 6)


def test_gauss_one_slice_and_all(gpu):
    Q = gpu
    u = Q.gauge_generate(L, 0.25, 71)
    v = Q.spinor_gaussian(L, 72)
    geo = O.Geometry(L)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(u)
    for t, niter in [(3, 1), (0, 5), (7, 4), (-1, 3)]:
        ref = v.copy()
        for _ in range(niter):
            ref = O.gauss3d(geo, ref, u, 4.0, t) if t >= 0 else O.gauss3d_all(geo, ref, u, 4.0)
        f = Q.Spinor(lat).upload(v)
        Q.gauss_smear(f, g, 4.0, t, niter)
        assert rel_l2(f.download(), ref) <= 1e-13


def test_gauss_point_source_50_iterations(gpu):
    """tests/test_smearing.c:82-116 pattern: a point source, 50 iterations, alpha 4."""
    Q = gpu
    u = Q.gauge_generate(L, 0.3, 73)
    src = Q.point_source(L, (2, 1, 3, 0), 1, 2)
    ref = src.copy()
    geo = O.Geometry(L)
    for _ in range(50):
        ref = O.gauss3d(geo, ref, u, 4.0, 2)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(u)
    f = Q.Spinor(lat).upload(src)
    Q.gauss_smear(f, g, 4.0, 2, 50)
    assert rel_l2(f.download(), ref) <= 1e-13


@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 2e-6), ("f16", 4e-3)])
@pytest.mark.parametrize("rows", [None, "2"])
def test_gauss_batched_rhs(gpu, monkeypatch, prec, tol, rows):
    """Every right-hand side of a field smeared in one call, on a lattice whose
    spatial extents all differ, one slice and all slices, odd and even iteration
    counts; rows = 2 splits the five right-hand sides over three passes."""
    Q = gpu
    if rows:
        monkeypatch.setenv("QCD_HIP_GAUSS_ROWS", rows)
    Lb = (6, 4, 6, 8)
    u = Q.gauge_generate(Lb, 0.25, 76)
    geo = O.Geometry(Lb)
    lat = Q.Lattice(Lb)
    p = Q.PREC[prec]
    g = Q.Gauge(lat, p).upload(u)
    vs = [Q.spinor_gaussian(Lb, 760 + r) for r in range(5)]
    for t, niter in [(4, 3), (1, 6), (-1, 2)]:
        f = Q.Spinor(lat, p, nrhs=5)
        for r, v in enumerate(vs):
            f.upload(v, rhs=r)
        Q.gauss_smear(f, g, 3.0, t, niter)
        for r, v in enumerate(vs):
            ref = v.copy()
            for _ in range(niter):
                ref = O.gauss3d(geo, ref, u, 3.0, t) if t >= 0 else O.gauss3d_all(geo, ref, u, 3.0)
            assert rel_l2(f.download(rhs=r), ref) <= tol, (t, niter, r)


def test_gauss_slices_of_growing_size(gpu):
    """The library keeps the slice buffers and the slice's links between
    calls; a later call on a larger slice must get larger buffers."""
    Q = gpu
    for k, Ls in enumerate([(4, 4, 2, 2), (4, 8, 8, 6), (6, 4, 4, 4), (4, 10, 8, 8)]):
        u = Q.gauge_generate(Ls, 0.25, 780 + k)
        v = Q.spinor_gaussian(Ls, 790 + k)
        lat = Q.Lattice(Ls)
        g = Q.Gauge(lat).upload(u)
        f = Q.Spinor(lat).upload(v)
        Q.gauss_smear(f, g, 2.0, 1, 2)
        geo = O.Geometry(Ls)
        ref = O.gauss3d(geo, O.gauss3d(geo, v, u, 2.0, 1), u, 2.0, 1)
        assert rel_l2(f.download(), ref) <= 1e-13, Ls


def test_dropin_gauss(gpu):
    Q = gpu
    u = Q.gauge_generate(L, 0.25, 74)
    v = Q.spinor_gaussian(L, 75)
    di = Q.DropIn(L)
    try:
        uu, vv = di.gauge(u), di.vector(v)
        lib = Q.lib()
        assert lib.qcd_gaussIteration3d(ctypes.byref(vv), ctypes.byref(uu), 2.5, 5) == 0
        assert rel_l2(di.view(vv, 12).reshape(-1, 4, 3).copy(), O.gauss3d(O.Geometry(L), v, u, 2.5, 5)) <= 1e-13
        w = di.view(vv, 12).reshape(-1, 4, 3).copy()
        assert lib.qcd_gaussIteration3dAll(ctypes.byref(vv), ctypes.byref(uu), 2.5, 1) == 0
        assert rel_l2(di.view(vv, 12).reshape(-1, 4, 3).copy(), O.gauss3d_all(O.Geometry(L), w, u, 2.5)) <= 1e-13
        # a slice not on this rank leaves the vector unchanged (qcd_smearing.c:49)
        before = di.view(vv, 12).copy()
        assert lib.qcd_gaussIteration3d(ctypes.byref(vv), ctypes.byref(uu), 2.5, 99) == 0
        assert np.array_equal(di.view(vv, 12), before)
    finally:
        di.close()


def test_device_blas(gpu):
    Q = gpu
    a, b = Q.spinor_gaussian(L, 81), Q.spinor_gaussian(L, 82)
    lat = Q.Lattice(L)
    fa, fb = Q.Spinor(lat).upload(a), Q.Spinor(lat).upload(b)
    assert abs(fa.norm2() - O.norm(a) ** 2) <= 1e-13 * O.norm(a) ** 2
    d = Q.dot(fa, fb)
    assert abs(d - O.dot(a, b)) <= 1e-12 * abs(O.norm(a) * O.norm(b))
    Q.axpy(fb, 0.5 - 2j, fa)
    assert rel_l2(fb.download(), b + (0.5 - 2j) * a) <= 1e-15
    fc = Q.Spinor(lat)
    Q.lincomb(fc, 1.5, fa, -1j, fb, t0=3)   # slice only: other sites stay zero
    got = fc.download()
    want = np.zeros_like(a)
    sl = np.arange(a.shape[0]) % L[0] == 3
    want[sl] = 1.5 * a[sl] - 1j * (b + (0.5 - 2j) * a)[sl]
    assert np.abs(got - want).max() <= 1e-14


def test_dropin_blas(gpu):
    """tests/test_blas.c:53-104 identities through the host-struct API."""
    Q = gpu
    a, b = Q.spinor_gaussian(L, 91), Q.spinor_gaussian(L, 92)
    di = Q.DropIn(L)
    try:
        lib = Q.lib()
        va, vb, vc = di.vector(a), di.vector(b), di.vector()
        assert abs(lib.qcd_normVector(ctypes.byref(va)) - O.norm(a)) <= 1e-13 * O.norm(a)
        z = lib.qcd_mulAdjointVectorVector(ctypes.byref(va), ctypes.byref(vb))
        assert abs(complex(z.re, z.im) - O.dot(a, b)) <= 1e-12 * O.norm(a) * O.norm(b)
        lib.qcd_subVector(ctypes.byref(vc), ctypes.byref(va), ctypes.byref(vb))
        assert np.abs(di.view(vc, 12) - (a - b).reshape(-1)).max() == 0.0
        lib.qcd_addVector(ctypes.byref(vc), ctypes.byref(vc), ctypes.byref(vb))   # sum aliases summand1
        assert np.abs(di.view(vc, 12) - a.reshape(-1)).max() <= 1e-15
        lib.qcd_axpyVector(ctypes.byref(vc), Q.QcdComplex(0.25, -1.0), ctypes.byref(va), ctypes.byref(vb))
        assert np.abs(di.view(vc, 12) - (b + (0.25 - 1j) * a).reshape(-1)).max() <= 1e-14
        lib.qcd_scaleVector3d(ctypes.byref(va), 2.0, 5)
        sl = np.arange(a.shape[0]) % L[0] == 5
        want = a.copy()
        want[sl] *= 2.0
        assert np.array_equal(di.view(va, 12).reshape(-1, 4, 3), want)
        lib.qcd_mulVectorC(ctypes.byref(vb), Q.QcdComplex(0.0, 1.0))
        assert np.abs(di.view(vb, 12) - (1j * b).reshape(-1)).max() == 0.0
    finally:
        di.close()


@pytest.mark.parametrize("Ls", [(4, 6, 4, 7), (6, 4, 6, 5), (4, 4, 4, 9)])
def test_gauss_z_chunks_of_unequal_length(gpu, Ls):
    """The smearing kernel walks z in chunks of ceil(Z / 4) (at least 2)
    planes; z extents of 7, 5 and 9 leave a shorter last chunk (and Z = 5 a
    chunk of one plane).  One slice and all slices against the oracle
    (lib/src/qcd_smearing.c:25-199)."""
    Q = gpu
    u = Q.gauge_generate(Ls, 0.25, 801)
    v = Q.spinor_gaussian(Ls, 802)
    lat = Q.Lattice(Ls)
    g = Q.Gauge(lat).upload(u)
    geo = O.Geometry(Ls)
    for t in (1, -1):
        f = Q.Spinor(lat).upload(v)
        Q.gauss_smear(f, g, 2.5, t, 3)
        ref = v.copy()
        for _ in range(3):
            ref = O.gauss3d(geo, ref, u, 2.5, t) if t >= 0 else O.gauss3d_all(geo, ref, u, 2.5)
        assert rel_l2(f.download(), ref) <= 1e-13, (Ls, t)
