"""3d APE smearing and the spatial SU(3) projection on the GPU
(lib/src/qcd_smearing.c:205-273, lib/src/qcd_blas.c:940-1364) against the
CPU oracle, whose projection and staples are pinned in tests/test_oracle.py."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

L = (16, 8, 8, 8)


def test_ape_steps_match_oracle(gpu):
    Q = gpu
    u = Q.gauge_generate(L, 0.3, 701)
    geo = O.Geometry(L)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(u)
    out = Q.Gauge(lat)
    ref1 = O.ape3d(geo, u, 0.5)
    Q.ape_smear(out, g, 0.5, 1)
    assert np.abs(out.download() - ref1).max() <= 1e-13
    ref3 = O.ape3d(geo, O.ape3d(geo, ref1, 0.5), 0.5)
    Q.ape_smear(out, g, 0.5, 3)
    assert np.abs(out.download() - ref3).max() <= 1e-13
    # the input field is left unchanged, and zero steps copy it
    assert np.array_equal(g.download(), u)
    Q.ape_smear(out, g, 0.5, 0)
    assert np.array_equal(out.download(), u)


def test_projection_of_perturbed_links(gpu):
    """alpha = 0: spatial links projected, temporal links copied bit for bit."""
    Q = gpu
    rng = np.random.default_rng(7)
    u = Q.gauge_generate(L, 0.3, 702)
    m = u + 0.2 * (rng.normal(size=u.shape) + 1j * rng.normal(size=u.shape))
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(m)
    out = Q.Gauge(lat)
    Q.ape_smear(out, g, 0.0, 1)
    got = out.download()
    assert np.array_equal(got[:, 0], m[:, 0])
    ref = O.project_su3(m[:, 1:])
    assert np.abs(got[:, 1:] - ref).max() <= 1e-13


def test_dropin_ape_and_projection(gpu):
    Q = gpu
    Ls = (8, 4, 4, 4)
    u = Q.gauge_generate(Ls, 0.4, 703)
    di = Q.DropIn(Ls)
    try:
        lib = Q.lib()
        uu, aa = di.gauge(u), di.gauge()
        assert lib.qcd_apeSmear3d(ctypes.byref(aa), ctypes.byref(uu), 0.5) == 0
        ref = O.ape3d(O.Geometry(Ls), u, 0.5)
        assert np.abs(di.view(aa, 36).reshape(u.shape) - ref).max() <= 1e-13
        # in place (apeu aliasing u) works here
        assert lib.qcd_apeSmear3d(ctypes.byref(uu), ctypes.byref(uu), 0.5) == 0
        assert np.abs(di.view(uu, 36).reshape(u.shape) - ref).max() <= 1e-13
        rng = np.random.default_rng(9)
        m = u + 0.1 * rng.normal(size=u.shape)
        pp = di.gauge(m)
        lib.qcd_projectSU33d(ctypes.byref(pp))
        got = di.view(pp, 36).reshape(u.shape)
        assert np.array_equal(got[:, 0], m[:, 0])
        assert np.abs(got[:, 1:] - O.project_su3(m[:, 1:])).max() <= 1e-13
    finally:
        di.close()


def test_ape_rejects_bad_arguments(gpu):
    Q = gpu
    lat = Q.Lattice((8, 4, 4, 4))
    g = Q.Gauge(lat)
    with pytest.raises(Q.QcdError):
        Q.ape_smear(g, g, 0.5, 1)
    with pytest.raises(Q.QcdError):
        Q.ape_smear(Q.Gauge(lat, Q.F32), g, 0.5, 1)
    with pytest.raises(Q.QcdError):
        Q.ape_smear(Q.Gauge(lat), g, 0.5, -1)


def test_plaquette_device_and_dropin(gpu):
    """qcd_calculatePlaquette (lib/src/qcd_observables.c:15-36) on the GPU against the oracle; 1 for U = 1."""
    Q = gpu
    u = Q.gauge_generate(L, 0.3, 704)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(u)
    got = ctypes.c_double()
    assert Q.lib().qcd_hip_plaquette(g.h, ctypes.byref(got)) == 0
    ref = O.plaquette(O.Geometry(L), u)
    assert abs(got.value - ref) <= 1e-13
    one = np.zeros_like(u)
    one[:, :, [0, 1, 2], [0, 1, 2]] = 1.0
    di = Q.DropIn(L)
    try:
        lib = Q.lib()
        lib.qcd_calculatePlaquette.restype = ctypes.c_double
        lib.qcd_calculatePlaquette.argtypes = [ctypes.POINTER(Q.QcdField)]
        assert abs(lib.qcd_calculatePlaquette(ctypes.byref(di.gauge(u))) - ref) <= 1e-13
        assert abs(lib.qcd_calculatePlaquette(ctypes.byref(di.gauge(one))) - 1.0) <= 1e-15
    finally:
        di.close()


@pytest.mark.parametrize("ndim", [3, 4])
def test_stout_steps_match_oracle(gpu, ndim):
    """Stout smearing (lib/src/qcd_smearing.c:374-476) on the GPU against the
    oracle (pinned by brute force with scipy's expm and gauge covariance)."""
    Q = gpu
    u = Q.gauge_generate(L, 0.3, 711)
    geo = O.Geometry(L)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(u)
    out = Q.Gauge(lat)
    ref1 = O.stout(geo, u, 0.2, ndim)
    Q.stout_smear(out, g, 0.2, ndim, 1)
    assert np.abs(out.download() - ref1).max() <= 1e-13
    ref2 = O.stout(geo, ref1, 0.2, ndim)
    Q.stout_smear(out, g, 0.2, ndim, 2)
    assert np.abs(out.download() - ref2).max() <= 1e-13
    assert np.array_equal(g.download(), u)
    # U = 1 stays exactly 1 (the reference's exponential gives NaN there)
    one = np.zeros_like(u)
    one[:, :, [0, 1, 2], [0, 1, 2]] = 1.0
    Q.stout_smear(out, Q.Gauge(lat).upload(one), 0.2, ndim, 1)
    assert np.array_equal(out.download(), one)


def test_dropin_stout_and_bad_arguments(gpu):
    Q = gpu
    Ls = (8, 4, 4, 4)
    u = Q.gauge_generate(Ls, 0.4, 712)
    di = Q.DropIn(Ls)
    try:
        lib = Q.lib()
        uu, ss = di.gauge(u), di.gauge()
        assert lib.qcd_stoutsmearing(ctypes.byref(ss), ctypes.byref(uu), 0.1, 4) == 0
        ref = O.stout(O.Geometry(Ls), u, 0.1, 4)
        assert np.abs(di.view(ss, 36).reshape(u.shape) - ref).max() <= 1e-13
        assert lib.qcd_stoutsmearing(ctypes.byref(uu), ctypes.byref(uu), 0.1, 4) == 0   # in place
        assert np.abs(di.view(uu, 36).reshape(u.shape) - ref).max() <= 1e-13
        assert lib.qcd_stoutsmearing(ctypes.byref(ss), ctypes.byref(uu), 0.1, 2) == 1
    finally:
        di.close()
    lat = Q.Lattice(Ls)
    g = Q.Gauge(lat)
    with pytest.raises(Q.QcdError):
        Q.stout_smear(g, g, 0.1, 3, 1)
    with pytest.raises(Q.QcdError):
        Q.stout_smear(Q.Gauge(lat), g, 0.1, 5, 1)


def _drand48_stream(seed, n):
    """The libc drand48 sequence after srand48(seed): X <- (a X + c) mod 2^48."""
    a, c, mask = 0x5DEECE66D, 0xB, (1 << 48) - 1
    x = ((seed << 16) | 0x330E) & mask
    out = np.empty(n)
    for i in range(n):
        x = (a * x + c) & mask
        out[i] = x / float(1 << 48)
    return out


def test_project_su3_all_links_and_random_gauge_field(gpu):
    """qcd_projectSU3 projects every link (all four directions) on the device,
    link by link equal to the oracle's projection; qcd_randomGaugeField draws
    drand48() entries in the reference's order (x, mu, c1, c2, re then im)
    and projects them (lib/src/qcd_blas.c:1771-1782)."""
    import ctypes
    Q = gpu
    L = (4, 2, 2, 2)
    V = 32
    rng = np.random.default_rng(91)
    m = (rng.normal(size=(V, 4, 3, 3)) + 1j * rng.normal(size=(V, 4, 3, 3)))
    di = Q.DropIn(L)
    lib = Q.lib()
    try:
        u = di.gauge(m)
        lib.qcd_projectSU3.argtypes = [ctypes.POINTER(Q.QcdField)]
        lib.qcd_projectSU3(ctypes.byref(u))
        got = di.view(u, 36).reshape(V, 4, 3, 3).copy()
        assert np.abs(got - O.project_su3(m)).max() <= 1e-13
        libc = ctypes.CDLL(None)
        libc.srand48.argtypes = [ctypes.c_long]
        libc.srand48(12345)
        lib.qcd_randomGaugeField.argtypes = [ctypes.POINTER(Q.QcdField)]
        lib.qcd_randomGaugeField(ctypes.byref(u))
        d = _drand48_stream(12345, V * 72)
        raw = (d[0::2] + 1j * d[1::2]).reshape(V, 4, 3, 3)
        got = di.view(u, 36).reshape(V, 4, 3, 3).copy()
        assert np.abs(got - O.project_su3(raw)).max() <= 1e-13
    finally:
        di.close()
