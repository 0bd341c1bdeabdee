"""The z-marching even-odd hop (k_eo_hop_zm, qcd_amd/csrc/eo.hip) against the
one-site kernels (k_eo_hop, k_eo_hop2) and the oracle.

The z-march computes every hop from the same stored values with the same
arithmetic and adds them in the same order, so its Schur operator must give
the one-site kernels' bits for every z chunk length, including chunks that do
not divide the z extent, chunks of the whole extent (both z wraps inside one
lane) and two rows per lane with an odd batch.  The one-site kernels are
pinned against the oracle's full operator (test_gpu_eo.py); one case here is
checked against the oracle directly as well."""
import numpy as np
import pytest

from conftest import rel_l2
from test_gpu_eo import schur_reference

pytestmark = pytest.mark.gpu


def _schur(Q, L, theta, prec, nrhs, seed, zlens):
    lat = Q.Lattice(L, theta=theta)
    half = Q.Lattice((L[0] // 2,) + tuple(L[1:]))
    u = Q.gauge_generate(L, 0.3, seed)
    g64 = Q.Gauge(lat).upload(u)
    p = Q.PREC[prec]
    g = g64
    if p != Q.F64:
        g = Q.Gauge(lat, p)
        Q.convert(g, g64)
    xs = [Q.eo_split(Q.spinor_gaussian(L, seed + 1 + r), L)[1] for r in range(nrhs)]
    x = Q.Spinor(half, p, nrhs=nrhs)
    for r, v in enumerate(xs):
        x.upload(v, rhs=r)
    y = Q.Spinor(half, p, nrhs=nrhs)
    m, tm = 1 / (2 * 0.13) - 4, 0.04
    out = {}
    try:
        for z in zlens:
            Q.eo_set_zlen(z)
            Q.dslash_schur(y, x, g, m, tm)
            out[z] = [y.download(r) for r in range(nrhs)]
    finally:
        Q.eo_set_zlen(0)
    return u, xs, m, tm, out


@pytest.mark.parametrize("L,theta", [((8, 4, 4, 8), (np.pi, 0.3, -0.2, 0.5)),
                                     ((4, 6, 2, 6), (np.pi, 0.0, 0.0, 0.0))])
@pytest.mark.parametrize("prec,nrhs", [("f64", 1), ("f64", 3), ("f32", 5), ("f16", 1), ("f16", 12)])
def test_zmarch_schur_bits_equal_one_site(gpu, L, theta, prec, nrhs):
    """z chunks of 2, 3 (ragged), 5 (ragged) and the whole extent against one
    site per lane, bit for bit, for every rhs."""
    Q = gpu
    zl = [1, 2, 3, 5, L[3]]
    _, _, _, _, out = _schur(Q, L, theta, prec, nrhs, 900 + nrhs, zl)
    for z in zl[1:]:
        for r in range(nrhs):
            assert np.array_equal(out[z][r], out[1][r]), (prec, nrhs, z, r)


def test_zmarch_schur_matches_oracle(gpu):
    """FP64, z chunks of 3 over an extent of 8, against the oracle's full operator."""
    Q = gpu
    L, theta = (8, 4, 4, 8), (np.pi, 0.3, -0.2, 0.5)
    u, xs, m, tm, out = _schur(Q, L, theta, "f64", 2, 950, [3])
    for r in range(2):
        ref = schur_reference(Q, L, u, xs[r], m, tm, theta)
        assert rel_l2(out[3][r], ref) <= 1e-13


@pytest.mark.parametrize("prec", ["f64", "f16"])
def test_zmarch_eo_solve_converges_like_one_site(gpu, prec):
    """An even-odd solve with z-marched hops (including the outer step's
    subtracting hops and the one with the |r|^2 reduction) converges in the
    same number of outer cycles to the same tolerance."""
    Q = gpu
    L = (8, 4, 4, 8)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat)
    Q.gauge_generate_device(g, 0.3, 960)
    b = Q.Spinor(lat, nrhs=3)
    for r in range(3):
        Q.point_source_device(b, r, (1, 2, 0, 3), r, r)
    stats = {}
    try:
        for z in (1, 3):
            Q.eo_set_zlen(z)
            x = Q.Spinor(lat, nrhs=3)
            S = Q.Solver(lat, g, n_restart=8, tol=1e-10, inner_prec=prec, batch=3, eo=1)
            stats[z] = S.solve(x, b, 0.13)
            S.close()
            worst = max(Q.b_minus_Dx(x, b, g, 0.13, rx=r, rb=r) for r in range(3))
            assert worst <= 1e-10, (z, worst)
    finally:
        Q.eo_set_zlen(0)
    assert abs(stats[3]["outer"] - stats[1]["outer"]) <= 1


@pytest.mark.parametrize("prec,nrhs", [("f64", 1), ("f64", 13), ("f16", 12)])
def test_zmarch_schur_bits_equal_one_site_16c32(gpu, prec, nrhs):
    """16^3x32: 16 column blocks per plane, several z chunks per row and, with
    13 / 12 rhs, row groups of more than one; z chunks of 3, 4 and 8 and the
    tuned choice (0) against one site per lane, bit for bit."""
    Q = gpu
    L, theta = (32, 16, 16, 16), (np.pi, 0.0, 0.0, 0.0)
    zl = [1, 3, 4, 8, 0]
    _, _, _, _, out = _schur(Q, L, theta, prec, nrhs, 970 + nrhs, zl)
    for z in zl[1:]:
        for r in range(nrhs):
            assert np.array_equal(out[z][r], out[1][r]), (prec, nrhs, z, r)
