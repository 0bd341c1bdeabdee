"""Even-odd (Schur complement) operator and the even-odd preconditioned solver.

The Schur complement D^ = M - H_oe M^-1 H_eo is pinned through the oracle's
full operator D (qcd_wilson.c:114-336): for any odd-site vector x_o, the
vector z = (-M^-1 H_eo x_o, x_o) satisfies D z = (0, D^ x_o), where H_eo x_o
is the even part of D (0, x_o)."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def minv(v, m, tm):
    """M^-1 v with M = m + 4 + i tm g5 (g5 = diag(1, 1, -1, -1)), v of shape [n, 4, 3]."""
    a = m + 4.0
    f = np.array([1 / (a + 1j * tm)] * 2 + [1 / (a - 1j * tm)] * 2)
    return v * f[None, :, None]


def schur_reference(Q, L, u, x_o, m, tm, theta):
    geo = O.Geometry(L, theta=theta)
    zero = np.zeros_like(x_o)
    y1 = O.dslash(geo, Q.eo_merge(zero, x_o, L), u, m, tm)
    h_eo, _ = Q.eo_split(y1, L)
    z = Q.eo_merge(-minv(h_eo, m, tm), x_o, L)
    y2 = O.dslash(geo, z, u, m, tm)
    ye, yo = Q.eo_split(y2, L)
    assert np.linalg.norm(ye) <= 1e-13 * np.linalg.norm(yo)
    return yo


@pytest.mark.parametrize("L,theta", [((16, 8, 8, 8), (np.pi, 0.0, 0.0, 0.0)),
                                     ((6, 4, 2, 8), (np.pi, 0.3, -0.2, 0.5))])
def test_schur_complement_matches_full_operator(gpu, L, theta):
    Q = gpu
    u = Q.gauge_generate(L, 0.3, 601)
    psi = Q.spinor_gaussian(L, 602)
    _, x_o = Q.eo_split(psi, L)
    m, tm = 1 / (2 * 0.13) - 4, 0.05
    ref = schur_reference(Q, L, u, x_o, m, tm, theta)
    lat = Q.Lattice(L, theta=theta)
    Lh = (L[0] // 2,) + tuple(L[1:])
    half = Q.Lattice(Lh)
    g64 = Q.Gauge(lat).upload(u)
    for prec, tol in (("f64", 1e-13), ("f32", 2e-6), ("f16", 4e-3)):
        p = Q.PREC[prec]
        g = g64
        if p != Q.F64:
            g = Q.Gauge(lat, p)
            Q.convert(g, g64)
        xi = Q.Spinor(half, p).upload(x_o)
        yo = Q.Spinor(half, p)
        Q.dslash_schur(yo, xi, g, m, tm)
        assert rel_l2(yo.download(), ref) <= tol, prec


def test_schur_rejects_mismatched_fields(gpu):
    Q = gpu
    L = (8, 4, 4, 4)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(Q.gauge_generate(L, 0.2, 1))
    full = Q.Spinor(lat)
    half = Q.Lattice((4, 4, 4, 4))
    with pytest.raises(Q.QcdError):
        Q.dslash_schur(Q.Spinor(half), full, g, 0.1)
    odd = Q.Lattice((7, 4, 4, 4))   # not an all-even lattice
    with pytest.raises(Q.QcdError):
        Q.Solver(odd, Q.Gauge(odd), eo=1)


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
def test_eo_solver_point_sources(gpu, prec):
    """Even-odd preconditioned solves reach ||b - D x|| / ||b|| <= 1e-10 (oracle),
    agree with the oracle's solution, and give the same bits alone and batched."""
    Q = gpu
    L = (16, 8, 8, 8)
    kappa, mu = 0.13, 0.02
    u = Q.gauge_generate(L, 0.3, 501)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(u)
    geo = O.Geometry(L)
    fb = Q.Spinor(lat, nrhs=4)
    srcs = []
    for r, (sp, c) in enumerate([(0, 0), (1, 2), (2, 1), (3, 0)]):
        s = Q.point_source(L, (3, 1, 2, 5), sp, c)
        srcs.append(s)
        fb.upload(s, rhs=r)
    sols = {}
    for batch in (1, 4):
        fx = Q.Spinor(lat, nrhs=4)
        S = Q.Solver(lat, g, n_restart=16, tol=1e-10, inner_prec=prec, eo=1, batch=batch, max_outer=500)
        st = S.solve(fx, fb, kappa, mu)
        assert st["relres"] <= 1e-10
        sols[batch] = [fx.download(r) for r in range(4)]
        S.close()
    for r in range(4):
        assert np.array_equal(sols[1][r], sols[4][r])
        res = O.b_minus_Dx(geo, sols[4][r], srcs[r], u, kappa, mu)
        assert res <= 1e-10
    rc, xref, _, _ = O.invert(geo, srcs[1], u, kappa, mu, n=32, op=0, keep_q1=False, max_res=1e-12)
    assert rc == 0 and rel_l2(sols[4][1], xref) <= 1e-7


def test_schur_batched_matches_single(gpu):
    """FP32/FP16 batches of the even-odd hops (two rhs per lane, an odd batch's
    last lane with one) give each rhs the bits of a one-rhs application."""
    Q = gpu
    L = (8, 8, 4, 4)
    u = Q.gauge_generate(L, 0.3, 611)
    lat = Q.Lattice(L)
    half = Q.Lattice((L[0] // 2,) + tuple(L[1:]))
    g64 = Q.Gauge(lat).upload(u)
    m, tm = 1 / (2 * 0.13) - 4, 0.05
    for prec in ("f32", "f16"):
        p = Q.PREC[prec]
        g = Q.Gauge(lat, p)
        Q.convert(g, g64)
        vs = [Q.spinor_gaussian((L[0] // 2,) + tuple(L[1:]), 620 + r) for r in range(5)]
        xb = Q.Spinor(half, p, nrhs=5)
        for r, v in enumerate(vs):
            xb.upload(v, rhs=r)
        yb = Q.Spinor(half, p, nrhs=5)
        Q.dslash_schur(yb, xb, g, m, tm)
        for r, v in enumerate(vs):
            x1 = Q.Spinor(half, p).upload(v)
            y1 = Q.Spinor(half, p)
            Q.dslash_schur(y1, x1, g, m, tm)
            assert np.array_equal(yb.download(r), y1.download()), (prec, r)
