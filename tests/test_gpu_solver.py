"""GPU propagator solves (invert.c:20-294) against the CPU oracle.

Every solve is accepted only through the FP64 residual ||b - D x|| / ||b||
(b_minus_Dx.c:152-167) computed by the CPU oracle; solutions are compared
with the oracle's own GCR solution (relative difference bounded by the
conditioning times the tolerance)."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

L = (16, 8, 8, 8)


@pytest.fixture(scope="module")
def setup(gpu):
    Q = gpu
    u = Q.gauge_generate(L, 0.3, 501)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(u)
    b = Q.point_source(L, (0, 0, 0, 0), 1, 2)
    rc, xref, outer, rel = O.invert(O.Geometry(L), b, u, 0.13, 0.0, n=32, op=0, keep_q1=False, max_res=1e-12)
    assert rc == 0
    return Q, u, lat, g, b, xref


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
@pytest.mark.parametrize("op", ["Q", "D"])
def test_solve_point_source(setup, prec, op):
    Q, u, lat, g, b, xref = setup
    kappa = 0.13
    S = Q.Solver(lat, g, n_restart=16, tol=1e-10, inner_prec=prec, op=op, max_outer=400)
    fb = Q.Spinor(lat).upload(b)
    fx = Q.Spinor(lat)
    st = S.solve(fx, fb, kappa)
    x = fx.download()
    res = O.b_minus_Dx(O.Geometry(L), x, b, u, kappa)
    assert res <= 1e-10 and st["relres"] <= 1e-10
    assert abs(Q.b_minus_Dx(fx, fb, g, kappa) - res) <= 1e-3 * res + 1e-15
    assert rel_l2(x, xref) <= 1e-7


def test_solve_multi_rhs_and_twisted(setup):
    """12 spin-colour sources in one call, twisted mass on (invert.c:255-294 loop)."""
    Q, u, lat, g, b, _ = setup
    kappa, mu = 0.13, 0.02
    fb = Q.Spinor(lat, nrhs=12)
    srcs = []
    for sp in range(4):
        for c in range(3):
            s = Q.point_source(L, (1, 2, 3, 4), sp, c)
            srcs.append(s)
            fb.upload(s, rhs=3 * sp + c)
    geo = O.Geometry(L)
    sols = {}
    for batch in (1, 5, 12):   # one at a time (invert.c), ragged batches, all at once
        fx = Q.Spinor(lat, nrhs=12)
        S = Q.Solver(lat, g, n_restart=24, tol=1e-10, inner_prec="f32", batch=batch)
        st = S.solve(fx, fb, kappa, mu)
        assert st["outer"] >= 12
        sols[batch] = [fx.download(r) for r in range(12)]
        for r in range(12):
            assert O.b_minus_Dx(geo, sols[batch][r], srcs[r], u, kappa, mu) <= 1e-10
    # a right-hand side gets the same answer alone and inside a batch
    for r in range(12):
        assert np.array_equal(sols[1][r], sols[12][r])
        assert np.array_equal(sols[5][r], sols[12][r])


def test_reference_q1_variant(setup):
    """keep_q1 reproduces invert.c's dropped first direction; it still converges."""
    Q, u, lat, g, b, xref = setup
    S = Q.Solver(lat, g, n_restart=32, tol=1e-10, inner_prec="f64", keep_q1=1)
    fb = Q.Spinor(lat).upload(b)
    fx = Q.Spinor(lat)
    S.solve(fx, fb, 0.13)
    assert O.b_minus_Dx(O.Geometry(L), fx.download(), b, u, 0.13) <= 1e-10


def test_nonconvergence_reported(setup):
    Q, u, lat, g, b, _ = setup
    S = Q.Solver(lat, g, n_restart=4, tol=1e-14, inner_prec="f16", max_outer=2)
    fb = Q.Spinor(lat).upload(b)
    fx = Q.Spinor(lat)
    st = S.solve(fx, fb, 0.13, check=False)
    assert st["rc"] == 6 and st["outer"] == 2


def test_f16_store_of_denormal_site_maxima(gpu):
    """Sites whose largest entry is a float denormal are stored as zero, not NaN
    (1/max would overflow; this broke FP16-inner GCR on point sources)."""
    Q = gpu
    lat = Q.Lattice(L)
    v = Q.spinor_gaussian(L, 77)
    v[5] *= 1e-41
    v[6] *= 1e-36
    v[7] = 0.0
    f = Q.Spinor(lat, Q.F16).upload(v)
    got = f.download()
    assert np.isfinite(got).all()
    assert np.all(got[5] == 0.0) and np.all(got[7] == 0.0)
    assert np.abs(got[6] - v[6]).max() <= 1e-3 * np.abs(v[6]).max()
    g = Q.Gauge(lat).upload(Q.gauge_generate(L, 0.3, 78))
    g16 = Q.Gauge(lat, Q.F16)
    Q.convert(g16, g)
    y = Q.Spinor(lat, Q.F16)
    Q.dslash(y, f, g16, 1 / (2 * 0.13) - 4, 0.0, Q.OP_Q)
    assert np.isfinite(y.download()).all()


def test_scratch_fields_outlive_destroyed_lattices(gpu):
    """b_minus_Dx borrows a library scratch spinor.  After a smaller lattice (and its
    fields) is destroyed, a larger lattice must get a scratch buffer of its own size."""
    Q = gpu
    import gc
    for Ls in ((8, 4, 4, 4), (16, 8, 8, 8), (8, 4, 4, 4), (16, 8, 8, 12)):
        u = Q.gauge_generate(Ls, 0.3, 901)
        x, b = Q.spinor_gaussian(Ls, 902), Q.spinor_gaussian(Ls, 903)
        lat = Q.Lattice(Ls)
        g = Q.Gauge(lat).upload(u)
        fx, fb = Q.Spinor(lat).upload(x), Q.Spinor(lat).upload(b)
        res = Q.b_minus_Dx(fx, fb, g, 0.13)
        ref = O.b_minus_Dx(O.Geometry(Ls), x, b, u, 0.13)
        assert abs(res - ref) <= 1e-12 * ref
        del g, fx, fb, lat
        gc.collect()


@pytest.mark.parametrize("prec,eo", [("f64", 0), ("f32", 1), ("f16", 1)])
def test_graph_replayed_cycles_match_eager(setup, prec, eo):
    """GCR cycles replayed from a captured HIP graph give the same bits as the
    kernel-by-kernel launches: over many replays per solve (with the outer
    FP64 residual, host syncs and slot-list uploads between them), after the
    active rhs count shrinks (re-capture), on a second solve that reuses the
    graph, at a new kappa, and after the gauge field is set again."""
    Q, u, lat, g, b, _ = setup
    fb = Q.Spinor(lat, nrhs=4)
    for r, (site, sp, c) in enumerate([((0, 0, 0, 0), 1, 2), ((3, 1, 2, 5), 0, 0),
                                       ((7, 7, 1, 0), 2, 1), ((15, 4, 6, 3), 3, 2)]):
        fb.upload(Q.point_source(L, site, sp, c), rhs=r)
    runs = {}
    for graph in (0, 1):
        S = Q.Solver(lat, g, n_restart=8, tol=1e-10, inner_prec=prec, eo=eo, batch=4, graph=graph,
                     max_outer=2000)
        out = []
        for kappa, mu in ((0.13, 0.0), (0.13, 0.0), (0.128, 0.03)):
            fx = Q.Spinor(lat, nrhs=4)
            st = S.solve(fx, fb, kappa, mu)
            assert st["relres"] <= 1e-10 and st["outer"] >= 8
            out.append([fx.download(r) for r in range(4)])
        S.set_gauge(g)
        fx = Q.Spinor(lat, nrhs=4)
        S.solve(fx, fb, 0.13)
        out.append([fx.download(r) for r in range(4)])
        S.close()
        runs[graph] = out
    for a, e in zip(runs[1], runs[0]):
        for r in range(4):
            assert np.array_equal(a[r], e[r])
    for r in range(4):   # the same solve twice, and after set_gauge, is the same
        assert np.array_equal(runs[1][0][r], runs[1][1][r])
        assert np.array_equal(runs[1][0][r], runs[1][3][r])
