"""Oracle parity at the sizes of every benchmarked configuration.

The small-lattice tests pin the arithmetic; these run the tuned kernels at
their production tile shapes (T = 48 and T = 96 row tiles, the 12-step
z-march with the LDS-carried centre, two-rhs lanes over 16 and 48 rhs) and
re-check every solve through the CPU oracle's own b_minus_Dx residual
(b_minus_Dx.c:152-172), so a bug shared by the GPU operator and the GPU
residual cannot pass.

Workloads (kappa, eps, sources as the benchmark configurations):
  16^3x32, kappa 0.130, eps 0.3, 12 point sources at the origin, FP64 and mixed;
  24^3x48, kappa 0.132, eps 0.3, 48 rhs = 12 point sources at 4 seeded sites,
           FP16-storage even-odd batched solve;
  48^3x96 Dslash, and a T = 96 lattice through 10 APE steps (alpha 0.5), 50 Gauss
           iterations (alpha 4.0) of 12 point sources and a 12-rhs mixed solve;
  a serving mix with 16^3x32 and 24^3x48 requests.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

L24 = (48, 24, 24, 24)
L16 = (32, 16, 16, 16)
L48 = (96, 48, 48, 48)
KAPPA24 = 0.132


def _m(kappa):
    return 1.0 / (2.0 * kappa) - 4.0


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def f24(gpu):
    """24^3x48 links (eps 0.3) and one Gaussian spinor, host and device."""
    Q = gpu
    lat = Q.Lattice(L24)
    g = Q.Gauge(lat)
    Q.gauge_generate_device(g, 0.3, 3001)
    u = g.download()
    psi = Q.spinor_gaussian(L24, 3002)
    return Q, lat, g, u, psi


def _lowprec(Q, lat, g64, prec):
    p = Q.PREC[prec]
    if p == Q.F64:
        return g64
    g = Q.Gauge(lat, p)
    Q.convert(g, g64)
    return g


# ------------------------------------------------------------ single-rhs Dslash
@pytest.mark.parametrize("cfg", [None, (4 | (12 << 8), 16), (3 | (8 << 8), 16), (2 | (6 << 8), 0)],
                         ids=["tuned", "zmarch-lds12-fence", "rowtile8-fence", "zmarch6"])
def test_fp64_24c48_d_and_q(f24, cfg):
    """FP64 D and Q at 24^3x48 (kappa 0.132, theta_t = pi) on the tuned kernel
    and on forced headline candidates (T = 48 row tiles: 4 rows x 48 = 192 lanes)."""
    Q, lat, g, u, psi = f24
    _say(f"24^3x48 FP64 config {cfg}")
    m, tm = _m(KAPPA24), 0.0
    x = Q.Spinor(lat).upload(psi)
    y = Q.Spinor(lat)
    if cfg is not None:
        Q.dslash_set_config(x, g, *cfg)
    geo = O.Geometry(L24)
    Q.dslash(y, x, g, m, tm, Q.OP_D)
    assert rel_l2(y.download(), O.dslash(geo, psi, u, m, tm)) <= 1e-13
    Q.dslash(y, x, g, m, 0.03, Q.OP_Q)
    assert rel_l2(y.download(), O.qtm(geo, psi, u, m, 0.03)) <= 1e-13
    if cfg is not None:
        Q.dslash_tune(y, x, g)   # leave the tuned configuration behind


@pytest.mark.parametrize("prec,tol", [("f32", 2e-6), ("f16", 4e-3)])
def test_lowprec_24c48_tuned(f24, prec, tol):
    """FP32 and FP16-storage D at 24^3x48 with the autotuned traversal."""
    Q, lat, g64, u, psi = f24
    _say(f"24^3x48 {prec} tuned")
    m = _m(KAPPA24)
    g = _lowprec(Q, lat, g64, prec)
    x = Q.Spinor(lat, Q.PREC[prec]).upload(psi)
    y = Q.Spinor(lat, Q.PREC[prec])
    Q.dslash_tune(y, x, g)
    Q.dslash(y, x, g, m, 0.0, Q.OP_D)
    assert rel_l2(y.download(), O.dslash(O.Geometry(L24), psi, u, m, 0.0)) <= tol


@pytest.mark.parametrize("cfg", [(3 | (6 << 8), 16), (4 | (12 << 8), 16)], ids=["rowtile6-fence", "zmarch-lds12"])
def test_fp64_t96_row_tiles(gpu, cfg):
    """T = 96: row tiles of 2 x 96 = 192 lanes with 75 KB of dynamic LDS (the
    48^3x96 kernel's shape) on a 96 x 8 x 6 x 10 lattice, z chunks ragged."""
    Q = gpu
    L = (96, 8, 6, 10)
    _say(f"96x8x6x10 FP64 config {cfg}")
    m, tm = _m(0.132), 0.01
    lat = Q.Lattice(L)
    g = Q.Gauge(lat)
    Q.gauge_generate_device(g, 0.3, 3101)
    u = g.download()
    psi = Q.spinor_gaussian(L, 3102)
    x = Q.Spinor(lat).upload(psi)
    y = Q.Spinor(lat)
    Q.dslash_set_config(x, g, *cfg)
    Q.dslash(y, x, g, m, tm, Q.OP_D)
    assert rel_l2(y.download(), O.dslash(O.Geometry(L), psi, u, m, tm)) <= 1e-13


def test_fp64_48c96_tuned(gpu):
    """The full 48^3x96 FP64 Dslash (10.6 M sites) with its autotuned kernel."""
    Q = gpu
    _say("48^3x96 FP64 tuned")
    m = _m(0.132)
    lat = Q.Lattice(L48)
    g = Q.Gauge(lat)
    Q.gauge_generate_device(g, 0.3, 5005)
    u = g.download()
    psi = Q.spinor_gaussian(L48, 5006)
    x = Q.Spinor(lat).upload(psi)
    y = Q.Spinor(lat)
    mp, occ, _ = Q.dslash_tune(y, x, g)
    _say(f"48^3x96 tuned map {mp} occ {occ}")
    Q.dslash(y, x, g, m, 0.0, Q.OP_D)
    out = y.download()
    del x, y
    assert rel_l2(out, O.dslash(O.Geometry(L48), psi, u, m, 0.0)) <= 1e-13


# ------------------------------------------------------------ batched Dslash, per rhs
@pytest.mark.parametrize("nrhs", [16, 48])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 2e-6), ("f16", 4e-3)])
def test_batched_dslash_per_rhs_16c32(gpu, nrhs, prec, tol):
    """Every rhs of a 16- and 48-rhs launch (the tuned batch traversal: two
    rhs per lane for FP32/FP16, rows grouped by 4) against the oracle."""
    Q = gpu
    L = L16
    _say(f"16^3x32 batched {prec} x {nrhs}")
    m = _m(0.130)
    lat = Q.Lattice(L)
    g64 = Q.Gauge(lat)
    Q.gauge_generate_device(g64, 0.3, 2002)
    u = g64.download()
    g = _lowprec(Q, lat, g64, prec)
    p = Q.PREC[prec]
    x = Q.Spinor(lat, p, nrhs=nrhs)
    y = Q.Spinor(lat, p, nrhs=nrhs)
    srcs = [Q.spinor_gaussian(L, 7000 + r) for r in range(nrhs)]
    for r, s in enumerate(srcs):
        x.upload(s, rhs=r)
    Q.dslash_tune(y, x, g)
    Q.dslash(y, x, g, m, 0.02, Q.OP_Q)
    geo = O.Geometry(L)
    for r in range(nrhs):
        assert rel_l2(y.download(r), O.qtm(geo, srcs[r], u, m, 0.02)) <= tol, r


def test_batched_f16_16rhs_24c48(f24):
    """16 FP16 rhs at 24^3x48 with the tuned batch traversal, each vs the oracle."""
    Q, lat, g64, u, psi = f24
    _say("24^3x48 batched f16 x 16")
    m = _m(KAPPA24)
    g = _lowprec(Q, lat, g64, "f16")
    x = Q.Spinor(lat, Q.F16, nrhs=16)
    y = Q.Spinor(lat, Q.F16, nrhs=16)
    srcs = [psi if r == 0 else Q.spinor_gaussian(L24, 3300 + r) for r in range(16)]
    for r, s in enumerate(srcs):
        x.upload(s, rhs=r)
    Q.dslash_tune(y, x, g)
    Q.dslash(y, x, g, m, 0.0, Q.OP_D)
    geo = O.Geometry(L24)
    for r in range(16):
        assert rel_l2(y.download(r), O.dslash(geo, srcs[r], u, m, 0.0)) <= 4e-3, r


# ------------------------------------------------------------ solves re-checked by the oracle
def _point_sources(Q, lat, L, sites):
    b = Q.Spinor(lat, nrhs=12 * len(sites))
    host = []
    for i, site in enumerate(sites):
        for r in range(12):
            Q.point_source_device(b, 12 * i + r, site, r // 3, r % 3)
            host.append(O.point_source(L, site, r // 3, r % 3))
    return b, host


def test_16c32_solves_fp64_and_mixed(gpu):
    """16^3x32, kappa 0.130: 12 origin point sources solved by the plain FP64
    solver one at a time (invert.c's structure, GCR(32)), by FP32-inner one at
    a time and by FP16-inner even-odd in one batch.  Every solution's residual
    is recomputed by the oracle; mixed solutions agree with the FP64 ones."""
    Q = gpu
    L, kappa = L16, 0.130
    lat = Q.Lattice(L)
    g = Q.Gauge(lat)
    Q.gauge_generate_device(g, 0.3, 2002)
    u = g.download()
    geo = O.Geometry(L)
    b, host = _point_sources(Q, lat, L, [(0, 0, 0, 0)])
    sols = {}
    for name, kw in (("f64", dict(inner_prec="f64", batch=1, n_restart=32)),
                     ("f32", dict(inner_prec="f32", batch=1, n_restart=32)),
                     ("f16eo", dict(inner_prec="f16", batch=12, n_restart=12, eo=1))):
        _say(f"16^3x32 solve {name}")
        x = Q.Spinor(lat, nrhs=12)
        S = Q.Solver(lat, g, tol=1e-10, max_outer=20000, **kw)
        st = S.solve(x, b, kappa)
        assert st["relres"] <= 1e-10
        sols[name] = [x.download(r) for r in range(12)]
        for r in range(12):
            assert O.b_minus_Dx(geo, sols[name][r], host[r], u, kappa) <= 1e-10, (name, r)
    for name in ("f32", "f16eo"):
        worst = max(rel_l2(sols[name][r], sols["f64"][r]) for r in range(12))
        _say(f"16^3x32 max ||x_{name} - x_f64|| / ||x_f64|| = {worst:.2e}")
        assert worst <= 1e-7


def test_24c48_48rhs_f16_eo_batched_solve(f24):
    """24^3x48, kappa 0.132: 48 rhs (12 point sources at each of 4 seeded
    sites) in one FP16-storage even-odd batch; every rhs re-checked by the
    oracle's b_minus_Dx."""
    Q, lat, g, u, _ = f24
    rng = np.random.default_rng(3003)
    sites = [tuple(int(rng.integers(0, L24[d])) for d in range(4)) for _ in range(4)]
    b, host = _point_sources(Q, lat, L24, sites)
    x = Q.Spinor(lat, nrhs=48)
    S = Q.Solver(lat, g, tol=1e-10, inner_prec="f16", batch=48, n_restart=12, eo=1, max_outer=20000)
    _say("24^3x48 48-rhs f16 even-odd solve")
    st = S.solve(x, b, KAPPA24)
    assert st["relres"] <= 1e-10
    geo = O.Geometry(L24)
    worst = 0.0
    for r in range(48):
        res = O.b_minus_Dx(geo, x.download(r), host[r], u, KAPPA24)
        worst = max(worst, res)
        assert res <= 1e-10, r
    _say(f"24^3x48 worst oracle residual {worst:.2e}")


def test_t96_ape_gauss_and_mixed_solve(gpu):
    """The 48^3x96 pipeline reduced to 96 x 8 x 8 x 8: 10 APE steps (alpha 0.5)
    against the oracle, 50 Gauss iterations (alpha 4.0) of 12 origin point
    sources with the APE links against the oracle, then the 12-rhs FP16-inner
    even-odd solve with the original links, re-checked by the oracle."""
    Q = gpu
    L, kappa = (96, 8, 8, 8), 0.132
    _say("96x8x8x8 pipeline")
    lat = Q.Lattice(L)
    g = Q.Gauge(lat)
    Q.gauge_generate_device(g, 0.3, 5005)
    u = g.download()
    geo = O.Geometry(L)
    ga = Q.Gauge(lat)
    Q.ape_smear(ga, g, 0.5, 10)
    ua = u
    for _ in range(10):
        ua = O.ape3d(geo, ua, 0.5)
    assert rel_l2(ga.download(), ua) <= 1e-13
    b, host = _point_sources(Q, lat, L, [(0, 0, 0, 0)])
    Q.gauss_smear(b, ga, 4.0, 0, 50)
    smeared = []
    for r in range(12):
        v = host[r]
        for _ in range(50):
            v = O.gauss3d(geo, v, ua, 4.0, 0)
        smeared.append(v)
        assert rel_l2(b.download(r), v) <= 1e-13, r
    x = Q.Spinor(lat, nrhs=12)
    S = Q.Solver(lat, g, tol=1e-10, inner_prec="f16", batch=12, n_restart=12, eo=1, max_outer=20000)
    st = S.solve(x, b, kappa)
    assert st["relres"] <= 1e-10
    for r in range(12):
        assert O.b_minus_Dx(geo, x.download(r), b.download(r), u, kappa) <= 1e-10, r


# ------------------------------------------------------------ serving at the larger volumes
def test_serve_larger_volumes_mix_equals_alone(gpu):
    """A serving mix with 16^3x32 and 24^3x48 requests beside an 8^3x16 one:
    every served solution passes the oracle's residual and equals the same
    request served alone bit for bit."""
    Q = gpu
    reqs = [
        {"arrival_s": 0.0, "L": (48, 24, 24, 24), "kappa": 0.131, "eps": 0.3, "seed": 9101, "nrhs": 12,
         "inner_prec": "f16", "n_restart": 12, "tol": 1e-10},
        {"arrival_s": 0.001, "L": (32, 16, 16, 16), "kappa": 0.126, "eps": 0.25, "seed": 9102, "nrhs": 1,
         "inner_prec": "f32", "n_restart": 12, "tol": 1e-10},
        {"arrival_s": 0.002, "L": (16, 8, 8, 8), "kappa": 0.121, "eps": 0.22, "seed": 9103, "nrhs": 12,
         "inner_prec": "f64", "n_restart": 12, "tol": 1e-10},
        {"arrival_s": 0.003, "L": (32, 16, 16, 16), "kappa": 0.129, "eps": 0.33, "seed": 9104, "nrhs": 12,
         "inner_prec": "f16", "n_restart": 12, "tol": 1e-10},
    ]
    _say("serving mix 24^3x48 + 16^3x32 + 8^3x16")
    mix, _, sols = Q.serve(reqs, workers=3, realtime=True, want_solutions=True)
    assert all(r["rc"] == 0 and r["relres"] <= 1e-10 for r in mix)
    for k, q in enumerate(reqs):
        alone, _, s1 = Q.serve([dict(q, arrival_s=0.0)], workers=1, realtime=False, want_solutions=True)
        assert alone[0]["iters"] == mix[k]["iters"]
        assert np.array_equal(s1[0], sols[k])
        L = q["L"]
        lat = Q.Lattice(L)
        g = Q.Gauge(lat)
        Q.gauge_generate_device(g, q["eps"], q["seed"])
        u = g.download()
        geo = O.Geometry(L)
        for r in range(q["nrhs"]):
            b = O.point_source(L, (0, 0, 0, 0), r // 3, r % 3)
            assert O.b_minus_Dx(geo, sols[k][r], b, u, q["kappa"]) <= 1e-10, (k, r)


def test_solver_bytes_is_what_the_solver_allocates(gpu):
    """qcd_hip_solver_bytes (the serving admission figure) equals the device
    bytes a solver allocates, for an even-odd FP16 solver with a short cycle
    (where the even-odd temporaries dominate) and for a plain FP64 one."""
    Q = gpu
    L = (16, 8, 8, 8)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat)
    Q.gauge_generate_device(g, 0.3, 1)
    for kw in (dict(inner_prec="f16", n_restart=4, batch=12, eo=1), dict(inner_prec="f64", n_restart=32, batch=1),
               dict(inner_prec="f32", n_restart=7, batch=5, eo=1)):
        Q.synchronize()
        before = Q.bytes_in_use()
        S = Q.Solver(lat, g, **kw)
        used = Q.bytes_in_use() - before
        assert used == Q.solver_bytes(L, **kw), kw
        S.close()
