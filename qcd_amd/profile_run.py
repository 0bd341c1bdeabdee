"""Run a fixed number of kernels for rocprofv3 (kernel trace or counters).

usage: python -m qcd_amd.profile_run --lattice 48 24 24 24 --prec f64 --iters 20
Runs `iters` Wilson Dslash applications and `iters` plain FP64 copies
(qcd_hip_lincomb with a = 1, b = 0: 192 B read + 192 B written per site) on
the same lattice, so that the counters of the copy calibrate those of Dslash.

       python -m qcd_amd.profile_run --lattice 48 24 24 24 --prec f16 --rhs 48 [--schur]
Runs `iters` batched launches over 48 right-hand sides instead (with --schur:
the even-odd Schur operator M - H_oe M^-1 H_eo on odd-site half vectors, two
hop launches per application).

       python -m qcd_amd.profile_run --solve --lattice 48 24 24 24 --prec f16 --rhs 48 --eo 1
Runs one propagator solve instead: 12 point sources (spin x colour) at each
of rhs/12 sites, kappa 0.132, eps 0.3, GCR(n_restart), tol 1e-10, all rhs in one
batch, so that a kernel trace shows where a solve spends its time.
"""
import argparse
import os
import sys

import numpy as np

import qcd_amd as Q


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lattice", type=int, nargs=4, default=[48, 24, 24, 24], help="T X Y Z")
    ap.add_argument("--prec", default="f64", choices=["f64", "f32", "f16"])
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=0,
                    help="untimed launches first (0 for counter runs, so every dispatch is a timed one)")
    ap.add_argument("--op", default="D", choices=["D", "Q"])
    ap.add_argument("--map", type=int, default=None,
                    help="force a traversal (0 XCD ranges, 1 y-slabs, 2 + 256*n z-march of n steps)")
    ap.add_argument("--occ", type=int, default=0, help="occupancy variant when --map is given")
    ap.add_argument("--sweep", action="store_true", help="time every candidate (QCD_HIP_TUNE_VERBOSE)")
    ap.add_argument("--solve", action="store_true", help="profile one batched propagator solve instead")
    ap.add_argument("--rhs", type=int, default=None,
                    help="right-hand sides: --solve a multiple of 12 (default 12); otherwise one batched launch")
    ap.add_argument("--schur", action="store_true", help="profile the even-odd Schur operator instead of D")
    ap.add_argument("--eo", type=int, default=1, help="--solve: even-odd preconditioning")
    ap.add_argument("--n-restart", type=int, default=32, help="--solve: GCR cycle length (invert.ini: 32)")
    ap.add_argument("--warm-solve", action="store_true",
                    help="--solve: run a short solve first, so the autotuner stays out of the profiled one")
    args = ap.parse_args(argv)
    L = tuple(args.lattice)
    Q.init(0)
    if args.solve:
        return solve(Q, L, args.prec, args.rhs or 12, args.eo, args.n_restart, args.warm_solve)
    if args.schur:
        return schur(Q, L, args.prec, args.rhs or 1, args.iters)
    nrhs = args.rhs or 1
    lat = Q.Lattice(L)
    u = Q.gauge_generate(L, 0.3, 4001)
    psi = Q.spinor_gaussian(L, 4002)
    g64 = Q.Gauge(lat).upload(u)
    p = Q.PREC[args.prec]
    g = g64
    if p != Q.F64:
        g = Q.Gauge(lat, p)
        Q.convert(g, g64)
    x = Q.Spinor(lat, p, nrhs=nrhs)
    for r in range(nrhs):
        x.upload(psi if r == 0 else Q.spinor_gaussian(L, 4002 + r), rhs=r)
    y = Q.Spinor(lat, p, nrhs=nrhs)
    m = 1 / (2 * 0.132) - 4
    if args.sweep:
        os.environ["QCD_HIP_TUNE_VERBOSE"] = "1"
        print("tuned:", Q.dslash_tune(y, x, g), flush=True)
    if args.map is not None:
        Q.dslash_set_config(x, g, args.map, args.occ)
    ms = Q.dslash_time(y, x, g, m, 0.0, Q.OP_Q if args.op == "Q" else Q.OP_D, warmup=args.warmup, iters=args.iters)
    a = Q.Spinor(lat).upload(psi)
    b = Q.Spinor(lat)
    for _ in range(args.iters):
        Q.lincomb(b, 1.0, a)
    Q.synchronize()
    V = int(np.prod(L))
    print(f"lattice {L} V={V} prec={args.prec} rhs={nrhs} dslash mean {ms*1e3:.2f} us "
          f"({ms*1e3/nrhs:.2f} us per rhs); copy bytes/launch {V*384}", flush=True)
    return 0


def schur(Q, L, prec, nrhs, iters):
    """iters applications of the even-odd Schur operator to nrhs odd-site half vectors."""
    import time
    T, X, Y, Z = L
    lat = Q.Lattice(L)
    half = Q.Lattice((T // 2, X, Y, Z))
    g64 = Q.Gauge(lat)
    Q.gauge_generate_device(g64, 0.3, 4001)
    p = Q.PREC[prec]
    g = g64
    if p != Q.F64:
        g = Q.Gauge(lat, p)
        Q.convert(g, g64)
    Vh = half.V
    x = Q.Spinor(half, p, nrhs=nrhs)
    for r in range(nrhs):
        x.upload(Q.spinor_gaussian((Vh, 1, 1, 1), 4100 + r), rhs=r)
    y = Q.Spinor(half, p, nrhs=nrhs)
    m = 1 / (2 * 0.132) - 4
    Q.dslash_schur(y, x, g, m)   # tunes and warms up
    Q.synchronize()
    t = time.perf_counter()
    for _ in range(iters):
        Q.dslash_schur(y, x, g, m)
    Q.synchronize()
    el = (time.perf_counter() - t) / iters
    a = Q.Spinor(lat).upload(Q.spinor_gaussian(L, 4002))   # the plain FP64 copy that calibrates the counters
    b = Q.Spinor(lat)
    for _ in range(iters):
        Q.lincomb(b, 1.0, a)
    Q.synchronize()
    print(f"lattice {L} prec={prec} rhs={nrhs} Schur operator {el*1e3:.3f} ms per application "
          f"({el*1e6/nrhs:.2f} us per rhs)", flush=True)
    return 0


def solve(Q, L, prec, nrhs, eo, n_restart=32, warm=False):
    lat = Q.Lattice(L)
    g = Q.Gauge(lat)
    Q.gauge_generate_device(g, 0.3, 3003)
    rng = np.random.default_rng(3003)
    sites = [tuple(int(rng.integers(0, L[d])) for d in range(4)) for _ in range(max(1, nrhs // 12))]
    b = Q.Spinor(lat, nrhs=12 * len(sites))
    for i, site in enumerate(sites):
        for r in range(12):
            Q.point_source_device(b, 12 * i + r, site, r // 3, r % 3)
    x = Q.Spinor(lat, nrhs=12 * len(sites))
    if warm:
        # a short solve first, so the autotuner's timing launches stay out of the profiled solve
        Q.Solver(lat, g, n_restart=n_restart, tol=1e-2, inner_prec=prec, batch=12 * len(sites), eo=eo).solve(x, b, 0.132)
        Q.synchronize()
    S = Q.Solver(lat, g, n_restart=n_restart, tol=1e-10, inner_prec=prec, batch=12 * len(sites), eo=eo,
                 max_outer=20000)
    st = S.solve(x, b, 0.132)
    Q.synchronize()
    print(f"lattice {L} prec={prec} rhs={12 * len(sites)} eo={eo}: {st['seconds']:.3f} s, "
          f"{st['inner_iters']} GCR iterations, {st['outer']} outer cycles, relres {st['relres']:.2e}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
