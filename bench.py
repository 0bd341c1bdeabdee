#!/usr/bin/env python3
"""Benchmark: Wilson Dslash HBM GB/s (headline), solves/s and latency legs.

One "step" is one application of the FP64 Wilson Dslash (hopping stencil +
mass term) to one spinor on a 24^3x48 lattice, inputs resident in HBM.
Algorithmic bytes per site: spinor in 192 + links 576 + spinor out 192 =
960 B (FP64).  Prints one JSON line (see the README of this repository).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
BYTES_PER_SITE = {"f64": 960, "f32": 480, "f16": 248}
METRIC = "Dslash HBM GB/s; solves/s mixed vs FP64 at rel. resid 1e-10; p50/p99 latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def dslash_leg(Q, L, prec, steps, warmup, seed=3001):
    """Mean Dslash time (HIP events on the library stream) and GB/s."""
    V = int(np.prod(L))
    lat = Q.Lattice(L)
    u = Q.gauge_generate(L, 0.3, seed)
    psi = Q.spinor_gaussian(L, seed + 1)
    g64 = Q.Gauge(lat).upload(u)
    p = Q.PREC[prec]
    if p == Q.F64:
        g = g64
    else:
        g = Q.Gauge(lat, p)
        Q.convert(g, g64)
    x = Q.Spinor(lat, p).upload(psi)
    y = Q.Spinor(lat, p)
    m = 1 / (2 * 0.132) - 4
    cfg = Q.dslash_tune(y, x, g)
    Q.synchronize()
    t0 = time.perf_counter()
    ms = Q.dslash_time(y, x, g, m, 0.0, Q.OP_D, warmup=warmup, iters=steps)
    wall = time.perf_counter() - t0
    gbps = V * BYTES_PER_SITE[prec] / (ms * 1e-3) / 1e9
    return {"prec": prec, "L": list(L), "ms": ms, "GBps": gbps, "wall_s": wall,
            "config": {"traversal": ["xcd-contiguous", "xcd-y-slab"][cfg[0]], "min_waves": cfg[1]}}


def cpu_baseline_leg(L, seconds=10.0):
    """The CPU oracle (plain-C restatement of qcd_wilson.c) timed on host cores."""
    import oracle as O
    threads = cpu_threads()
    O.set_threads(threads)
    u = O.gauge_generate(L, 0.3, 3001)
    psi = O.spinor_gaussian(L, 3002)
    geo = O.Geometry(L)
    m = 1 / (2 * 0.132) - 4
    O.dslash(geo, psi, u, m)   # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        O.dslash(geo, psi, u, m)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 200:
            break
    V = int(np.prod(L))
    return {"value": V * 960 / (el / n) / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{n} FP64 Dslash applies of the CPU oracle (OpenMP, {threads} threads) on "
                      f"{L[1]}^3x{L[0]} in {el:.1f} s, 960 B/site"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs one rank on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import qcd_amd as Q
    if not os.path.exists(Q.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()
    Q.init(0)
    dev = Q.device_name()
    log(f"device: {dev}")

    L = (48, 24, 24, 24)
    head = dslash_leg(Q, L, "f64", args.steps, args.warmup)
    log(f"f64 dslash 24^3x48: {head['ms']*1e3:.1f} us, {head['GBps']:.0f} GB/s")
    extra = {}
    for prec in ("f32", "f16"):
        r = dslash_leg(Q, L, prec, args.steps, args.warmup)
        extra[prec] = {"ms": r["ms"], "GBps": r["GBps"], "frac": r["GBps"] / HBM_PEAK_GBPS, "config": r["config"]}
        log(f"{prec} dslash 24^3x48: {r['ms']*1e3:.1f} us, {r['GBps']:.0f} GB/s")
    L5 = (96, 48, 48, 48)
    big = dslash_leg(Q, L5, "f64", max(10, args.steps // 10), max(2, args.warmup // 10))
    log(f"f64 dslash 48^3x96: {big['ms']*1e3:.1f} us, {big['GBps']:.0f} GB/s")
    extra["f64_48c96"] = {"ms": big["ms"], "GBps": big["GBps"], "frac": big["GBps"] / HBM_PEAK_GBPS,
                          "config": big["config"]}
    extra["f64"] = {"config": head["config"]}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_leg(L)
        log(f"cpu baseline: {cpu['value']:.2f} GB/s ({cpu['sample']})")

    V = int(np.prod(L))
    out = {
        "metric": METRIC,
        "value": round(head["GBps"], 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: U = reunitarize(exp(i 0.3 H)) seeded links, N(0,1) spinor",
        "config": {"workload": "FP64 Wilson Dslash (hopping + mass), 24^3x48, kappa 0.132, "
                               "theta=(pi,0,0,0), one RHS, inputs resident in HBM",
                   "lattice": "24^3x48", "sites": V, "bytes_per_site": 960},
        "roofline": {"bound": "hbm", "achieved": round(head["GBps"], 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(head["GBps"] / HBM_PEAK_GBPS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "dslash": extra,
        "device": dev,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
