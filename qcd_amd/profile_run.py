"""Run a fixed number of kernels for rocprofv3 (kernel trace or counters).

usage: python -m qcd_amd.profile_run --lattice 48 24 24 24 --prec f64 --iters 20
Runs `iters` Wilson Dslash applications and `iters` plain FP64 copies
(qcd_hip_lincomb with a = 1, b = 0: 192 B read + 192 B written per site) on
the same lattice, so that the counters of the copy calibrate those of Dslash.
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
    ap.add_argument("--op", default="D", choices=["D", "Q"])
    ap.add_argument("--map", type=int, default=None,
                    help="force a traversal (0 XCD ranges, 1 y-slabs, 2 + 256*n z-march of n steps)")
    ap.add_argument("--occ", type=int, default=0, help="occupancy variant when --map is given")
    ap.add_argument("--sweep", action="store_true", help="time every candidate (QCD_HIP_TUNE_VERBOSE)")
    args = ap.parse_args(argv)
    L = tuple(args.lattice)
    Q.init(0)
    lat = Q.Lattice(L)
    u = Q.gauge_generate(L, 0.3, 4001)
    psi = Q.spinor_gaussian(L, 4002)
    g64 = Q.Gauge(lat).upload(u)
    p = Q.PREC[args.prec]
    g = g64
    if p != Q.F64:
        g = Q.Gauge(lat, p)
        Q.convert(g, g64)
    x = Q.Spinor(lat, p).upload(psi)
    y = Q.Spinor(lat, p)
    m = 1 / (2 * 0.132) - 4
    if args.sweep:
        os.environ["QCD_HIP_TUNE_VERBOSE"] = "1"
        print("tuned:", Q.dslash_tune(y, x, g), flush=True)
    if args.map is not None:
        Q.dslash_set_config(x, g, args.map, args.occ)
    ms = Q.dslash_time(y, x, g, m, 0.0, Q.OP_Q if args.op == "Q" else Q.OP_D, warmup=0, iters=args.iters)
    a = Q.Spinor(lat).upload(psi)
    b = Q.Spinor(lat)
    for _ in range(args.iters):
        Q.lincomb(b, 1.0, a)
    Q.synchronize()
    V = int(np.prod(L))
    print(f"lattice {L} V={V} prec={args.prec} dslash mean {ms*1e3:.2f} us; copy bytes/launch {V*384}",
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
