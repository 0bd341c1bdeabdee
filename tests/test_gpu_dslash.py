"""Parity of the gfx950 Wilson Dslash with the CPU oracle (lib/src/qcd_wilson.c:114-336).

FP64 must match to <= 1e-13 relative L2 on the same seeded fields; FP32 and
FP16-storage operators are checked against their storage precision.  At the
full 24^3x48 size, size-independent properties (linearity, gamma5-hermiticity)
are checked on the GPU alone.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

L_8C16 = (16, 8, 8, 8)   # 8^3 x 16, tests/test_wilson.c pattern


def _fields(L, eps, su, sp):
    import qcd_amd as Q
    return Q.gauge_generate(L, eps, su), Q.spinor_gaussian(L, sp)


@pytest.mark.parametrize("op", ["D", "Q"])
@pytest.mark.parametrize("tm", [0.0, 0.05])
def test_fp64_8c16_parity(gpu, op, tm):
    Q = gpu
    L = L_8C16
    kappa = 0.125
    m = 1 / (2 * kappa) - 4
    u, psi = _fields(L, 0.2, 11, 12)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(u)
    x = Q.Spinor(lat).upload(psi)
    y = Q.Spinor(lat)
    Q.dslash(y, x, g, m, tm, Q.OP_Q if op == "Q" else Q.OP_D)
    geo = O.Geometry(L)
    ref = O.qtm(geo, psi, u, m, tm) if op == "Q" else O.dslash(geo, psi, u, m, tm)
    assert rel_l2(y.download(), ref) <= 1e-13


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (6, 4, 2, 8), (2, 2, 2, 2), (8, 6, 4, 10)])
def test_fp64_ragged_lattices_and_phases(gpu, L):
    """Odd-looking extents (2 = both neighbours equal), phases in every direction."""
    Q = gpu
    theta = (np.pi, 0.3, -1.1, 2.0)
    m = 1 / (2 * 0.13) - 4
    u, psi = _fields(L, 0.35, 21, 22)
    lat = Q.Lattice(L, theta)
    g = Q.Gauge(lat).upload(u)
    x = Q.Spinor(lat).upload(psi)
    y = Q.Spinor(lat)
    Q.dslash(y, x, g, m, 0.02, Q.OP_D)
    ref = O.dslash(O.Geometry(L, theta), psi, u, m, 0.02)
    assert rel_l2(y.download(), ref) <= 1e-13


def test_dropin_apply_matches_oracle(gpu):
    """qcd_applyWilsonTMOp / qcd_applyQTMOp through host qcd_vector structs."""
    import ctypes
    Q = gpu
    L = L_8C16
    kappa, tm = 0.125, 0.01
    m = 1 / (2 * kappa) - 4
    u, psi = _fields(L, 0.2, 31, 32)
    di = Q.DropIn(L)
    try:
        uu = di.gauge(u)
        v = di.vector(psi)
        dv = di.vector()
        qv = di.vector()
        lib = Q.lib()
        for fn in (lib.qcd_applyWilsonTMOp, lib.qcd_applyWilsonTMOpNoScripts, lib.qcd_applyWilsonTMOpNoOptim):
            assert fn(ctypes.byref(dv), ctypes.byref(v), ctypes.byref(uu), m, tm) == 0
            got = di.view(dv, 12).reshape(-1, 4, 3).copy()
            assert rel_l2(got, O.dslash(O.Geometry(L), psi, u, m, tm)) <= 1e-13
        assert lib.qcd_applyQTMOp(ctypes.byref(qv), ctypes.byref(v), ctypes.byref(uu), m, tm) == 0
        assert rel_l2(di.view(qv, 12).reshape(-1, 4, 3).copy(), O.qtm(O.Geometry(L), psi, u, m, tm)) <= 1e-13
        # a bound (resident) gauge field gives the same answer
        assert lib.qcd_hip_bindGaugeField(ctypes.byref(uu)) == 0
        assert lib.qcd_applyWilsonTMOp(ctypes.byref(dv), ctypes.byref(v), ctypes.byref(uu), m, tm) == 0
        assert rel_l2(di.view(dv, 12).reshape(-1, 4, 3).copy(), O.dslash(O.Geometry(L), psi, u, m, tm)) <= 1e-13
        lib.qcd_hip_unbindGaugeField()
        # reference error contract: uninitialised input -> return 1
        bad = Q.QcdField()
        assert lib.qcd_applyWilsonTMOp(ctypes.byref(dv), ctypes.byref(bad), ctypes.byref(uu), m, tm) == 1
    finally:
        di.close()


def test_c_caller_program(gpu, tmp_path):
    """A plain-C caller (gcc, include/qcd.h) gets the oracle's answer."""
    import os
    import subprocess
    from conftest import ROOT
    L = L_8C16
    out = tmp_path / "dpsi.bin"
    exe = os.path.join(ROOT, "build", "bin", "wilson_dropin")
    r = subprocess.run([exe, *map(str, L), "0.125", "0.0", "0.2", "41", "42", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    import qcd_amd as Q
    u, psi = Q.gauge_generate(L, 0.2, 41), Q.spinor_gaussian(L, 42)
    got = np.fromfile(out, dtype=np.complex128).reshape(-1, 4, 3)
    assert rel_l2(got, O.dslash(O.Geometry(L), psi, u, 1 / (2 * 0.125) - 4, 0.0)) <= 1e-13
    gap = float(r.stdout.strip().splitlines()[-1].split("=")[-1])
    assert gap <= 1e-15


@pytest.mark.parametrize("prec,tol", [("f32", 2e-6), ("f16", 4e-3)])
def test_low_precision_operator(gpu, prec, tol):
    """FP32 / FP16-storage (per-site norm) Dslash vs the FP64 oracle."""
    Q = gpu
    L = L_8C16
    m = 1 / (2 * 0.13) - 4
    u, psi = _fields(L, 0.3, 51, 52)
    lat = Q.Lattice(L)
    p = Q.PREC[prec]
    g64 = Q.Gauge(lat).upload(u)
    g = Q.Gauge(lat, p)
    Q.convert(g, g64)
    x = Q.Spinor(lat, p).upload(psi)
    y = Q.Spinor(lat, p)
    Q.dslash(y, x, g, m, 0.0, Q.OP_Q)
    assert rel_l2(y.download(), O.qtm(O.Geometry(L), psi, u, m, 0.0)) <= tol


def test_full_size_properties(gpu):
    """24^3x48: linearity and gamma5-hermiticity <phi, g5 D psi> = <g5 D phi, psi>."""
    Q = gpu
    L = (48, 24, 24, 24)
    m = 1 / (2 * 0.132) - 4
    u = Q.gauge_generate(L, 0.3, 61)
    a, b = Q.spinor_gaussian(L, 62), Q.spinor_gaussian(L, 63)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(u)
    xa, xb = Q.Spinor(lat).upload(a), Q.Spinor(lat).upload(b)
    ya, yb = Q.Spinor(lat), Q.Spinor(lat)
    Q.dslash(ya, xa, g, m, 0.0, Q.OP_D)
    Q.dslash(yb, xb, g, m, 0.0, Q.OP_D)
    g5 = np.array([1, 1, -1, -1])[None, :, None]
    Da, Db = ya.download(), yb.download()
    lhs = np.vdot(a, g5 * Db)
    rhs = np.vdot(g5 * Da, b)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # linearity: D(2a - 3i b) = 2 Da - 3i Db
    xs = Q.Spinor(lat).upload(2 * a - 3j * b)
    ys = Q.Spinor(lat)
    Q.dslash(ys, xs, g, m, 0.0, Q.OP_D)
    assert rel_l2(ys.download(), 2 * Da - 3j * Db) <= 1e-14


def test_launch_configs_bit_identical(gpu):
    """Autotuner candidates: FP64 gives the same bits for every traversal and
    occupancy; FP32/FP16 agree to rounding (the compiler's FMA contraction
    differs between code shapes)."""
    Q = gpu
    L = (16, 8, 8, 8)
    m = 1 / (2 * 0.13) - 4
    u, psi = _fields(L, 0.3, 71, 72)
    lat = Q.Lattice(L)
    for prec, tol in (("f64", 0.0), ("f32", 1e-6), ("f16", 1e-3)):
        p = Q.PREC[prec]
        g64 = Q.Gauge(lat).upload(u)
        g = g64
        if p != Q.F64:
            g = Q.Gauge(lat, p)
            Q.convert(g, g64)
        x = Q.Spinor(lat, p).upload(psi)
        y = Q.Spinor(lat, p)
        by_kernel = {}
        for map_ in (0, 1, 2 | (1 << 8), 2 | (3 << 8), 2 | (5 << 8), 3 | (1 << 8), 3 | (2 << 8), 3 | (5 << 8),
                     4 | (1 << 8), 4 | (3 << 8), 4 | (8 << 8)):
            for occ in (0, 2, 3, 16):
                Q.dslash_set_config(x, g, map_, occ)
                Q.dslash(y, x, g, m, 0.01, Q.OP_Q)
                by_kernel.setdefault(map_ & 0xff >= 2, []).append(y.download())
        ref = by_kernel[True][0]
        for outs in by_kernel.values():
            for o in outs:
                if tol == 0.0:
                    assert np.array_equal(o, ref)
                else:
                    assert rel_l2(o, ref) <= tol
        mp, occ, ms = Q.dslash_tune(y, x, g)
        assert (mp & 0xff) in (0, 1, 2, 3, 4) and occ in (0, 2, 3, 16) and ms > 0


def test_two_rhs_lanes_match_one_rhs_lanes(gpu):
    """Batched FP32/FP16 launches with two right-hand sides per lane (one link
    load for both) give the same bits as one rhs per lane, for an odd batch
    (the last lane carries a single rhs) and for Q and D; FP64 batches and
    single-rhs launches fall back to the one-rhs z-march."""
    Q = gpu
    L = (16, 8, 8, 8)
    m = 1 / (2 * 0.13) - 4
    u, _ = _fields(L, 0.3, 81, 82)
    lat = Q.Lattice(L)
    g64 = Q.Gauge(lat).upload(u)
    for prec in ("f64", "f32", "f16"):
        p = Q.PREC[prec]
        g = g64
        if p != Q.F64:
            g = Q.Gauge(lat, p)
            Q.convert(g, g64)
        x = Q.Spinor(lat, p, nrhs=5)
        for r in range(5):
            x.upload(Q.spinor_gaussian(L, 90 + r), rhs=r)
        y = Q.Spinor(lat, p, nrhs=5)
        for op in (Q.OP_Q, Q.OP_D):
            outs = []
            for map_ in (2 | (3 << 8), 5 | (3 << 8)):
                Q.dslash_set_config(x, g, map_, 16)
                Q.dslash(y, x, g, m, 0.02, op)
                outs.append([y.download(r) for r in range(5)])
            for a, b in zip(*outs):
                assert np.array_equal(a, b), prec


def test_fp64_batch_traversals_match_one_rhs(gpu):
    """FP64 batches: the traversals the tuner times for batches (z-march of 4,
    6, 8 and 12 steps, LDS carry of 8, row tiles of 4 and 8; with and without
    the scheduling fence) and the tuned batch choice give every rhs of a
    5-rhs batch the bits of a one-rhs launch, on a lattice with ragged z
    chunks and row tiles; rhs 0 equals the oracle."""
    Q = gpu
    L = (16, 6, 4, 20)
    m = 1 / (2 * 0.13) - 4
    u, _ = _fields(L, 0.3, 181, 182)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(u)
    psis = [Q.spinor_gaussian(L, 190 + r) for r in range(5)]
    x = Q.Spinor(lat, nrhs=5)
    for r, v in enumerate(psis):
        x.upload(v, rhs=r)
    y = Q.Spinor(lat, nrhs=5)
    x1, y1 = Q.Spinor(lat), Q.Spinor(lat)
    for op in (Q.OP_Q, Q.OP_D):
        one = []
        for v in psis:
            x1.upload(v)
            Q.dslash(y1, x1, g, m, 0.02, op)
            one.append(y1.download())
        ref = (O.qtm if op == Q.OP_Q else O.dslash)(O.Geometry(L), psis[0], u, m, 0.02)
        assert rel_l2(one[0], ref) <= 1e-13
        for map_ in (2 | (4 << 8), 2 | (6 << 8), 2 | (8 << 8), 2 | (12 << 8), 4 | (8 << 8), 3 | (4 << 8), 3 | (8 << 8)):
            for occ in (0, 16):
                Q.dslash_set_config(x, g, map_, occ)
                Q.dslash(y, x, g, m, 0.02, op)
                for r in range(5):
                    assert np.array_equal(y.download(r), one[r]), (op, map_ & 0xff, map_ >> 8, occ, r)
        Q.dslash_tune(y, x, g)
        Q.dslash(y, x, g, m, 0.02, op)
        for r in range(5):
            assert np.array_equal(y.download(r), one[r]), (op, "tuned", r)


@pytest.mark.parametrize("L", [(16, 8, 8, 8), (16, 6, 4, 20)])
def test_every_tuner_candidate(gpu, L):
    """Every traversal the autotuner can pick (qcd_hip_dslash_candidates) with
    every occupancy variant, on a lattice shorter in z than the longest
    z-marches and on one whose z extent (20) leaves ragged chunks of 12 and 16
    and whose 6 x-rows split into row tiles of 4 and 2: FP64 bit-identical
    across candidates and equal to the oracle; FP32 and FP16 within storage
    precision of the oracle."""
    Q = gpu
    m = 1 / (2 * 0.13) - 4
    u, psi = _fields(L, 0.3, 171, 172)
    lat = Q.Lattice(L)
    g64 = Q.Gauge(lat).upload(u)
    cands = Q.dslash_candidates()
    assert len(cands) >= 20 and all((c & 0xff) in (2, 3, 4) for c in cands)
    ref = O.qtm(O.Geometry(L), psi, u, m, 0.01)
    for prec, tol in (("f64", 1e-13), ("f32", 2e-6), ("f16", 4e-3)):
        p = Q.PREC[prec]
        g = g64
        if p != Q.F64:
            g = Q.Gauge(lat, p)
            Q.convert(g, g64)
        x = Q.Spinor(lat, p).upload(psi)
        y = Q.Spinor(lat, p)
        first = None
        for map_ in cands:
            for occ in (0, 2, 3, 16):
                Q.dslash_set_config(x, g, map_, occ)
                Q.dslash(y, x, g, m, 0.01, Q.OP_Q)
                out = y.download()
                assert rel_l2(out, ref) <= tol, (prec, map_ & 0xff, map_ >> 8, occ)
                if prec == "f64":
                    if first is None:
                        first = out
                    assert np.array_equal(out, first), (map_ & 0xff, map_ >> 8, occ)


def test_bound_gauge_field_follows_in_place_writes(gpu):
    """A bound (resident) gauge field that a drop-in call rewrites in place
    (APE, stout, SU(3) projection, copy, zero) is refreshed on the device: the
    next operator call sees the new links."""
    import ctypes
    Q = gpu
    L = (8, 4, 4, 6)
    m, tm = 1 / (2 * 0.13) - 4, 0.0
    u, psi = _fields(L, 0.3, 181, 182)
    u2 = Q.gauge_generate(L, 0.25, 183)
    geo = O.Geometry(L)
    di = Q.DropIn(L)
    lib = Q.lib()
    try:
        uu = di.gauge(u)
        other = di.gauge(u2)
        v = di.vector(psi)
        dv = di.vector()

        def apply_and_check(expect_u):
            assert lib.qcd_applyWilsonTMOp(ctypes.byref(dv), ctypes.byref(v), ctypes.byref(uu), m, tm) == 0
            got = di.view(dv, 12).reshape(-1, 4, 3).copy()
            assert rel_l2(got, O.dslash(geo, psi, expect_u, m, tm)) <= 1e-13

        assert lib.qcd_hip_bindGaugeField(ctypes.byref(uu)) == 0
        apply_and_check(u)
        assert lib.qcd_apeSmear3d(ctypes.byref(uu), ctypes.byref(uu), 0.5) == 0
        ua = O.ape3d(geo, u, 0.5)
        apply_and_check(ua)
        assert lib.qcd_stoutsmearing(ctypes.byref(uu), ctypes.byref(uu), 0.1, 4) == 0
        us = O.stout(geo, ua, 0.1, 4)
        apply_and_check(us)
        lib.qcd_copyGaugeField(ctypes.byref(uu), ctypes.byref(other))
        apply_and_check(u2)
        di.view(uu, 36)[:] = (1.7 * u2).reshape(-1)   # links no longer in SU(3)
        assert lib.qcd_hip_bindGaugeField(ctypes.byref(uu)) == 0
        lib.qcd_projectSU33d(ctypes.byref(uu))
        up = di.view(uu, 36).reshape(-1, 4, 3, 3).copy()
        assert np.abs(up[:, 1:] - u2[:, 1:]).max() <= 1e-12   # spatial links projected back
        apply_and_check(up)
        lib.qcd_zeroGaugeField(ctypes.byref(uu))
        apply_and_check(np.zeros_like(u))
        lib.qcd_hip_unbindGaugeField()
    finally:
        di.close()


def test_dropin_phases_come_from_the_gauge_geometry(gpu):
    """qcd_applyWilsonTMOp takes the boundary phases from u->geo->theta
    (lib/src/qcd_wilson.c:174-177), also when the vector's geometry carries
    other phases."""
    import ctypes
    Q = gpu
    L = (8, 4, 4, 6)
    m, tm = 1 / (2 * 0.13) - 4, 0.0
    u, psi = _fields(L, 0.3, 191, 192)
    th_u, th_v = (np.pi, 0.4, 0.0, -0.7), (0.0, 0.0, 0.0, 0.0)
    du, dv_ = Q.DropIn(L, th_u), Q.DropIn(L, th_v)
    lib = Q.lib()
    try:
        uu = du.gauge(u)
        v = dv_.vector(psi)
        out = dv_.vector()
        assert lib.qcd_applyWilsonTMOp(ctypes.byref(out), ctypes.byref(v), ctypes.byref(uu), m, tm) == 0
        got = dv_.view(out, 12).reshape(-1, 4, 3).copy()
        assert rel_l2(got, O.dslash(O.Geometry(L, th_u), psi, u, m, tm)) <= 1e-13
    finally:
        dv_.close()
        du.close()
