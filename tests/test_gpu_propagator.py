"""Propagator columns on the device (lib/src/qcd_blas.c:182-233): the 12
columns (nu, c2) of a qcd_propagator gathered into 12 rhs of a device field and
scattered back, and the operator applied to all 12 columns in one launch."""
import ctypes

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _random_prop(V, seed):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=(V, 4, 4, 3, 3)) + 1j * rng.normal(size=(V, 4, 4, 3, 3))).astype(np.complex128)


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
def test_propagator_round_trip_and_columns(gpu, prec):
    """Upload then download gives the propagator back (bit for bit in FP64);
    rhs 3 nu + c2 of the field is column (nu, c2), i.e. P[:, :, nu, :, c2]."""
    Q = gpu
    L = (8, 4, 6, 4)
    V = int(np.prod(L))
    P = _random_prop(V, 5)
    lat = Q.Lattice(L)
    f = Q.Spinor(lat, Q.PREC[prec], nrhs=14)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert Q.lib().qcd_hip_propagator_upload(f.h, 1, dp(P)) == 0
    tol = {"f64": 0.0, "f32": 1e-7, "f16": 1e-3}[prec]
    for nu in range(4):
        for c2 in range(3):
            col = np.ascontiguousarray(P[:, :, nu, :, c2])
            got = f.download(1 + 3 * nu + c2)
            if tol == 0.0:
                assert np.array_equal(got, col)
            else:
                assert rel_l2(got, col) <= tol
    back = np.zeros_like(P)
    assert Q.lib().qcd_hip_propagator_download(f.h, 1, dp(back)) == 0
    if tol == 0.0:
        assert np.array_equal(back, P)
    else:
        assert rel_l2(back, P) <= tol
    # out-of-range columns are rejected
    assert Q.lib().qcd_hip_propagator_upload(f.h, 3, dp(P)) != 0


def test_propagator_operator_matches_oracle_per_column(gpu):
    """qcd_hip_applyWilsonTMOpPropagator / QTMOp: each of the 12 columns
    equals the oracle's D (Q) of that column, as the reference's
    column-by-column loop would give."""
    Q = gpu
    L = (16, 8, 8, 8)
    V = int(np.prod(L))
    m, tm = 1 / (2 * 0.13) - 4, 0.02
    u = Q.gauge_generate(L, 0.3, 601)
    P = _random_prop(V, 602)
    geo = O.Geometry(L)
    di = Q.DropIn(L)
    lib = Q.lib()
    try:
        uu = di.gauge(u)
        p = di.propagator()
        dpp = di.propagator()
        di.view(p, 144)[:] = P.reshape(-1)
        for fn, ref in ((lib.qcd_hip_applyWilsonTMOpPropagator, O.dslash),
                        (lib.qcd_hip_applyQTMOpPropagator, O.qtm)):
            assert fn(ctypes.byref(dpp), ctypes.byref(p), ctypes.byref(uu), m, tm) == 0
            out = di.view(dpp, 144).reshape(V, 4, 4, 3, 3)
            for nu in range(4):
                for c2 in range(3):
                    col = np.ascontiguousarray(P[:, :, nu, :, c2])
                    assert rel_l2(out[:, :, nu, :, c2], ref(geo, col, u, m, tm)) <= 1e-13, (nu, c2)
        assert lib.qcd_hip_applyWilsonTMOpPropagator(ctypes.byref(p), ctypes.byref(p), ctypes.byref(uu), m, tm) == 1
    finally:
        di.close()
