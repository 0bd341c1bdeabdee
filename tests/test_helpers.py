"""The reference's small host helpers (include/qcd_wilson.h:627-632,
include/qcd_blas.h:224-273) on the CPU: per-site spinor operations against
the gamma table of lib/src/qcd_gamma.c:46-88, 3x3 matrix helpers, propagator
and gauge-field arithmetic against numpy, and the one-matrix SU(3)
projection against the SVD polar factor."""
import ctypes

import numpy as np
import pytest

import qcd_amd as Q
from test_oracle import G, G5

C16 = Q.QcdComplex
M3 = (C16 * 3) * 3
SP = (C16 * 3) * 4


def _m3(a):
    m = M3()
    for i in range(3):
        for j in range(3):
            m[i][j] = C16(a[i, j].real, a[i, j].imag)
    return m


def _np3(m):
    return np.array([[complex(m[i][j].re, m[i][j].im) for j in range(3)] for i in range(3)])


def _sp(a):
    v = SP()
    for s in range(4):
        for c in range(3):
            v[s][c] = C16(a[s, c].real, a[s, c].imag)
    return v


def _npsp(v):
    return np.array([[complex(v[s][c].re, v[s][c].im) for c in range(3)] for s in range(4)])


@pytest.fixture(scope="module")
def lib():
    L = Q.lib()
    for name in ("qcd_onePlusGamma", "qcd_oneMinusGamma"):
        getattr(L, name).argtypes = [ctypes.POINTER(SP), ctypes.POINTER(SP), ctypes.c_uint16]
    for name in ("qcd_applyU", "qcd_applyUDagger"):
        getattr(L, name).argtypes = [ctypes.POINTER(SP), ctypes.POINTER(M3), ctypes.POINTER(SP)]
    L.qcd_massTerms.argtypes = [ctypes.POINTER(SP), ctypes.POINTER(SP), ctypes.c_double, ctypes.c_double]
    L.qcd_phaseMul.argtypes = [ctypes.POINTER(SP), C16]
    for name in ("qcd_sub3x3", "qcd_add3x3", "qcd_addAdjoint3x3", "qcd_mul3x3", "qcd_mulAdjoint3x3"):
        getattr(L, name).argtypes = [ctypes.POINTER(M3)] * 3
    for name in ("qcd_zero3x3", "qcd_unit3x3", "qcd_minusUnit3x3", "qcd_dagger3x3", "qcd_projectSU33x3"):
        getattr(L, name).argtypes = [ctypes.POINTER(M3)]
    L.qcd_scale3x3.argtypes = [ctypes.POINTER(M3), ctypes.c_double]
    L.qcd_cScale3x3.argtypes = [ctypes.POINTER(M3), C16]
    L.qcd_trace3x3.argtypes = [ctypes.POINTER(M3)]
    L.qcd_trace3x3.restype = C16
    L.qcd_copy3x3.argtypes = [ctypes.POINTER(M3), ctypes.POINTER(M3)]
    pf = ctypes.POINTER(Q.QcdField)
    L.qcd_copyPropagatorPropagator.argtypes = [pf, pf]
    L.qcd_mulPropagatorC.argtypes = [pf, C16]
    L.qcd_mulPropagatorC3d.argtypes = [pf, C16, ctypes.c_uint32]
    L.qcd_conjPropagator.argtypes = [pf]
    L.qcd_subGaugeField.argtypes = [pf, pf, pf]
    L.qcd_addGaugeField.argtypes = [pf, pf, pf]
    L.qcd_scaleGaugeField.argtypes = [pf, ctypes.c_double]
    L.qcd_normGaugeField.argtypes = [pf]
    L.qcd_normGaugeField.restype = ctypes.c_double
    return L


def test_spinor_helpers_follow_the_gamma_table(lib):
    rng = np.random.default_rng(1)
    psi = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
    u = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    out = SP()
    for mu in range(4):
        lib.qcd_onePlusGamma(ctypes.byref(out), ctypes.byref(_sp(psi)), mu)
        assert np.abs(_npsp(out) - (np.eye(4) + G[mu]) @ psi).max() <= 1e-14
        lib.qcd_oneMinusGamma(ctypes.byref(out), ctypes.byref(_sp(psi)), mu)
        assert np.abs(_npsp(out) - (np.eye(4) - G[mu]) @ psi).max() <= 1e-14
    lib.qcd_applyU(ctypes.byref(out), ctypes.byref(_m3(u)), ctypes.byref(_sp(psi)))
    assert np.abs(_npsp(out) - psi @ u.T).max() <= 1e-14
    lib.qcd_applyUDagger(ctypes.byref(out), ctypes.byref(_m3(u)), ctypes.byref(_sp(psi)))
    assert np.abs(_npsp(out) - psi @ u.conj()).max() <= 1e-14
    lib.qcd_massTerms(ctypes.byref(out), ctypes.byref(_sp(psi)), 3.7, 0.2)
    assert np.abs(_npsp(out) - (3.7 * psi + 1j * 0.2 * G5 @ psi)).max() <= 1e-14
    v = _sp(psi)
    lib.qcd_phaseMul(ctypes.byref(v), C16(0.6, -0.8))
    assert np.abs(_npsp(v) - psi * (0.6 - 0.8j)).max() <= 1e-15


def test_3x3_helpers(lib):
    rng = np.random.default_rng(2)
    a = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    b = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    o = M3()
    for fn, want in (("qcd_sub3x3", a - b), ("qcd_add3x3", a + b), ("qcd_addAdjoint3x3", a + b.conj().T),
                     ("qcd_mul3x3", a @ b), ("qcd_mulAdjoint3x3", a @ b.conj().T)):
        getattr(lib, fn)(ctypes.byref(o), ctypes.byref(_m3(a)), ctypes.byref(_m3(b)))
        assert np.abs(_np3(o) - want).max() <= 1e-14, fn
    ma = _m3(a)   # the product may alias a factor
    lib.qcd_mul3x3(ctypes.byref(ma), ctypes.byref(ma), ctypes.byref(_m3(b)))
    assert np.abs(_np3(ma) - a @ b).max() <= 1e-14
    for fn, want in (("qcd_zero3x3", np.zeros((3, 3))), ("qcd_unit3x3", np.eye(3)), ("qcd_minusUnit3x3", -np.eye(3)),
                     ("qcd_dagger3x3", a.conj().T)):
        m = _m3(a)
        getattr(lib, fn)(ctypes.byref(m))
        assert np.abs(_np3(m) - want).max() <= 1e-15, fn
    m = _m3(a)
    lib.qcd_scale3x3(ctypes.byref(m), 2.5)
    assert np.abs(_np3(m) - 2.5 * a).max() <= 1e-14
    m = _m3(a)
    lib.qcd_cScale3x3(ctypes.byref(m), C16(0.3, 1.1))
    assert np.abs(_np3(m) - (0.3 + 1.1j) * a).max() <= 1e-14
    t = lib.qcd_trace3x3(ctypes.byref(_m3(a)))
    assert abs(complex(t.re, t.im) - np.trace(a)) <= 1e-14
    lib.qcd_copy3x3(ctypes.byref(o), ctypes.byref(_m3(a)))
    assert np.array_equal(_np3(o), a)


def test_projection_of_one_matrix_is_the_polar_factor(lib):
    rng = np.random.default_rng(3)
    for _ in range(20):
        m = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        w, _, vh = np.linalg.svd(m)
        polar = w @ vh
        polar = polar * np.exp(-1j * np.angle(np.linalg.det(polar)) / 3)
        g = _m3(m)
        lib.qcd_projectSU33x3(ctypes.byref(g))
        p = _np3(g)
        assert np.abs(p - polar).max() <= 1e-13
        assert abs(np.linalg.det(p) - 1) <= 1e-13


def test_propagator_and_gauge_arithmetic(lib):
    L = (4, 2, 2, 2)
    V = 32
    rng = np.random.default_rng(4)
    di = Q.DropIn(L)
    try:
        P = rng.normal(size=V * 144) + 1j * rng.normal(size=V * 144)
        p, q = di.propagator(), di.propagator()
        di.view(p, 144)[:] = P
        lib.qcd_copyPropagatorPropagator(ctypes.byref(q), ctypes.byref(p))
        assert np.array_equal(di.view(q, 144), P)
        lib.qcd_mulPropagatorC(ctypes.byref(q), C16(0.5, 2.0))
        assert np.abs(di.view(q, 144) - (0.5 + 2j) * P).max() <= 1e-14
        di.view(q, 144)[:] = P
        lib.qcd_mulPropagatorC3d(ctypes.byref(q), C16(-1.0, 0.0), 3)
        want = P.reshape(V, 144).copy()
        want[np.arange(V) % L[0] == 3] *= -1
        assert np.array_equal(di.view(q, 144).reshape(V, 144), want)
        lib.qcd_conjPropagator(ctypes.byref(q))
        assert np.array_equal(di.view(q, 144).reshape(V, 144), want.conj())
        U1 = Q.gauge_generate(L, 0.4, 5).reshape(-1)
        U2 = Q.gauge_generate(L, 0.4, 6).reshape(-1)
        a, b, c = di.gauge(U1), di.gauge(U2), di.gauge()
        lib.qcd_subGaugeField(ctypes.byref(c), ctypes.byref(a), ctypes.byref(b))
        assert np.array_equal(di.view(c, 36), U1 - U2)
        lib.qcd_addGaugeField(ctypes.byref(c), ctypes.byref(a), ctypes.byref(b))
        assert np.array_equal(di.view(c, 36), U1 + U2)
        lib.qcd_scaleGaugeField(ctypes.byref(c), 0.25)
        assert np.abs(di.view(c, 36) - 0.25 * (U1 + U2)).max() <= 1e-15
        n = lib.qcd_normGaugeField(ctypes.byref(a))
        tr = np.trace(U1.reshape(V, 4, 3, 3), axis1=2, axis2=3)
        assert abs(n - np.abs(tr).sum()) <= 1e-12 * n
    finally:
        di.close()
