"""CPU oracle for the Wilson Dslash / Gauss smearing / GCR path.

TEST INFRASTRUCTURE ONLY: a ctypes front end to ``liboracle.so`` (a plain-C
restatement of the reference algorithm, built from ``qcd_oracle.c`` by the
Makefile beside it).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package; the product
library never does.

Arrays are numpy ``complex128`` in the reference AoS layouts:
spinor ``[V, 4, 3]``, gauge ``[V, 4, 3, 3]``, site ``t + T*(x + X*(y + Y*z))``.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle.so with gcc (make)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        vp = ctypes.c_void_p
        d = ctypes.c_double
        i = ctypes.c_int
        lg = ctypes.c_long
        sig = {
            "orc_geo_create": (vp, [ctypes.POINTER(ctypes.c_int), _dp]),
            "orc_geo_destroy": (None, [vp]),
            "orc_geo_volume": (lg, [vp]),
            "orc_set_threads": (None, [i]),
            "orc_dslash_half": (None, [vp, _dp, _dp, _dp, d, d]),
            "orc_dslash_full": (None, [vp, _dp, _dp, _dp, d, d]),
            "orc_qtm": (None, [vp, _dp, _dp, _dp, d, d]),
            "orc_gauss3d": (None, [vp, _dp, _dp, d, i]),
            "orc_gauss3d_all": (None, [vp, _dp, _dp, d]),
            "orc_scale": (None, [_dp, d, lg]),
            "orc_sub": (None, [_dp, _dp, _dp, lg]),
            "orc_add": (None, [_dp, _dp, _dp, lg]),
            "orc_axpy": (None, [_dp, d, d, _dp, _dp, lg]),
            "orc_mulc": (None, [_dp, d, d, lg]),
            "orc_dot": (None, [_dp, _dp, _dp, lg]),
            "orc_norm": (d, [_dp, lg]),
            "orc_gcr": (i, [vp, _dp, _dp, _dp, d, d, i, i, i]),
            "orc_invert": (i, [vp, _dp, _dp, _dp, d, d, i, i, i, d, i,
                               ctypes.POINTER(ctypes.c_int), _dp]),
            "orc_b_minus_Dx": (d, [vp, _dp, _dp, _dp, d, d]),
            "orc_gauge_generate": (None, [_dp, lg, d, ctypes.c_uint64]),
            "orc_spinor_gaussian": (None, [_dp, lg, ctypes.c_uint64]),
            "orc_point_source": (None, [_dp, ctypes.POINTER(ctypes.c_int), i, i, i, i, i, i]),
            "orc_plaquette": (d, [vp, _dp]),
            "orc_project_su3": (None, [_dp]),
            "orc_ape3d": (None, [vp, _dp, _dp, d]),
            "orc_stout_exp": (None, [_dp]),
            "orc_stout": (i, [vp, _dp, _dp, d, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Geometry:
    """Neighbour tables of an L = (T, X, Y, Z) lattice with boundary phases theta."""

    def __init__(self, L, theta=(np.pi, 0.0, 0.0, 0.0)):
        self.L = tuple(int(x) for x in L)
        self.theta = tuple(float(x) for x in theta)
        self.V = int(np.prod(self.L))
        Lc = (ctypes.c_int * 4)(*self.L)
        th = np.ascontiguousarray(self.theta, dtype=np.float64)
        self._h = lib().orc_geo_create(Lc, _p(th))
        if not self._h:
            raise MemoryError("orc_geo_create failed")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_geo_destroy(self._h)
            self._h = None

    def spinor(self):
        return np.zeros((self.V, 4, 3), dtype=np.complex128)

    def gauge(self):
        return np.zeros((self.V, 4, 3, 3), dtype=np.complex128)


def set_threads(n):
    lib().orc_set_threads(int(n))


def dslash(geo, psi, u, m, tm=0.0, form="half"):
    out = np.empty_like(psi)
    f = lib().orc_dslash_half if form == "half" else lib().orc_dslash_full
    f(geo._h, _p(out), _p(psi), _p(u), float(m), float(tm))
    return out


def qtm(geo, psi, u, m, tm=0.0):
    out = np.empty_like(psi)
    lib().orc_qtm(geo._h, _p(out), _p(psi), _p(u), float(m), float(tm))
    return out


def gauss3d(geo, v, u, alpha, t):
    v = np.array(v, dtype=np.complex128, order="C")
    lib().orc_gauss3d(geo._h, _p(v), _p(u), float(alpha), int(t))
    return v


def gauss3d_all(geo, v, u, alpha):
    v = np.array(v, dtype=np.complex128, order="C")
    lib().orc_gauss3d_all(geo._h, _p(v), _p(u), float(alpha))
    return v


def dot(a, b):
    out = np.zeros(2)
    lib().orc_dot(_p(out), _p(a), _p(b), a.shape[0])
    return complex(out[0], out[1])


def norm(a):
    return lib().orc_norm(_p(a), a.shape[0])


def axpy(a, x, y):
    out = np.empty_like(y)
    lib().orc_axpy(_p(out), float(np.real(a)), float(np.imag(a)), _p(x), _p(y), x.shape[0])
    return out


def gcr(geo, src, u, kappa, mu=0.0, n=32, op=0, keep_q1=True):
    sol = np.zeros_like(src)
    rc = lib().orc_gcr(geo._h, _p(sol), _p(src), _p(u), float(kappa), float(mu), int(n), int(op), int(keep_q1))
    return rc, sol


def invert(geo, b, u, kappa, mu=0.0, n=32, op=0, keep_q1=True, max_res=1e-10, max_outer=10000):
    x = np.zeros_like(b)
    outer = ctypes.c_int(0)
    rel = np.zeros(1)
    rc = lib().orc_invert(geo._h, _p(x), _p(b), _p(u), float(kappa), float(mu), int(n), int(op),
                          int(keep_q1), float(max_res), int(max_outer), ctypes.byref(outer), _p(rel))
    return rc, x, outer.value, float(rel[0])


def b_minus_Dx(geo, x, b, u, kappa, mu=0.0):
    return lib().orc_b_minus_Dx(geo._h, _p(x), _p(b), _p(u), float(kappa), float(mu))


def gauge_generate(L, eps, seed):
    V = int(np.prod(L))
    u = np.zeros((V, 4, 3, 3), dtype=np.complex128)
    lib().orc_gauge_generate(_p(u), V, float(eps), ctypes.c_uint64(seed))
    return u


def spinor_gaussian(L, seed):
    V = int(np.prod(L))
    v = np.zeros((V, 4, 3), dtype=np.complex128)
    lib().orc_spinor_gaussian(_p(v), V, ctypes.c_uint64(seed))
    return v


def point_source(L, site, spin, col):
    V = int(np.prod(L))
    v = np.zeros((V, 4, 3), dtype=np.complex128)
    Lc = (ctypes.c_int * 4)(*L)
    lib().orc_point_source(_p(v), Lc, *[int(s) for s in site], int(spin), int(col))
    return v


def plaquette(geo, u):
    return lib().orc_plaquette(geo._h, _p(u))


def project_su3(m):
    """SU(3) projection (lib/src/qcd_blas.c:940-1364) of every 3x3 matrix of m (..., 3, 3)."""
    a = np.array(m, dtype=np.complex128, order="C").reshape(-1, 3, 3)
    for k in range(a.shape[0]):
        blk = np.ascontiguousarray(a[k])
        lib().orc_project_su3(_p(blk))
        a[k] = blk
    return a.reshape(np.shape(m))


def stout_exp(q):
    """exp(iQ) of Hermitian traceless 3x3 matrices q (..., 3, 3) (lib/src/qcd_smearing.c:293-370)."""
    a = np.array(q, dtype=np.complex128, order="C").reshape(-1, 3, 3)
    for k in range(a.shape[0]):
        blk = np.ascontiguousarray(a[k])
        lib().orc_stout_exp(_p(blk))
        a[k] = blk
    return a.reshape(np.shape(q))


def stout(geo, u, rho, ndim=3):
    """One stout smearing step (lib/src/qcd_smearing.c:374-476); returns the new field."""
    out = np.empty_like(u)
    if lib().orc_stout(geo._h, _p(out), _p(np.ascontiguousarray(u)), float(rho), int(ndim)):
        raise ValueError("ndim should be either 3 or 4")
    return out


def ape3d(geo, u, alpha):
    """One 3d APE smearing step (lib/src/qcd_smearing.c:205-273); returns the new field."""
    out = np.empty_like(u)
    lib().orc_ape3d(geo._h, _p(out), _p(np.ascontiguousarray(u)), float(alpha))
    return out
