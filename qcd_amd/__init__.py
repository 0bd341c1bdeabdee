"""MI355X-native Wilson Dslash / Gauss smearing / propagator solves.

Python front end to ``libqcd_hip.so`` (built from ``qcd_amd/csrc`` for gfx950).
Everything here is a thin ctypes layer over the C ABI declared in
``include/qcd_hip.h`` (device-resident fast path) and ``include/qcd*.h`` (the
reference's drop-in API).  There is no CPU fallback: if the shared library is
missing, importing succeeds but every call raises ``RuntimeError``.

Host arrays are numpy ``complex128`` in the reference layouts
(``include/qcd.h``): spinor ``[V, 4, 3]``, gauge ``[V, 4, 3, 3]``, site index
``t + T*(x + X*(y + Y*z))`` with ``L = (T, X, Y, Z)``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqcd_hip.so")

F64, F32, F16 = 0, 1, 2
OP_D, OP_Q = 0, 1
PREC = {"f64": F64, "f32": F32, "f16": F16}

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class QcdError(RuntimeError):
    pass


# ------------------------------------------------------------ reference structs
class QcdGeometry(ctypes.Structure):
    """ctypes mirror of ``qcd_geometry`` (include/qcd.h)."""
    _fields_ = [
        ("V", ctypes.c_uint64), ("V3", ctypes.c_uint64), ("L", ctypes.c_uint16 * 4),
        ("lV", ctypes.c_uint64), ("lV3", ctypes.c_uint64), ("lL", ctypes.c_uint16 * 4),
        ("Pos", ctypes.c_uint16 * 4), ("theta", ctypes.c_double * 4), ("myid", ctypes.c_uint32),
        ("nproc", ctypes.c_uint32), ("Pplus", ctypes.c_uint32 * 4), ("Pminus", ctypes.c_uint32 * 4),
        ("plus", _vp), ("minus", _vp), ("plus3d", _vp), ("minus3d", _vp), ("eo", _vp),
        ("startBminus", ctypes.c_uint64 * 4), ("startBplus", ctypes.c_uint64 * 4),
        ("startBminus3d", ctypes.c_uint64 * 4), ("startBplus3d", ctypes.c_uint64 * 4),
        ("boundarySize", ctypes.c_uint64), ("boundarySize3d", ctypes.c_uint64),
        ("edge", _vp), ("edgePoints", ctypes.c_uint64), ("edge0", _vp), ("edge0Points", ctypes.c_uint64),
        ("requests", ctypes.c_int * 32), ("numOfRequests", ctypes.c_int),
        ("stypeV", ctypes.c_int * 4), ("rtypeV", ctypes.c_int * 4),
        ("stypeU", ctypes.c_int * 4), ("rtypeU", ctypes.c_int * 4),
        ("stypeT", ctypes.c_int * 4), ("rtypeT", ctypes.c_int * 4),
        ("stypeP", ctypes.c_int * 4), ("rtypeP", ctypes.c_int * 4),
        ("initialized", ctypes.c_uint16),
    ]


class QcdField(ctypes.Structure):
    """ctypes mirror of ``qcd_vector`` / ``qcd_gaugeField`` / ``qcd_propagator``."""
    _fields_ = [("D", _vp), ("Bplus", _vp * 4), ("Bminus", _vp * 4),
                ("geo", ctypes.POINTER(QcdGeometry)), ("initialized", ctypes.c_uint16)]


class QcdComplex(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


class SolverParams(ctypes.Structure):
    """qcd_hip_solver_params (include/qcd_hip.h)."""
    _fields_ = [("n_restart", ctypes.c_int), ("max_outer", ctypes.c_int), ("tol", ctypes.c_double),
                ("inner_prec", ctypes.c_int), ("op", ctypes.c_int), ("keep_q1", ctypes.c_int),
                ("batch", ctypes.c_int), ("eo", ctypes.c_int), ("graph", ctypes.c_int)]


class Request(ctypes.Structure):
    """qcd_hip_request (include/qcd_hip.h)."""
    _fields_ = [("arrival_s", ctypes.c_double), ("L", ctypes.c_int * 4), ("kappa", ctypes.c_double),
                ("eps", ctypes.c_double), ("seed", ctypes.c_uint64), ("nrhs", ctypes.c_int),
                ("inner_prec", ctypes.c_int), ("n_restart", ctypes.c_int), ("tol", ctypes.c_double)]


class Result(ctypes.Structure):
    """qcd_hip_result (include/qcd_hip.h)."""
    _fields_ = [("queued_s", ctypes.c_double), ("service_s", ctypes.c_double), ("latency_s", ctypes.c_double),
                ("rc", ctypes.c_int), ("outer", ctypes.c_int), ("iters", ctypes.c_int), ("relres", ctypes.c_double),
                ("bytes", ctypes.c_double), ("xnorm2", ctypes.c_double)]


class SolverStats(ctypes.Structure):
    """qcd_hip_solver_stats (include/qcd_hip.h)."""
    _fields_ = [("outer", ctypes.c_int), ("inner_iters", ctypes.c_int), ("relres", ctypes.c_double),
                ("seconds", ctypes.c_double)]


def lib():
    """Load libqcd_hip.so (raises QcdError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise QcdError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    i, d, lg, sz = ctypes.c_int, ctypes.c_double, ctypes.c_long, ctypes.c_size_t
    pf = ctypes.POINTER(QcdField)
    pg = ctypes.POINTER(QcdGeometry)
    sig = {
        "qcd_hip_last_error": (ctypes.c_char_p, []),
        "qcd_hip_init": (i, [i]),
        "qcd_hip_device_name": (i, [ctypes.c_char_p, i]),
        "qcd_hip_synchronize": (i, []),
        "qcd_hip_bytes_in_use": (sz, []),
        "qcd_hip_lattice_create": (i, [ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_int), _dp]),
        "qcd_hip_lattice_destroy": (None, [_vp]),
        "qcd_hip_lattice_volume": (lg, [_vp]),
        "qcd_hip_spinor_create": (i, [ctypes.POINTER(_vp), _vp, i, i]),
        "qcd_hip_gauge_create": (i, [ctypes.POINTER(_vp), _vp, i]),
        "qcd_hip_field_destroy": (None, [_vp]),
        "qcd_hip_spinor_upload": (i, [_vp, i, _dp]),
        "qcd_hip_spinor_download": (i, [_vp, i, _dp]),
        "qcd_hip_gauge_upload": (i, [_vp, _dp]),
        "qcd_hip_gauge_download": (i, [_vp, _dp]),
        "qcd_hip_field_convert": (i, [_vp, _vp]),
        "qcd_hip_gauge_generate_host": (i, [_dp, lg, d, ctypes.c_uint64]),
        "qcd_hip_spinor_gaussian_host": (i, [_dp, lg, ctypes.c_uint64]),
        "qcd_hip_dslash": (i, [_vp, _vp, _vp, d, d, i]),
        "qcd_hip_dslash_schur": (i, [_vp, _vp, _vp, d, d]),
        "qcd_hip_eo_set_zlen": (i, [i]),
        "qcd_hip_ape_smear": (i, [_vp, _vp, d, i]),
        "qcd_hip_stout_smear": (i, [_vp, _vp, d, i, i]),
        "qcd_hip_plaquette": (i, [_vp, _dp]),
        "qcd_hip_dslash_time": (i, [_vp, _vp, _vp, d, d, i, i, i, _dp]),
        "qcd_hip_dslash_tune": (i, [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_int), _dp]),
        "qcd_hip_dslash_set_config": (i, [_vp, _vp, i, i]),
        "qcd_hip_dslash_candidates": (i, [ctypes.POINTER(ctypes.c_int), i]),
        "qcd_hip_gauss_smear": (i, [_vp, _vp, d, i, i]),
        "qcd_hip_norm2": (i, [_vp, i, _dp]),
        "qcd_hip_dot": (i, [_vp, i, _vp, i, _dp]),
        "qcd_hip_axpy": (i, [_vp, i, d, d, _vp, i]),
        "qcd_hip_lincomb": (i, [_vp, i, d, d, _vp, i, d, d, _vp, i, i]),
        "qcd_hip_solver_default_params": (None, [ctypes.POINTER(SolverParams)]),
        "qcd_hip_solver_create": (i, [ctypes.POINTER(_vp), _vp, ctypes.POINTER(SolverParams)]),
        "qcd_hip_solver_destroy": (None, [_vp]),
        "qcd_hip_solver_bytes": (sz, [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(SolverParams)]),
        "qcd_hip_solver_set_gauge": (i, [_vp, _vp]),
        "qcd_hip_solver_solve": (i, [_vp, _vp, _vp, d, d, ctypes.POINTER(SolverStats)]),
        "qcd_hip_invert": (i, [_vp, _vp, _vp, d, d, ctypes.POINTER(SolverParams), ctypes.POINTER(SolverStats)]),
        "qcd_hip_b_minus_Dx": (i, [_vp, i, _vp, i, _vp, d, d, _dp]),
        "qcd_hip_b_minus_Dx_noalloc": (i, [_vp, i, _vp, i, _vp, d, d, _vp, i, _dp]),
        "qcd_hip_gauge_generate": (i, [_vp, d, ctypes.c_uint64]),
        "qcd_hip_point_source": (i, [_vp, i, ctypes.POINTER(ctypes.c_int), i, i]),
        "qcd_hip_z4_source": (i, [_vp, i, i, ctypes.c_uint64]),
        "qcd_rng_init": (None, [ctypes.c_uint, i]),
        "qcd_rng_finalize": (None, []),
        "qcd_get_rand_4": (ctypes.c_ulong, []),
        "qcd_z2Vector": (None, [pf, i]),
        "qcd_hip_release_cache": (None, []),
        "qcd_hip_serve": (i, [ctypes.POINTER(Request), i, i, sz, i, ctypes.POINTER(Result), _dp,
                              ctypes.POINTER(_dp)]),
        # drop-in reference API
        "qcd_initGeometry": (i, [pg, ctypes.POINTER(ctypes.c_uint16), ctypes.POINTER(ctypes.c_uint16), _dp, i, i]),
        "qcd_destroyGeometry": (None, [pg]),
        "qcd_initVector": (i, [pf, pg]),
        "qcd_destroyVector": (None, [pf]),
        "qcd_initGaugeField": (i, [pf, pg]),
        "qcd_destroyGaugeField": (None, [pf]),
        "qcd_initPropagator": (i, [pf, pg]),
        "qcd_destroyPropagator": (None, [pf]),
        "qcd_applyWilsonTMOp": (i, [pf, pf, pf, d, d]),
        "qcd_applyWilsonTMOpNoScripts": (i, [pf, pf, pf, d, d]),
        "qcd_applyWilsonTMOpNoOptim": (i, [pf, pf, pf, d, d]),
        "qcd_applyQTMOp": (i, [pf, pf, pf, d, d]),
        "qcd_hip_applyWilsonTMOpPropagator": (i, [pf, pf, pf, d, d]),
        "qcd_hip_applyQTMOpPropagator": (i, [pf, pf, pf, d, d]),
        "qcd_hip_propagator_upload": (i, [_vp, i, _dp]),
        "qcd_hip_propagator_download": (i, [_vp, i, _dp]),
        "qcd_hip_bindGaugeField": (i, [pf]),
        "qcd_hip_unbindGaugeField": (None, []),
        "qcd_gaussIteration3d": (i, [pf, pf, d, ctypes.c_uint32]),
        "qcd_apeSmear3d": (i, [pf, pf, d]),
        "qcd_stoutsmearing": (i, [pf, pf, d, i]),
        "qcd_projectSU33d": (None, [pf]),
        "qcd_gaussIteration3dAll": (i, [pf, pf, d, ctypes.c_uint16]),
        "qcd_zeroVector": (None, [pf]),
        "qcd_zeroGaugeField": (None, [pf]),
        "qcd_copyGaugeField": (None, [pf, pf]),
        "qcd_copyVector": (None, [pf, pf]),
        "qcd_copyVectorPropagator": (None, [pf, pf, ctypes.c_uint16, ctypes.c_uint16]),
        "qcd_copyPropagatorVector": (None, [pf, pf, ctypes.c_uint16, ctypes.c_uint16]),
        "qcd_mulVectorC": (None, [pf, QcdComplex]),
        "qcd_scaleVector": (None, [pf, d]),
        "qcd_scaleVector3d": (None, [pf, d, ctypes.c_uint32]),
        "qcd_subVector": (None, [pf, pf, pf]),
        "qcd_addVector": (None, [pf, pf, pf]),
        "qcd_axpyVector": (None, [pf, QcdComplex, pf, pf]),
        "qcd_addVector3d": (None, [pf, pf, pf, ctypes.c_uint32]),
        "qcd_mulAdjointVectorVector": (QcdComplex, [pf, pf]),
        "qcd_normVector": (d, [pf]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        msg = lib().qcd_hip_last_error().decode(errors="replace")
        raise QcdError(f"{what} failed ({rc}): {msg}")


def _p(a):
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(_dp)


def init(device=-1):
    _check(lib().qcd_hip_init(int(device)), "qcd_hip_init")


def device_name():
    buf = ctypes.create_string_buffer(256)
    _check(lib().qcd_hip_device_name(buf, 256), "qcd_hip_device_name")
    return buf.value.decode()


def synchronize():
    _check(lib().qcd_hip_synchronize(), "qcd_hip_synchronize")


def bytes_in_use():
    return int(lib().qcd_hip_bytes_in_use())


# ------------------------------------------------------------ device objects
class Lattice:
    """A T x X x Y x Z lattice with fermion boundary phases theta."""

    def __init__(self, L, theta=(np.pi, 0.0, 0.0, 0.0)):
        self.L = tuple(int(x) for x in L)
        self.theta = tuple(float(x) for x in theta)
        self.V = int(np.prod(self.L))
        h = _vp()
        th = np.ascontiguousarray(self.theta, dtype=np.float64)
        _check(lib().qcd_hip_lattice_create(ctypes.byref(h), (ctypes.c_int * 4)(*self.L), _p(th)),
               "qcd_hip_lattice_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qcd_hip_lattice_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class _Field:
    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qcd_hip_field_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Spinor(_Field):
    """nrhs spinor fields resident on the GPU in precision prec."""

    def __init__(self, lat, prec=F64, nrhs=1):
        self.lat, self.prec, self.nrhs = lat, prec, nrhs
        h = _vp()
        _check(lib().qcd_hip_spinor_create(ctypes.byref(h), lat.h, int(prec), int(nrhs)), "qcd_hip_spinor_create")
        self.h = h

    def upload(self, a, rhs=0):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.size == self.lat.V * 12
        _check(lib().qcd_hip_spinor_upload(self.h, int(rhs), _p(a)), "qcd_hip_spinor_upload")
        return self

    def download(self, rhs=0):
        a = np.empty((self.lat.V, 4, 3), dtype=np.complex128)
        _check(lib().qcd_hip_spinor_download(self.h, int(rhs), _p(a)), "qcd_hip_spinor_download")
        return a

    def norm2(self, rhs=0):
        out = np.zeros(1)
        _check(lib().qcd_hip_norm2(self.h, int(rhs), _p(out)), "qcd_hip_norm2")
        return float(out[0])


class Gauge(_Field):
    """A gauge field resident on the GPU in precision prec."""

    def __init__(self, lat, prec=F64):
        self.lat, self.prec = lat, prec
        h = _vp()
        _check(lib().qcd_hip_gauge_create(ctypes.byref(h), lat.h, int(prec)), "qcd_hip_gauge_create")
        self.h = h

    def upload(self, u):
        u = np.ascontiguousarray(u, dtype=np.complex128)
        assert u.size == self.lat.V * 36
        _check(lib().qcd_hip_gauge_upload(self.h, _p(u)), "qcd_hip_gauge_upload")
        return self

    def download(self):
        u = np.empty((self.lat.V, 4, 3, 3), dtype=np.complex128)
        _check(lib().qcd_hip_gauge_download(self.h, _p(u)), "qcd_hip_gauge_download")
        return u


def convert(dst, src):
    _check(lib().qcd_hip_field_convert(dst.h, src.h), "qcd_hip_field_convert")


def gauge_generate(L, eps, seed):
    """Host generator: U = reunitarize(exp(i eps H)), H seeded traceless Hermitian."""
    V = int(np.prod(L))
    u = np.empty((V, 4, 3, 3), dtype=np.complex128)
    _check(lib().qcd_hip_gauge_generate_host(_p(u), V, float(eps), ctypes.c_uint64(seed)), "gauge_generate")
    return u


def gauge_generate_device(gauge64, eps, seed):
    """The generator recipe run on the GPU into an FP64 gauge field."""
    _check(lib().qcd_hip_gauge_generate(gauge64.h, float(eps), ctypes.c_uint64(seed)), "qcd_hip_gauge_generate")


def z4_source(f, rhs, t_slice, seed):
    """rhs of an FP64 device spinor = Z4 noise on time slice t_slice (all slices when negative)."""
    _check(lib().qcd_hip_z4_source(f.h, int(rhs), int(t_slice), int(seed)), "qcd_hip_z4_source")


def point_source_device(f, rhs, site, spin, col):
    """rhs of an FP64 device spinor = delta at site (t, x, y, z) on (spin, col)."""
    _check(lib().qcd_hip_point_source(f.h, int(rhs), (ctypes.c_int * 4)(*site), int(spin), int(col)),
           "qcd_hip_point_source")


def spinor_gaussian(L, seed):
    V = int(np.prod(L))
    v = np.empty((V, 4, 3), dtype=np.complex128)
    _check(lib().qcd_hip_spinor_gaussian_host(_p(v), V, ctypes.c_uint64(seed)), "spinor_gaussian")
    return v


def point_source(L, site, spin, col):
    V = int(np.prod(L))
    T, X, Y, _ = L
    t, x, y, z = site
    v = np.zeros((V, 4, 3), dtype=np.complex128)
    v[t + T * (x + X * (y + Y * z)), spin, col] = 1.0
    return v


# ------------------------------------------------------------ operators / BLAS
def dslash(out, inp, gauge, m, tm=0.0, op=OP_D):
    _check(lib().qcd_hip_dslash(out.h, inp.h, gauge.h, float(m), float(tm), int(op)), "qcd_hip_dslash")


def eo_set_zlen(zlen):
    """z steps per lane of the even-odd hops (0 = automatic, 1 = one site per lane)."""
    _check(lib().qcd_hip_eo_set_zlen(int(zlen)), "qcd_hip_eo_set_zlen")


def dslash_schur(out, inp, gauge, m, tm=0.0):
    """Even-odd Schur complement D^ = M - H_oe M^-1 H_eo on odd-site half-lattice fields."""
    _check(lib().qcd_hip_dslash_schur(out.h, inp.h, gauge.h, float(m), float(tm)), "qcd_hip_dslash_schur")


def ape_smear(out, inp, alpha, nsteps=1):
    """nsteps of 3d APE smearing of the FP64 gauge field inp into out (lib/src/qcd_smearing.c:205-273)."""
    _check(lib().qcd_hip_ape_smear(out.h, inp.h, float(alpha), int(nsteps)), "qcd_hip_ape_smear")


def stout_smear(out, inp, rho, ndim=3, nsteps=1):
    """nsteps of stout smearing of the FP64 gauge field inp into out (lib/src/qcd_smearing.c:374-476)."""
    _check(lib().qcd_hip_stout_smear(out.h, inp.h, float(rho), int(ndim), int(nsteps)), "qcd_hip_stout_smear")


def eo_split(a, L):
    """(even, odd) half-lattice arrays of a full-lattice array [V, ...]: entry l >> 1 of parity p."""
    T, X, Y, Z = L
    t, x, y, z = np.meshgrid(np.arange(T), np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    lex = (t + T * (x + X * (y + Y * z))).reshape(-1)
    par = ((t + x + y + z) % 2).reshape(-1)
    order = np.argsort(lex)
    lex, par = lex[order], par[order]
    return a[lex[par == 0]], a[lex[par == 1]]


def eo_merge(ev, od, L):
    """Inverse of eo_split."""
    T, X, Y, Z = L
    V = T * X * Y * Z
    l = np.arange(V)
    t, x = l % T, (l // T) % X
    y, z = (l // (T * X)) % Y, l // (T * X * Y)
    par = (t + x + y + z) % 2
    out = np.empty((V,) + ev.shape[1:], dtype=ev.dtype)
    out[par == 0] = ev
    out[par == 1] = od
    return out


def dslash_time(out, inp, gauge, m, tm=0.0, op=OP_D, warmup=5, iters=20):
    ms = np.zeros(1)
    _check(lib().qcd_hip_dslash_time(out.h, inp.h, gauge.h, float(m), float(tm), int(op), int(warmup), int(iters),
                                     _p(ms)), "qcd_hip_dslash_time")
    return float(ms[0])


def dslash_tune(out, inp, gauge):
    """Autotune the Dslash launch configuration; returns (map, occ, ms)."""
    cfg = (ctypes.c_int * 2)()
    ms = np.zeros(1)
    _check(lib().qcd_hip_dslash_tune(out.h, inp.h, gauge.h, cfg, _p(ms)), "qcd_hip_dslash_tune")
    return int(cfg[0]), int(cfg[1]), float(ms[0])


def dslash_candidates():
    """The traversal candidates (map values) the autotuner times."""
    n = lib().qcd_hip_dslash_candidates(None, 0)
    buf = (ctypes.c_int * n)()
    lib().qcd_hip_dslash_candidates(buf, n)
    return list(buf)


def dslash_set_config(inp, gauge, map_, occ):
    _check(lib().qcd_hip_dslash_set_config(inp.h, gauge.h, int(map_), int(occ)), "qcd_hip_dslash_set_config")


def gauss_smear(v, gauge, alpha, t, niter):
    _check(lib().qcd_hip_gauss_smear(v.h, gauge.h, float(alpha), int(t), int(niter)), "qcd_hip_gauss_smear")


def dot(a, b, ra=0, rb=0):
    out = np.zeros(2)
    _check(lib().qcd_hip_dot(a.h, int(ra), b.h, int(rb), _p(out)), "qcd_hip_dot")
    return complex(out[0], out[1])


def axpy(y, a, x, ry=0, rx=0):
    _check(lib().qcd_hip_axpy(y.h, int(ry), float(np.real(a)), float(np.imag(a)), x.h, int(rx)), "qcd_hip_axpy")


def lincomb(out, a, x, b=0.0, y=None, ro=0, rx=0, ry=0, t0=-1):
    _check(lib().qcd_hip_lincomb(out.h, int(ro), float(np.real(a)), float(np.imag(a)), x.h, int(rx),
                                 float(np.real(b)), float(np.imag(b)), y.h if y is not None else None, int(ry),
                                 int(t0)), "qcd_hip_lincomb")


# ------------------------------------------------------------ solver
def solver_params(**kw):
    """Default qcd_hip_solver_params with keyword overrides (inner_prec may be 'f64'/'f32'/'f16')."""
    p = SolverParams()
    lib().qcd_hip_solver_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k == "inner_prec" and isinstance(v, str):
            v = PREC[v]
        if k == "op" and isinstance(v, str):
            v = OP_Q if v == "Q" else OP_D
        setattr(p, k, v)
    return p


def solver_bytes(L, **kw):
    """Device bytes a Solver with these parameters allocates on lattice extents L."""
    p = solver_params(**kw)
    return int(lib().qcd_hip_solver_bytes((ctypes.c_int * 4)(*[int(x) for x in L]), ctypes.byref(p)))


class Solver:
    """Resident solver workspace (invert.c's GCR + FP64 outer refinement)."""

    def __init__(self, lat, gauge64, **kw):
        self.lat = lat
        self.params = solver_params(**kw)
        h = _vp()
        _check(lib().qcd_hip_solver_create(ctypes.byref(h), lat.h, ctypes.byref(self.params)),
               "qcd_hip_solver_create")
        self.h = h
        self.set_gauge(gauge64)

    def set_gauge(self, gauge64):
        self.gauge = gauge64
        _check(lib().qcd_hip_solver_set_gauge(self.h, gauge64.h), "qcd_hip_solver_set_gauge")

    def solve(self, x, b, kappa, mu=0.0, check=True):
        st = SolverStats()
        rc = lib().qcd_hip_solver_solve(self.h, x.h, b.h, float(kappa), float(mu), ctypes.byref(st))
        if check:
            _check(rc, "qcd_hip_solver_solve")
        return {"rc": rc, "outer": st.outer, "inner_iters": st.inner_iters, "relres": st.relres,
                "seconds": st.seconds}

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qcd_hip_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def b_minus_Dx(x, b, gauge64, kappa, mu=0.0, rx=0, rb=0):
    rel = np.zeros(1)
    _check(lib().qcd_hip_b_minus_Dx(x.h, int(rx), b.h, int(rb), gauge64.h, float(kappa), float(mu), _p(rel)),
           "qcd_hip_b_minus_Dx")
    return float(rel[0])


def b_minus_Dx_noalloc(x, b, gauge64, kappa, r, mu=0.0, rx=0, rb=0, rr=0):
    """b_minus_Dx without device allocations; r[rr] receives b - D x."""
    rel = np.zeros(1)
    _check(lib().qcd_hip_b_minus_Dx_noalloc(x.h, int(rx), b.h, int(rb), gauge64.h, float(kappa), float(mu), r.h,
                                            int(rr), _p(rel)), "qcd_hip_b_minus_Dx_noalloc")
    return float(rel[0])


# ------------------------------------------------------------ serving
def serve(requests, workers=4, budget_bytes=200 << 30, realtime=True, want_solutions=False):
    """Run the serving loop over a list of request dicts (keys of qcd_hip_request;
    inner_prec may be a string).  Returns (results list of dicts, wall seconds,
    solutions or None)."""
    n = len(requests)
    arr = (Request * n)()
    for k, r in enumerate(requests):
        q = arr[k]
        q.arrival_s = float(r.get("arrival_s", 0.0))
        for d in range(4):
            q.L[d] = int(r["L"][d])
        q.kappa = float(r["kappa"])
        q.eps = float(r["eps"])
        q.seed = int(r["seed"])
        q.nrhs = int(r.get("nrhs", 1))
        ip = r.get("inner_prec", "f16")
        q.inner_prec = PREC[ip] if isinstance(ip, str) else int(ip)
        q.n_restart = int(r.get("n_restart", 32))
        q.tol = float(r.get("tol", 1e-10))
    res = (Result * n)()
    wall = np.zeros(1)
    sols = None
    sol_ptrs = None
    if want_solutions:
        sols = [np.zeros((r.get("nrhs", 1), int(np.prod(r["L"])), 4, 3), dtype=np.complex128) for r in requests]
        sol_ptrs = (_dp * n)(*[a.ctypes.data_as(_dp) for a in sols])
    rc = lib().qcd_hip_serve(arr, n, int(workers), ctypes.c_size_t(int(budget_bytes)), int(bool(realtime)), res,
                             _p(wall), sol_ptrs)
    _check(rc, "qcd_hip_serve")
    out = [{f: getattr(res[k], f) for f, _ in Result._fields_} for k in range(n)]
    return out, float(wall[0]), sols


# ------------------------------------------------------------ drop-in helpers
class DropIn:
    """Host structs of the reference API, allocated by the library's qcd_init*."""

    def __init__(self, L, theta=(np.pi, 0.0, 0.0, 0.0)):
        self.L = tuple(int(x) for x in L)
        self.V = int(np.prod(self.L))
        self.geo = QcdGeometry()
        Lc = (ctypes.c_uint16 * 4)(*self.L)
        Pc = (ctypes.c_uint16 * 4)(1, 1, 1, 1)
        th = np.ascontiguousarray(theta, dtype=np.float64)
        rc = lib().qcd_initGeometry(ctypes.byref(self.geo), Lc, Pc, _p(th), 0, 1)
        if rc:
            raise QcdError("qcd_initGeometry failed")
        self._fields = []

    def vector(self, data=None):
        v = QcdField()
        if lib().qcd_initVector(ctypes.byref(v), ctypes.byref(self.geo)):
            raise QcdError("qcd_initVector failed")
        self._fields.append(("v", v))
        if data is not None:
            self.view(v, 12)[:] = np.ascontiguousarray(data, dtype=np.complex128).reshape(-1)
        return v

    def gauge(self, data=None):
        u = QcdField()
        if lib().qcd_initGaugeField(ctypes.byref(u), ctypes.byref(self.geo)):
            raise QcdError("qcd_initGaugeField failed")
        self._fields.append(("u", u))
        if data is not None:
            self.view(u, 36)[:] = np.ascontiguousarray(data, dtype=np.complex128).reshape(-1)
        return u

    def propagator(self):
        p = QcdField()
        if lib().qcd_initPropagator(ctypes.byref(p), ctypes.byref(self.geo)):
            raise QcdError("qcd_initPropagator failed")
        self._fields.append(("p", p))
        return p

    def view(self, f, per_site):
        """numpy complex128 view of the host data of a drop-in field."""
        buf = (ctypes.c_double * (2 * per_site * self.V)).from_address(f.D)
        return np.frombuffer(buf, dtype=np.complex128)

    def close(self):
        if _lib is None:
            return
        for kind, f in self._fields:
            if kind == "v":
                _lib.qcd_destroyVector(ctypes.byref(f))
            elif kind == "u":
                _lib.qcd_destroyGaugeField(ctypes.byref(f))
            else:
                _lib.qcd_destroyPropagator(ctypes.byref(f))
        self._fields = []
        if self.geo.initialized:
            _lib.qcd_destroyGeometry(ctypes.byref(self.geo))
            self.geo.initialized = 0
