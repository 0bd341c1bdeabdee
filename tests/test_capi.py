"""The C-ABI boundary without a GPU: exported symbols, host-side geometry,
allocation and error contracts of the drop-in API (lib/src/qcd_init.c)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

import qcd_amd as Q

HEADERS = ["qcd.h", "qcd_wilson.h", "qcd_smearing.h", "qcd_blas.h", "qcd_io.h", "qcd_observables.h", "qcd_rng.h",
           "qcd_stochastic.h", "qcd_hip.h"]


def declared_functions():
    """Names of every non-inline function declared in include/*.h."""
    names = set()
    for h in HEADERS:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(qcd_\w+)\s*\(", text, flags=re.M):
            line = text[m.start():text.find("\n", m.start())]
            if line.lstrip().startswith(("#", "typedef", "static")):
                continue
            names.add(m.group(1))
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", Q.LIB_PATH], capture_output=True, text=True, check=True)
    return {ln.split()[-1] for ln in out.stdout.splitlines() if " T " in ln}


def test_every_declared_function_is_exported():
    decl = declared_functions()
    assert len(decl) > 60
    missing = decl - exported_symbols()
    assert not missing, f"declared but not exported: {sorted(missing)}"
    lib = Q.lib()
    for name in decl:
        getattr(lib, name)


def test_c_header_compiles_and_links(tmp_path):
    """A C99 caller including qcd.h and qcd_hip.h compiles with gcc -Wall -Werror."""
    src = tmp_path / "t.c"
    src.write_text('#include <qcd.h>\n#include <qcd_hip.h>\n'
                   'int main(void){qcd_geometry g; (void)g; qcd_real_8 u[18]={1,0,0,0,0,0,0,0,1,0,0,0,0,0,0,0,1,0};'
                   'qcd_real_8 x[24],y[24]; int i; for(i=0;i<24;i++) y[i]=i;'
                   'qcd_APPLY_U(u,x,y); return x[5]==5.0 ? 0 : 1;}\n')
    exe = tmp_path / "t"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", os.path.dirname(Q.LIB_PATH), "-lqcd_hip", "-Wl,-rpath," + os.path.dirname(Q.LIB_PATH),
                    "-lm", "-o", str(exe)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_geometry_tables_match_reference_layout():
    L = (6, 4, 2, 8)
    di = Q.DropIn(L, theta=(np.pi, 0.0, 0.1, 0.0))
    try:
        g = di.geo
        V = int(np.prod(L))
        assert g.V == g.lV == V and g.V3 == V // L[0] and list(g.L) == list(L) and g.initialized == 1
        plus = np.ctypeslib.as_array((ctypes.c_uint64 * (4 * V)).from_address(g.plus)).reshape(V, 4)
        minus = np.ctypeslib.as_array((ctypes.c_uint64 * (4 * V)).from_address(g.minus)).reshape(V, 4)
        T, X, Y, Z = L
        for l in range(V):
            c = [l % T, (l // T) % X, (l // (T * X)) % Y, l // (T * X * Y)]
            for mu in range(4):
                cp, cm = list(c), list(c)
                cp[mu] = (cp[mu] + 1) % L[mu]
                cm[mu] = (cm[mu] - 1) % L[mu]
                assert plus[l, mu] == cp[0] + T * (cp[1] + X * (cp[2] + Y * cp[3]))
                assert minus[l, mu] == cm[0] + T * (cm[1] + X * (cm[2] + Y * cm[3]))
        # surface sites of the block (lib/src/qcd_init.c:212-225)
        assert g.edgePoints == V - (T - 2) * (X - 2) * max(Y - 2, 0) * (Z - 2)
        assert g.edge0Points == V // T - (X - 2) * max(Y - 2, 0) * (Z - 2)
        v = di.vector()
        u = di.gauge()
        p = di.propagator()
        assert v.initialized == u.initialized == p.initialized == 1
        di.view(v, 12)[:] = 1.0
        Q.lib().qcd_zeroVector(ctypes.byref(v))
        assert not di.view(v, 12).any()
    finally:
        di.close()


def test_geometry_rejects_what_the_reference_rejects():
    lib = Q.lib()
    g = Q.QcdGeometry()
    th = np.zeros(4)
    P1 = (ctypes.c_uint16 * 4)(1, 1, 1, 1)
    for L, P, nproc in [((4, 4, 4, 4), (2, 1, 1, 1), 2),      # more than one rank
                        ((4, 1, 4, 4), (1, 1, 1, 1), 1),      # extent below 2
                        ((3, 3, 3, 3), (1, 1, 1, 1), 1)]:     # odd volume
        rc = lib.qcd_initGeometry(ctypes.byref(g), (ctypes.c_uint16 * 4)(*L), (ctypes.c_uint16 * 4)(*P),
                                  th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 0, nproc)
        assert rc == 1
    assert lib.qcd_initGeometry(ctypes.byref(g), (ctypes.c_uint16 * 4)(4, 4, 4, 4), P1,
                                th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 0, 1) == 0
    lib.qcd_destroyGeometry(ctypes.byref(g))


def test_propagator_column_round_trip():
    """qcd_copyPropagatorVector / qcd_copyVectorPropagator (lib/src/qcd_blas.c:182-212)."""
    L = (4, 2, 2, 2)
    di = Q.DropIn(L)
    try:
        lib = Q.lib()
        p = di.propagator()
        pv = di.view(p, 144).reshape(-1, 4, 4, 3, 3)
        pv[:] = np.random.default_rng(0).normal(size=pv.shape)
        v = di.vector()
        lib.qcd_copyVectorPropagator(ctypes.byref(v), ctypes.byref(p), 2, 1)
        assert np.array_equal(di.view(v, 12).reshape(-1, 4, 3), pv[:, :, 2, :, 1])
        w = Q.spinor_gaussian(L, 5)
        di.view(v, 12)[:] = w.reshape(-1)
        lib.qcd_copyPropagatorVector(ctypes.byref(p), ctypes.byref(v), 3, 0)
        assert np.array_equal(pv[:, :, 3, :, 0], w)
    finally:
        di.close()


def test_library_without_gpu_fails_loudly():
    """Device entry points report an error (no silent CPU path) when no GPU is visible."""
    rc = Q.lib().qcd_hip_init(0)
    if rc == 0:
        pytest.skip("a GPU is visible")
    assert b"device" in Q.lib().qcd_hip_last_error().lower()


def test_eo_split_merge_round_trip():
    """Half-lattice order used by the even-odd solver: site l of parity p is entry l >> 1."""
    L = (6, 4, 2, 8)
    V = int(np.prod(L))
    a = np.arange(V * 12, dtype=np.float64).reshape(V, 4, 3)
    ev, od = Q.eo_split(a, L)
    assert ev.shape == od.shape == (V // 2, 4, 3)
    assert np.array_equal(Q.eo_merge(ev, od, L), a)
    T, X, Y = L[0], L[1], L[2]
    for l in range(V):
        t, x, y, z = l % T, (l // T) % X, (l // (T * X)) % Y, l // (T * X * Y)
        half = ev if (t + x + y + z) % 2 == 0 else od
        assert np.array_equal(half[l >> 1], a[l])


def test_drivers_fail_loudly_without_gpu(tmp_path):
    """The invert driver has no CPU path: without a GPU it reports an error and fails."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    exe = os.path.join(ROOT, "build", "bin", "invert")
    if not os.path.exists(exe):
        pytest.skip("drivers not built")
    ini = tmp_path / "invert.ini"
    ini.write_text("<processors_txyz>1 1 1 1</processors_txyz>\n<lattice_txyz>4 4 4 4</lattice_txyz>\n"
                   "<source_type>HMC_PROPAGATOR</source_type>\n<source>none</source>\n"
                   "<solution_type>HMC_PROPAGATOR</solution_type>\n<solution>none</solution>\n"
                   "<cfg_name>none</cfg_name>\n<N_restart>8</N_restart>\n<kappa>0.13</kappa>\n<mu>0</mu>\n")
    r = subprocess.run([exe, str(ini)], capture_output=True, text=True)
    assert r.returncode != 0 and "Error" in r.stderr
