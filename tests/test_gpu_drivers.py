"""The invert and b_minus_Dx drivers (counterparts of the reference's invert.c
and b_minus_Dx.c) end to end: files written with the library's own writers,
the programs run as a user would, the solutions checked by the CPU oracle."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT
from test_io import setup_lib

pytestmark = pytest.mark.gpu

L = (16, 8, 8, 8)
BIN = os.path.join(ROOT, "build", "bin")


def minus_i_g5(v):
    out = -1j * v
    out[:, 2:] *= -1
    return out


def test_invert_then_b_minus_Dx(gpu, tmp_path):
    Q = gpu
    lib = setup_lib()
    kappa, mu = 0.13, 0.02
    u = Q.gauge_generate(L, 0.3, 801)
    srcs = [Q.point_source(L, (1, 2, 3, 4), r // 3, r % 3) for r in range(12)]
    di = Q.DropIn(L)
    try:
        uu = di.gauge(u)
        assert lib.qcd_writeGaugeLime(str(tmp_path / "conf.lime").encode(), ctypes.byref(uu), b"synthetic") == 0
        vecs = (Q.QcdField * 12)()
        for r in range(12):
            assert lib.qcd_initVector(ctypes.byref(vecs[r]), ctypes.byref(di.geo)) == 0
            di.view(vecs[r], 12)[:] = srcs[r].reshape(-1)
        assert lib.qcd_writePropagatorHMC(str(tmp_path / "src.hmc").encode(), vecs, 0.0, kappa, mu) == 0
        for r in range(12):
            lib.qcd_destroyVector(ctypes.byref(vecs[r]))
        ini = tmp_path / "invert.ini"
        ini.write_text(f"<processors_txyz>1 1 1 1</processors_txyz>\n<lattice_txyz>{' '.join(map(str, L))}"
                       "</lattice_txyz>\n<source_type>HMC_PROPAGATOR</source_type>\n"
                       f"<source>{tmp_path / 'src.hmc'}</source>\n<solution_type>HMC_PROPAGATOR</solution_type>\n"
                       f"<solution>{tmp_path / 'sol.hmc'}</solution>\n<cfg_name>{tmp_path / 'conf.lime'}</cfg_name>\n"
                       f"<N_restart>32</N_restart>\n<kappa>{kappa}</kappa>\n<mu>{mu}</mu>\n"
                       "<precision>f16</precision>\n<maxRes>1e-10</maxRes>\n")
        r = subprocess.run([os.path.join(BIN, "invert"), str(ini)], capture_output=True, text=True, timeout=300)
        print(r.stdout[-2000:], r.stderr[-2000:])
        assert r.returncode == 0
        assert r.stdout.count("relative residue") == 12
        # invert.c's system: Q x = src, i.e. D x = -i g5 src, checked by the oracle
        geo = O.Geometry(L)
        v = di.vector()
        for k in (0, 5, 11):
            assert lib.qcd_getVector(str(tmp_path / "sol.hmc").encode(), 1, k // 3, k % 3, ctypes.byref(v)) == 0
            x = di.view(v, 12).reshape(-1, 4, 3).copy()
            assert O.b_minus_Dx(geo, x, minus_i_g5(srcs[k]), u, kappa, mu) <= 1e-10
        # b_minus_Dx on solution (nu, c2) = (2, 1) against b = -i g5 src
        k = 7
        bb = di.vector(minus_i_g5(srcs[k]))
        assert lib.qcd_writeVector(str(tmp_path / "b.lime").encode(), 4, 0, 0, ctypes.byref(bb)) == 0
        ini2 = tmp_path / "bmdx.ini"

        def run_check(nu, c2, out):
            ini2.write_text(f"<processors_txyz>1 1 1 1</processors_txyz>\n<lattice_txyz>{' '.join(map(str, L))}"
                            f"</lattice_txyz>\n<cfg_name>{tmp_path / 'conf.lime'}</cfg_name>\n"
                            f"<src_name>{tmp_path / 'b.lime'}</src_name>\n<sol_name>{tmp_path / 'sol.hmc'}</sol_name>\n"
                            f"<src_col>{c2}</src_col>\n<src_spin>{nu}</src_spin>\n<kappa>{kappa}</kappa>\n"
                            f"<muT>{mu}<muT>\n<eps>1e-9</eps>\n<out_name>{out}</out_name>\n")
            return subprocess.run([os.path.join(BIN, "b_minus_Dx"), str(ini2)], capture_output=True, text=True,
                                  timeout=120)

        ok = run_check(k // 3, k % 3, tmp_path / "x.lime")
        print(ok.stdout[-1500:], ok.stderr[-1500:])
        assert ok.returncode == 0 and "plaquette" in ok.stdout
        rel = float(ok.stdout.split("|Dx-b| / |b|=")[1].split()[0])
        assert rel <= 1e-10
        assert lib.qcd_getVector(str(tmp_path / "x.lime").encode(), 3, 0, 0, ctypes.byref(v)) == 0
        assert lib.qcd_getVector(str(tmp_path / "sol.hmc").encode(), 1, k // 3, k % 3, ctypes.byref(bb)) == 0
        assert np.array_equal(di.view(v, 12), di.view(bb, 12))
        # the wrong solution vector for this source is reported and fails
        bad = run_check(0, 0, tmp_path / "y.lime")
        assert bad.returncode != 0 and "HIGH RESIDUE" in bad.stdout
    finally:
        di.close()
