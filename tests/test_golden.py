"""Committed regression fixtures (tests/golden/wilson_4x4x2x2.npz, written by
tests/golden/make_golden.py from the pinned oracle) against the oracle on the
CPU and the HIP path on the GPU."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "wilson_4x4x2x2.npz")


@pytest.fixture(scope="module")
def gold():
    d = np.load(GOLDEN)   # plain arrays, no pickles
    return {k: d[k] for k in d.files}


def test_oracle_reproduces_fixtures(gold):
    L = tuple(int(v) for v in gold["L"])
    geo = O.Geometry(L)
    u = O.gauge_generate(L, float(gold["eps"]), int(gold["seed_u"]))
    psi = O.spinor_gaussian(L, int(gold["seed_psi"]))
    assert np.array_equal(u, gold["u"]) and np.array_equal(psi, gold["psi"])
    m = 1 / (2 * float(gold["kappa"])) - 4
    mu = float(gold["mu"])
    assert rel_l2(O.dslash(geo, psi, u, m, mu), gold["dpsi"]) <= 1e-15
    assert rel_l2(O.qtm(geo, psi, u, m, mu), gold["qpsi"]) <= 1e-15
    assert rel_l2(O.gauss3d(geo, psi, u, 4.0, 1), gold["gauss_t1"]) <= 1e-15
    assert np.abs(O.ape3d(geo, u, 0.5) - gold["ape"]).max() <= 1e-14
    assert np.abs(O.stout(geo, u, 0.1, 4) - gold["stout4"]).max() <= 1e-14
    assert abs(O.plaquette(geo, u) - float(gold["plaquette"])) <= 1e-15
    assert O.b_minus_Dx(geo, gold["x"], gold["b"], u, float(gold["kappa"]), mu) <= 1e-10


def test_product_generators_reproduce_fixtures(gold):
    import qcd_amd as Q
    L = tuple(int(v) for v in gold["L"])
    assert np.array_equal(Q.gauge_generate(L, float(gold["eps"]), int(gold["seed_u"])), gold["u"])
    assert np.array_equal(Q.spinor_gaussian(L, int(gold["seed_psi"])), gold["psi"])


@pytest.mark.gpu
def test_gpu_reproduces_fixtures(gpu, gold):
    Q = gpu
    L = tuple(int(v) for v in gold["L"])
    kappa, mu = float(gold["kappa"]), float(gold["mu"])
    m = 1 / (2 * kappa) - 4
    lat = Q.Lattice(L)
    g = Q.Gauge(lat).upload(gold["u"])
    x = Q.Spinor(lat).upload(gold["psi"])
    y = Q.Spinor(lat)
    Q.dslash(y, x, g, m, mu, Q.OP_D)
    assert rel_l2(y.download(), gold["dpsi"]) <= 1e-13
    Q.dslash(y, x, g, m, mu, Q.OP_Q)
    assert rel_l2(y.download(), gold["qpsi"]) <= 1e-13
    Q.gauss_smear(x, g, 4.0, 1, 1)
    assert rel_l2(x.download(), gold["gauss_t1"]) <= 1e-13
    out = Q.Gauge(lat)
    Q.ape_smear(out, g, 0.5, 1)
    assert np.abs(out.download() - gold["ape"]).max() <= 1e-13
    Q.stout_smear(out, g, 0.1, 4, 1)
    assert np.abs(out.download() - gold["stout4"]).max() <= 1e-13
    fb = Q.Spinor(lat).upload(gold["b"])
    fx = Q.Spinor(lat)
    S = Q.Solver(lat, g, n_restart=8, tol=1e-10, inner_prec="f16", eo=1)
    st = S.solve(fx, fb, kappa, mu)
    assert st["relres"] <= 1e-10
    assert rel_l2(fx.download(), gold["x"]) <= 1e-7
