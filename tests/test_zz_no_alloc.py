"""The allocation-free residual entry point (qcd_hip_b_minus_Dx_noalloc)
against the allocating one (qcd_hip_b_minus_Dx, b_minus_Dx.c:152-172) and the
oracle: two calls in a row, the second on a smaller lattice, so the static
scratch array is reused at a different size."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_noalloc_residual_matches_the_allocating_one(gpu):
    Q = gpu
    kappa, mu = 0.131, 0.015
    for L, seed in (((32, 16, 16, 16), 801), ((16, 8, 8, 8), 811)):
        lat = Q.Lattice(L)
        g = Q.Gauge(lat)
        Q.gauge_generate_device(g, 0.3, seed)
        u = g.download()
        xh = Q.spinor_gaussian(L, seed + 1)
        bh = Q.spinor_gaussian(L, seed + 2)
        x = Q.Spinor(lat).upload(xh)
        b = Q.Spinor(lat, nrhs=2).upload(bh, rhs=1)
        r = Q.Spinor(lat, nrhs=3)
        ref = Q.b_minus_Dx(x, b, g, kappa, mu, rb=1)
        Q.synchronize()
        held = Q.bytes_in_use()
        got = Q.b_minus_Dx_noalloc(x, b, g, kappa, r, mu, rb=1, rr=2)
        assert Q.bytes_in_use() == held
        print(f"{L}: noalloc {got!r} allocating {ref!r}", flush=True)
        assert got == ref
        res = bh - O.dslash(O.Geometry(L), xh, u, 1 / (2 * kappa) - 4, mu)
        assert rel_l2(r.download(2), res) <= 1e-13
        assert abs(got - np.linalg.norm(res) / np.linalg.norm(bh)) <= 1e-13 * got
        assert O.b_minus_Dx(O.Geometry(L), xh, bh, u, kappa, mu) == pytest.approx(got, rel=1e-13)
