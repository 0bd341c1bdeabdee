"""The matrix-core (MFMA) multi-rhs hops (qcd_amd/csrc/mma.hip) against the
oracle, rhs by rhs (lib/src/qcd_wilson.c:114-336 applied to each column).

FP16 storage with per-site norms and FP16 links: every rhs within 4e-3
relative L2 of the FP64 oracle, like the vector-ALU FP16 kernels.  A lane's
result depends only on its own (site, rhs), so a rhs gets the same bits alone
and inside any batch."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from test_gpu_eo import schur_reference

pytestmark = pytest.mark.gpu

MMA = 6   # traversal id of the matrix-core Dslash


def _fields(Q, L, nrhs, seed, theta=(np.pi, 0.0, 0.0, 0.0)):
    lat = Q.Lattice(L, theta)
    u = Q.gauge_generate(L, 0.3, seed)
    g64 = Q.Gauge(lat).upload(u)
    g = Q.Gauge(lat, Q.F16)
    Q.convert(g, g64)
    srcs = [Q.spinor_gaussian(L, seed + 1 + r) for r in range(nrhs)]
    x = Q.Spinor(lat, Q.F16, nrhs=nrhs)
    for r, s in enumerate(srcs):
        x.upload(s, rhs=r)
    return lat, u, g, srcs, x


@pytest.mark.parametrize("L,theta", [((16, 8, 8, 8), (np.pi, 0.0, 0.0, 0.0)),
                                     ((6, 2, 2, 5), (np.pi, 0.4, -0.3, 1.1)),
                                     ((5, 4, 2, 6), (np.pi, 0.2, 0.0, -0.7))])
@pytest.mark.parametrize("nrhs", [1, 3, 5, 16])
def test_mma_dslash_per_rhs(gpu, L, theta, nrhs):
    """D and Q with twisted mass over 1, 3, 5 and 16 rhs (ragged groups of 4),
    on lattices whose volumes (120, 240) are not multiples of the 16-site block,
    with phases in every direction and an odd time extent (hop links in site
    order rather than parity-major); each rhs against the oracle and against
    the vector-ALU two-rhs kernel."""
    Q = gpu
    m, tm = 1 / (2 * 0.13) - 4, 0.02
    lat, u, g, srcs, x = _fields(Q, L, nrhs, 700 + nrhs, theta)
    y = Q.Spinor(lat, Q.F16, nrhs=nrhs)
    geo = O.Geometry(L, theta)
    for op, ref in ((Q.OP_D, O.dslash), (Q.OP_Q, O.qtm)):
        Q.dslash_set_config(x, g, MMA, 0)
        Q.dslash(y, x, g, m, tm, op)
        mma = [y.download(r) for r in range(nrhs)]
        Q.dslash_set_config(x, g, 5 | (3 << 8), 16)
        Q.dslash(y, x, g, m, tm, op)
        for r in range(nrhs):
            want = ref(geo, srcs[r], u, m, tm)
            assert rel_l2(mma[r], want) <= 4e-3, (op, r)
            assert rel_l2(mma[r], y.download(r)) <= 4e-3, (op, r)


def test_mma_rhs_bits_do_not_depend_on_the_batch(gpu):
    """rhs 2 of a 7-rhs batch equals the same vector applied alone."""
    Q = gpu
    L = (16, 8, 8, 8)
    m = 1 / (2 * 0.13) - 4
    lat, u, g, srcs, x = _fields(Q, L, 7, 760)
    y = Q.Spinor(lat, Q.F16, nrhs=7)
    Q.dslash_set_config(x, g, MMA, 0)
    Q.dslash(y, x, g, m, 0.01, Q.OP_Q)
    x1 = Q.Spinor(lat, Q.F16).upload(srcs[2])
    y1 = Q.Spinor(lat, Q.F16)
    Q.dslash(y1, x1, g, m, 0.01, Q.OP_Q)
    assert np.array_equal(y.download(2), y1.download())


@pytest.mark.parametrize("L,nrhs", [((32, 16, 16, 16), 48), ((48, 24, 24, 24), 16)])
def test_mma_dslash_config_sizes(gpu, L, nrhs):
    """16^3x32 over 48 rhs and 24^3x48 over 16 rhs (the batched configurations),
    every rhs against the oracle."""
    Q = gpu
    m = 1 / (2 * 0.132) - 4
    print(f"MMA Dslash {L} x {nrhs}", flush=True)
    lat = Q.Lattice(L)
    g64 = Q.Gauge(lat)
    Q.gauge_generate_device(g64, 0.3, 7100)
    u = g64.download()
    g = Q.Gauge(lat, Q.F16)
    Q.convert(g, g64)
    x = Q.Spinor(lat, Q.F16, nrhs=nrhs)
    seeds = [7200 + r for r in range(nrhs)]
    for r, sd in enumerate(seeds):
        x.upload(Q.spinor_gaussian(L, sd), rhs=r)
    y = Q.Spinor(lat, Q.F16, nrhs=nrhs)
    Q.dslash_set_config(x, g, MMA, 0)
    Q.dslash(y, x, g, m, 0.0, Q.OP_D)
    geo = O.Geometry(L)
    for r, sd in enumerate(seeds):
        assert rel_l2(y.download(r), O.dslash(geo, Q.spinor_gaussian(L, sd), u, m, 0.0)) <= 4e-3, r
    Q.dslash_tune(y, x, g)   # leave a tuned configuration behind


@pytest.mark.parametrize("nrhs", [16, 21])
def test_mma_even_odd_schur(gpu, nrhs):
    """The FP16 Schur operator M - H_oe M^-1 H_eo of a batch of 16 or more rhs
    (both hops on the matrix cores; 21: a ragged last group of 4) rhs by rhs
    against the oracle's full operator, and the bits of each rhs equal to those
    of the same rhs in another matrix-core batch."""
    Q = gpu
    L, theta = (8, 4, 6, 4), (np.pi, 0.2, 0.0, -0.6)
    m, tm = 1 / (2 * 0.13) - 4, 0.05
    u = Q.gauge_generate(L, 0.3, 780)
    lat = Q.Lattice(L, theta)
    Lh = (L[0] // 2,) + tuple(L[1:])
    half = Q.Lattice(Lh)
    g64 = Q.Gauge(lat).upload(u)
    g = Q.Gauge(lat, Q.F16)
    Q.convert(g, g64)
    vs = [Q.eo_split(Q.spinor_gaussian(L, 790 + r), L)[1] for r in range(nrhs)]

    def schur(vecs):
        xb = Q.Spinor(half, Q.F16, nrhs=len(vecs))
        for r, v in enumerate(vecs):
            xb.upload(v, rhs=r)
        yb = Q.Spinor(half, Q.F16, nrhs=len(vecs))
        Q.dslash_schur(yb, xb, g, m, tm)
        return [yb.download(r) for r in range(len(vecs))]

    ys = schur(vs)
    for r, v in enumerate(vs):
        assert rel_l2(ys[r], schur_reference(Q, L, u, v, m, tm, theta)) <= 4e-3, r
    # the same vectors in reverse order, in a batch of 17 (rows move between
    # groups of 4 and between lanes)
    rev = schur(vs[::-1][:17])
    for k, y in enumerate(rev):
        assert np.array_equal(y, ys[nrhs - 1 - k]), k


def test_mma_z_chunks_of_unequal_length(gpu):
    """The matrix-core kernels walk z in chunks of 12 planes: z extents of 13
    (full lattice, D over 5 rhs) and 14 (even-odd Schur over 16 rhs) leave a
    last chunk of 1 and 2 planes.  Every rhs against the oracle."""
    Q = gpu
    m, tm = 1 / (2 * 0.13) - 4, 0.03
    L = (8, 4, 4, 13)
    lat, u, g, srcs, x = _fields(Q, L, 5, 830, (np.pi, 0.3, 0.0, -0.4))
    y = Q.Spinor(lat, Q.F16, nrhs=5)
    Q.dslash_set_config(x, g, MMA, 0)
    Q.dslash(y, x, g, m, tm, Q.OP_D)
    geo = O.Geometry(L, (np.pi, 0.3, 0.0, -0.4))
    for r in range(5):
        assert rel_l2(y.download(r), O.dslash(geo, srcs[r], u, m, tm)) <= 4e-3, r
    L, theta = (8, 4, 4, 14), (np.pi, 0.0, 0.2, 0.0)
    u = Q.gauge_generate(L, 0.3, 840)
    lat = Q.Lattice(L, theta)
    half = Q.Lattice((L[0] // 2,) + tuple(L[1:]))
    g64 = Q.Gauge(lat).upload(u)
    g = Q.Gauge(lat, Q.F16)
    Q.convert(g, g64)
    vs = [Q.eo_split(Q.spinor_gaussian(L, 850 + r), L)[1] for r in range(16)]
    xb = Q.Spinor(half, Q.F16, nrhs=16)
    for r, v in enumerate(vs):
        xb.upload(v, rhs=r)
    yb = Q.Spinor(half, Q.F16, nrhs=16)
    Q.dslash_schur(yb, xb, g, m, tm)
    for r, v in enumerate(vs):
        assert rel_l2(yb.download(r), schur_reference(Q, L, u, v, m, tm, theta)) <= 4e-3, r
