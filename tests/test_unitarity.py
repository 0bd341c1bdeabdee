"""Unitarity of the gauge fields the library makes, in the reference's own
check: ||U^+ U v - v|| for random colour vectors v (tests/test_smearing.c:72-80),
together with U^+ U = 1 and det U = 1 link by link.

Fields: the host generator (CPU), and on the GPU the device generator, APE
smearing of the spatial links (qcd_smearing.c:205-273, projected back to SU(3))
and 3d / 4d stout smearing (qcd_smearing.c:374-476, U' = exp(iQ) U)."""
import numpy as np
import pytest


def unitarity_errors(u, seed=7):
    """max |U^+ U - 1|, max |det U - 1| and max ||U^+ U v - v|| / ||v|| over links."""
    links = u.reshape(-1, 3, 3)
    uhu = np.einsum("lji,ljk->lik", links.conj(), links)
    e_uu = np.abs(uhu - np.eye(3)[None]).max()
    e_det = np.abs(np.linalg.det(links) - 1.0).max()
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(links.shape[0], 3)) + 1j * rng.normal(size=(links.shape[0], 3))
    w = np.einsum("lij,lj->li", uhu, v)
    e_v = (np.linalg.norm(w - v, axis=1) / np.linalg.norm(v, axis=1)).max()
    return e_uu, e_det, e_v


def test_host_generator_is_su3():
    import qcd_amd as Q
    for L, eps, seed in (((4, 4, 4, 4), 0.2, 11), ((6, 4, 2, 8), 0.35, 12)):
        u = Q.gauge_generate(L, eps, seed)
        e_uu, e_det, e_v = unitarity_errors(u)
        assert e_uu <= 1e-14 and e_det <= 1e-14 and e_v <= 1e-14, (L, e_uu, e_det, e_v)


@pytest.mark.gpu
def test_device_generator_and_smeared_fields_are_su3(gpu):
    Q = gpu
    L = (8, 6, 4, 6)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat)
    Q.gauge_generate_device(g, 0.3, 21)
    u0 = g.download()
    out = Q.Gauge(lat)
    fields = {"device generator": u0}
    Q.ape_smear(out, g, 0.5, 10)
    fields["APE, 10 steps"] = out.download()
    for nd in (3, 4):
        Q.stout_smear(out, g, 0.1, nd, 4)
        fields[f"stout {nd}d, 4 steps"] = out.download()
    for name, u in fields.items():
        e_uu, e_det, e_v = unitarity_errors(u)
        assert e_uu <= 1e-13 and e_det <= 1e-13 and e_v <= 1e-13, (name, e_uu, e_det, e_v)
    # the smearing changed the spatial links and, for APE, kept the temporal ones
    ape = fields["APE, 10 steps"]
    assert np.array_equal(ape[:, 0], u0[:, 0])
    assert np.abs(ape[:, 1:] - u0[:, 1:]).max() > 1e-3
