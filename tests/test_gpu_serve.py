"""Serving loop: mixed requests overlap on the GPU and each result equals the
same request served alone; the device field generator follows the host recipe."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_device_generator_matches_host_recipe(gpu):
    Q = gpu
    L = (8, 4, 6, 4)
    lat = Q.Lattice(L)
    g = Q.Gauge(lat)
    Q.gauge_generate_device(g, 0.3, 99)
    dev = g.download()
    host = Q.gauge_generate(L, 0.3, 99)
    assert np.abs(dev - host).max() <= 1e-13
    U = dev.reshape(-1, 3, 3)
    assert np.abs(U @ np.conj(np.transpose(U, (0, 2, 1))) - np.eye(3)).max() <= 1e-14


def _requests():
    reqs = []
    specs = [((16, 8, 8, 8), 0.125, 0.25, 1, "f16"), ((16, 8, 8, 8), 0.128, 0.30, 12, "f32"),
             ((8, 8, 8, 8), 0.12, 0.2, 1, "f64"), ((16, 8, 8, 8), 0.13, 0.35, 12, "f16"),
             ((12, 6, 6, 6), 0.126, 0.28, 1, "f32"), ((16, 8, 8, 8), 0.122, 0.22, 1, "f16")]
    for k, (L, kappa, eps, nrhs, prec) in enumerate(specs):
        reqs.append({"arrival_s": 0.002 * k, "L": L, "kappa": kappa, "eps": eps, "seed": 7000 + k,
                     "nrhs": nrhs, "inner_prec": prec, "n_restart": 16, "tol": 1e-10})
    return reqs


def test_serve_mix_equals_alone_and_converges(gpu):
    Q = gpu
    reqs = _requests()
    mix, wall, sols = Q.serve(reqs, workers=4, realtime=True, want_solutions=True)
    assert all(r["rc"] == 0 and r["relres"] <= 1e-10 for r in mix)
    assert all(r["latency_s"] >= r["service_s"] > 0 for r in mix)
    for k, q in enumerate(reqs):
        alone, _, s1 = Q.serve([dict(q, arrival_s=0.0)], workers=1, realtime=False, want_solutions=True)
        assert alone[0]["iters"] == mix[k]["iters"]
        assert np.array_equal(s1[0], sols[k])
    # the FP64 residual of each served solution, recomputed by the CPU oracle
    for k, q in enumerate(reqs):
        L = q["L"]
        lat = Q.Lattice(L)
        g = Q.Gauge(lat)
        Q.gauge_generate_device(g, q["eps"], q["seed"])
        u = g.download()
        geo = O.Geometry(L)
        for r in range(q["nrhs"]):
            b = Q.point_source(L, (0, 0, 0, 0), r // 3, r % 3)
            assert O.b_minus_Dx(geo, sols[k][r], b, u, q["kappa"]) <= 1e-10


def test_serve_admits_by_memory_budget(gpu):
    """A budget that fits one request at a time serialises the stream."""
    Q = gpu
    reqs = _requests()[:3]
    res, _, _ = Q.serve(reqs, workers=3, realtime=False)
    need = max(r["bytes"] for r in res)
    tight, _, _ = Q.serve(reqs, workers=3, budget_bytes=int(need * 1.01), realtime=False)
    spans = sorted((r["queued_s"], r["queued_s"] + r["service_s"]) for r in tight)
    for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
        assert s1 >= e0 - 1e-3   # never two at once
    with pytest.raises(Q.QcdError):
        Q.serve(reqs, workers=1, budget_bytes=int(need * 0.5), realtime=False)
