"""Serving loop on the GPU: continuously batched requests give the same result as
each request alone (bit for bit), under a roomy and under a tight HBM budget."""
import numpy as np
import pytest

from mphimplicit_amd import abi, scenes
from mphimplicit_amd.scene import Scene
from mphimplicit_amd.serve import Request, Server

pytestmark = pytest.mark.gpu

KEEP = ("Position", "Velocity", "PressureP")


def requests(precision):
    spec = [  # (arrival s, steps, seed, dim, nf, mu)
        (0.00, 3, 11, 2, 400, 0.1),
        (0.00, 1, 12, 3, 700, 1.0),
        (0.00, 4, 13, 3, 1500, 0.05),
        (0.01, 2, 14, 2, 2500, 3.0),
        (0.02, 5, 15, 3, 300, 0.5),
        (0.02, 2, 16, 2, 900, 0.02),
    ]
    out = []
    for k, (t, steps, seed, dim, nf, mu) in enumerate(spec):
        g_dir = (1.0, -9.8, 0.0) if dim == 2 else (0.5, 0.0, -9.8)
        p, g = scenes.box_scene(seed, dim, nf, mu, g_dir, steps=steps)
        p.precision = precision
        p.cr_tolerance = 1e-14
        out.append(Request(rid=k, arrival=t, steps=steps, params=p, grid=g, nf=int((g.type < 2).sum()), dim=dim))
    return out


def solo(r):
    s = Scene(r.params, r.grid)
    for _ in range(r.steps):
        s.step(1)
    out = {n: s.read(n) for n in KEEP}
    s.close()
    return out


@pytest.mark.parametrize("precision,budget", [(abi.PRECISION_FP64, 64 << 30), (abi.PRECISION_MIXED, 64 << 30),
                                              (abi.PRECISION_FP64, 1)])
def test_served_equals_alone(precision, budget):
    reqs = requests(precision)
    print("[serve] precision %d budget %d: %d requests" % (precision, budget, len(reqs)), flush=True)
    srv = Server(budget_bytes=budget, keep_fields=KEEP)
    try:
        res = srv.run(reqs, duration_s=120.0)
    finally:
        srv.close()
    assert res["requests_done"] == len(reqs), res
    if budget == 1:  # one request at a time: every batch step had one member
        assert res["batch_steps"] == sum(r.steps for r in reqs)
    else:            # overlapping requests shared batch steps
        assert res["batch_steps"] < sum(r.steps for r in reqs)
    for r in srv.finished:
        assert r.steps_done == r.steps
        ref = solo(r)
        for n in KEEP:
            assert np.array_equal(r.result[n].view(np.uint8), ref[n].view(np.uint8)), (r.rid, n)
        assert r.latency >= 0
