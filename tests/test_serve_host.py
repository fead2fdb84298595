"""Host logic of the serving loop and the batched / serving workload generators
(CPU only: the scheduler runs against stand-in Scene and Batch objects)."""
import math

import numpy as np
import pytest

from mphimplicit_amd import scenes, serve


class FakeScene:
    live = 0

    def __init__(self, params, grid, device=0):
        self.params, self.grid = params, grid
        self.time = grid.time
        FakeScene.live += 1

    def scalar(self, name):
        assert name == "Time"
        return self.time

    def device_bytes(self):
        return 3000 * self.grid.np

    def read(self, name):
        return np.array([self.time])

    def close(self):
        FakeScene.live -= 1


class FakeStats:
    def __init__(self, done, status=0):
        self.steps_done = done
        self.status = status


class FakeBatch:
    def __init__(self, device=0):
        self.members = []
        self.max_members = 0

    def add(self, s):
        self.members.append(s)
        self.max_members = max(self.max_members, len(self.members))

    def remove(self, s):
        self.members.remove(s)

    def step(self):
        out = []
        for s in self.members:
            p = s.params
            if getattr(p, "diverges", False):  # its solve fails: not advanced
                out.append(FakeStats(0, -5))
            elif s.time < p.end_time + 1e-5 * p.dt:
                s.time += p.dt
                out.append(FakeStats(1))
            else:
                out.append(FakeStats(0))
        return out

    def close(self):
        self.members = []


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(serve, "Scene", FakeScene)
    monkeypatch.setattr(serve, "Batch", FakeBatch)
    FakeScene.live = 0
    yield


def small_requests(n=8, arrival_gap=0.0):
    out = []
    for k in range(n):
        p, g = scenes.box_scene(100 + k, 2 if k % 2 else 3, 200 + 50 * k, 0.1, (0, -9.8, 0) if k % 2 else (0, 0, -9.8),
                                steps=1 + k % 4)
        out.append(serve.Request(rid=k, arrival=k * arrival_gap, steps=1 + k % 4, params=p, grid=g,
                                 nf=int((g.type < 2).sum()), dim=p.dim))
    return out


def test_all_requests_served_once(fake):
    reqs = small_requests()
    srv = serve.Server(keep_fields=("Velocity",), build_ahead=False)
    res = srv.run(reqs, duration_s=30)
    assert res["requests_done"] == len(reqs)
    assert sorted(r.rid for r in srv.finished) == list(range(len(reqs)))
    for r in srv.finished:
        assert r.steps_done == r.steps and r.done >= r.admitted >= 0 and r.latency >= 0
    assert FakeScene.live == 0                      # every scene was released
    assert srv.batch.max_members == len(reqs)       # all arrived at once: all in one batch
    assert res["batch_steps"] == max(r.steps for r in reqs)


def test_budget_limits_concurrency(fake):
    """After every admission the admitted footprint fits the budget (or one scene
    runs alone), and the budget does bind: not all requests run together."""
    reqs = small_requests()
    est = serve.FootprintModel().estimate
    one = max(est(r.dim, r.nf, r.grid.np, True) for r in reqs)
    budget = int(1.5 * one)
    srv = serve.Server(budget_bytes=budget, build_ahead=False)
    seen = []
    admit = srv._admit

    def checked(now):
        admit(now)
        n = len(srv.active) + len(srv.building)
        seen.append((n, srv.used_bytes()))
        assert srv.used_bytes() <= budget or n == 1

    srv._admit = checked
    res = srv.run(reqs, duration_s=30)
    assert res["requests_done"] == len(reqs)
    assert 1 <= srv.batch.max_members < len(reqs)
    assert max(n for n, _ in seen) == srv.batch.max_members


def test_tiny_budget_serialises(fake):
    reqs = small_requests(5)
    srv = serve.Server(budget_bytes=1, build_ahead=False)
    res = srv.run(reqs, duration_s=30)
    assert res["requests_done"] == 5
    assert srv.batch.max_members == 1
    assert res["batch_steps"] == sum(r.steps for r in reqs)


def test_footprint_model_learns_upwards():
    m = serve.FootprintModel()
    e = m.estimate(3, 1000, 3000, True)
    m.observe(3, e, 2 * e)
    assert m.estimate(3, 1000, 3000, True) == pytest.approx(2 * e, rel=1e-6)
    m.observe(3, 2 * e, e)  # never shrinks
    assert m.estimate(3, 1000, 3000, True) == pytest.approx(2 * e, rel=1e-6)


def test_batch_sweep_distribution():
    sw = scenes.batch_sweep(16, steps=3)
    assert len(sw) == 16
    dims = [p.dim for p, _ in sw]
    assert set(dims) <= {2, 3}
    for p, g in sw:
        nf = int((g.type < 2).sum())
        assert 1500 <= nf <= 230000
        assert 0.01 <= p.shear_viscosity[0] <= 10.0
        gn = math.sqrt(sum(p.gravity[d] ** 2 for d in range(3)))
        assert gn == pytest.approx(9.8)
        if p.dim == 2:
            assert p.gravity[2] == 0.0
        assert p.end_time == pytest.approx(2 * p.dt)
        # particles inside the domain, walls around the fluid
        lo, hi = np.asarray(g.domain_min), np.asarray(g.domain_max)
        assert (g.pos >= lo).all() and (g.pos <= hi).all()
    assert scenes.batch_sweep(4, steps=3)[2][1].pos.tobytes() == sw[2][1].pos.tobytes()  # seeded


def test_serve_stream_statistics():
    st = scenes.serve_stream(4000, rate=200.0)
    t = np.array([x[0] for x in st])
    assert (np.diff(t) > 0).all()
    assert 4000 / t[-1] == pytest.approx(200.0, rel=0.08)          # Poisson rate
    steps = np.array([x[1] for x in st])
    assert steps.min() == 1 and steps.max() == 20
    nf = np.array([x[4] for x in st])
    assert nf.min() >= 5000 and nf.max() <= 500000
    assert np.median(nf) == pytest.approx(50000, rel=0.1)          # lognormal median
    mu = np.array([x[5] for x in st])
    assert mu.min() >= 0.01 and mu.max() <= 10.0
    assert 0.4 < np.mean([x[3] == 2 for x in st]) < 0.6


@pytest.mark.parametrize("dim,nf", [(2, 300), (2, 5000), (3, 700), (3, 20000)])
def test_box_scene_counts(dim, nf):
    """The counts the footprint estimate uses before a request's scene is built."""
    p, g = scenes.box_scene(1, dim, nf, 0.1, (0.0, -9.8, 0.0) if dim == 2 else (0.0, 0.0, -9.8))
    assert scenes.box_scene_counts(dim, nf) == (int((g.type < 2).sum()), g.np)


def test_lazy_requests_built_at_admission(fake):
    reqs = serve.make_requests(scenes.serve_stream(6, rate=1000.0))
    assert all(r.params is None and r.grid is None and r.np > r.nf > 0 for r in reqs)
    srv = serve.Server()
    res = srv.run(reqs, duration_s=30)
    assert res["requests_done"] == 6
    assert all(r.params is None for r in srv.finished)  # released again after retiring


def test_build_ahead_serves_all(fake):
    """Scenes built on the worker thread join the batch as they become ready:
    every request is served exactly once and every scene is released."""
    reqs = small_requests(12, arrival_gap=0.001)
    srv = serve.Server(keep_fields=("Velocity",))
    res = srv.run(reqs, duration_s=30)
    srv.close()
    assert res["requests_done"] == 12 and res["requests_failed"] == 0
    assert sorted(r.rid for r in srv.finished) == list(range(12))
    for r in srv.finished:
        assert r.steps_done == r.steps and r.done >= r.built >= r.admitted >= 0
    assert FakeScene.live == 0


@pytest.mark.parametrize("ahead", [False, True])
def test_failing_requests_fail_alone(fake, ahead):
    """A request whose build raises and one whose solve fails (the batch reports a
    member status) are retired as failed; the others are served as usual."""
    reqs = small_requests(6)

    def broken():
        raise ValueError("bad scene")

    reqs[2].params = reqs[2].grid = None
    reqs[2].make = broken
    reqs[4].params.diverges = True
    srv = serve.Server(build_ahead=ahead)
    res = srv.run(reqs, duration_s=30)
    srv.close()
    assert res["requests_failed"] == 2 and res["requests_done"] == 4
    failed = {r.rid: r.error for r in srv.failed}
    assert set(failed) == {2, 4}
    assert "bad scene" in failed[2] and "non-finite" in failed[4]
    assert sorted(r.rid for r in srv.finished) == [0, 1, 3, 5]
    assert FakeScene.live == 0
