"""Serving loop: scene-step requests over the batched executor (continuous batching).

A request is a scene (parameters + particles) and a number of time steps.  The
server keeps every admitted scene resident in HBM inside one ``mph_batch``; each
iteration of the loop advances all admitted scenes by one time step together
(lockstep CR solves, one launch per kernel for the whole batch).  Requests that
arrive while others are on the GPU wait in arrival order and are admitted between
batch steps as soon as a slot is free and their estimated HBM footprint fits the
budget; finished requests leave the batch and free their memory at once.  A
scene's arithmetic does not depend on its batch-mates, so each result is the same
as running that request alone.

The reference has no serving layer (one scene per process, source/main.cpp:313-521);
this module is the host-side scheduler around the C ABI.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .abi import OPERATOR_ASSEMBLED, PRECISION_MIXED, Params
from .scene import Batch, Scene


@dataclass
class Request:
    rid: int
    arrival: float          # seconds since the start of the run
    steps: int
    params: Params | None   # None until built by `make` (at admission)
    grid: object            # files.GridData, or None until built
    nf: int = 0
    dim: int = 3
    np: int = 0             # total particles (for the footprint estimate before building)
    make: object = None     # () -> (params, grid), called at admission when params is None
    # filled in by the server
    admitted: float = -1.0
    done: float = -1.0
    steps_done: int = 0
    scene: Scene | None = None
    est_bytes: int = 0
    result: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class FootprintModel:
    """HBM bytes of a scene before it exists: per-fluid-particle and per-particle
    terms by dimension (particle matrix and its FP32 copy, neighbour lists, state),
    corrected online by the largest measured/estimated ratio seen so far."""

    PER_FLUID = {2: 6.0e3, 3: 28.0e3}
    PER_PARTICLE = 2.0e3

    def __init__(self):
        self.scale = {2: 1.0, 3: 1.0}

    def estimate(self, dim: int, nf: int, np_: int, mixed: bool) -> int:
        per_fluid = self.PER_FLUID[dim] * (1.0 if mixed else 0.68)
        return int(self.scale[dim] * (per_fluid * nf + self.PER_PARTICLE * np_))

    def observe(self, dim: int, est: int, actual: int) -> None:
        if est > 0:
            self.scale[dim] = max(self.scale[dim], self.scale[dim] * actual / est)


class Server:
    """Continuous-batching scheduler of scene-step requests on one device."""

    def __init__(self, device: int = 0, budget_bytes: int = 200 << 30, max_active: int = 256,
                 keep_fields: tuple = (), lazy_mixed: bool = True):
        self.device = device
        self.budget = int(budget_bytes)
        self.max_active = max_active
        self.keep_fields = keep_fields
        self.lazy_mixed = lazy_mixed  # precision assumed for requests not yet built
        self.batch = Batch(device)
        self.model = FootprintModel()
        self.pending: list[Request] = []
        self.active: list[Request] = []
        self.finished: list[Request] = []
        self.particle_steps = 0
        self.fluid_steps = 0
        self.peak_bytes = 0
        self.batch_steps = 0

    # -- accounting ---------------------------------------------------------
    def used_bytes(self) -> int:
        return sum(r.est_bytes for r in self.active)

    def _admit(self, now: float) -> None:
        while self.pending and len(self.active) < self.max_active:
            r = self.pending[0]
            np_ = r.grid.np if r.grid is not None else r.np
            mixed = r.params.precision == PRECISION_MIXED if r.params is not None else self.lazy_mixed
            est = self.model.estimate(r.dim, r.nf, np_, mixed)
            if self.active and self.used_bytes() + est > self.budget:
                break  # wait for memory (an empty server admits anything)
            self.pending.pop(0)
            if r.params is None:
                r.params, r.grid = r.make()
            r.scene = Scene(r.params, r.grid, self.device)
            r.est_bytes = est
            r.admitted = now
            self.batch.add(r.scene)
            self.active.append(r)

    def _retire(self, r: Request, now: float) -> None:
        self.batch.remove(r.scene)
        for name in self.keep_fields:
            r.result[name] = r.scene.read(name)
        r.scene.close()
        r.scene = None
        if r.make is not None:  # rebuilt on demand; keep finished requests small
            r.params = r.grid = None
        r.done = now
        self.active.remove(r)
        self.finished.append(r)

    def step(self, clock) -> None:
        """One time step of every admitted request."""
        stats = self.batch.step()
        self.batch_steps += 1
        now = clock()
        for r, st in zip(list(self.active), stats):
            r.steps_done += st.steps_done
            self.particle_steps += st.steps_done * r.grid.np
            self.fluid_steps += st.steps_done * r.nf
            actual = r.scene.device_bytes()
            self.model.observe(r.dim, r.est_bytes, actual)
            r.est_bytes = max(r.est_bytes, actual)
        self.peak_bytes = max(self.peak_bytes, self.used_bytes())
        for r in list(self.active):
            if r.steps_done >= r.steps or r.scene.scalar("Time") >= r.params.end_time + 1.0e-5 * r.params.dt:
                self._retire(r, now)

    # -- the loop -----------------------------------------------------------
    def run(self, requests: list[Request], duration_s: float = 60.0, log=None) -> dict:
        """Serve `requests` (sorted by arrival) against the wall clock: a request
        becomes pending at its arrival time; the run stops after duration_s or
        when every request has finished."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0  # noqa: E731
        nxt = 0
        while True:
            now = clock()
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                self.pending.append(reqs[nxt])
                nxt += 1
            self._admit(now)
            if self.active:
                self.step(clock)
                if log and self.batch_steps % 10 == 0:
                    log("[serve] t=%.1fs active %d pending %d done %d" %
                        (clock(), len(self.active), len(self.pending), len(self.finished)))
            elif nxt < len(reqs):
                time.sleep(max(0.0, min(0.05, reqs[nxt].arrival - clock())))
            else:
                break
            if clock() > duration_s:
                break
        wall = clock()
        lat = np.array([r.latency for r in self.finished]) if self.finished else np.zeros(1)
        return {
            "wall_s": wall,
            "requests_arrived": nxt,
            "requests_done": len(self.finished),
            "requests_active_at_end": len(self.active),
            "requests_pending_at_end": len(self.pending),
            "fluid_particle_steps_per_s": self.fluid_steps / wall if wall > 0 else 0.0,
            "particle_steps_per_s": self.particle_steps / wall if wall > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "batch_steps": self.batch_steps,
            "peak_bytes": self.peak_bytes,
        }

    def close(self) -> None:
        for r in list(self.active):
            self.batch.remove(r.scene)
            r.scene.close()
        self.active = []
        self.batch.close()


def make_requests(stream, precision=PRECISION_MIXED, tol=1e-16, operator_form=OPERATOR_ASSEMBLED) -> list[Request]:
    """Requests from scenes.serve_stream tuples (arrival, steps, seed, dim, nf, mu);
    each scene is built when the server admits its request."""
    from . import scenes

    def factory(seed, dim, nf, mu, steps):
        def make():
            g_dir = (0.0, -9.8, 0.0) if dim == 2 else (0.0, 0.0, -9.8)
            p, g = scenes.box_scene(seed, dim, nf, mu, g_dir, steps=steps)
            p.precision = precision
            p.cr_tolerance = tol
            p.operator_form = operator_form
            return p, g
        return make

    out = []
    for k, (t, steps, seed, dim, nf, mu) in enumerate(stream):
        nfl, npart = scenes.box_scene_counts(dim, nf)
        out.append(Request(rid=k, arrival=t, steps=steps, params=None, grid=None, nf=nfl, dim=dim, np=npart,
                           make=factory(seed, dim, nf, mu, steps)))
    return out
