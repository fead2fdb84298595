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

Admitted requests are built on a worker thread (the scene's particles in Python,
then mph_scene_create: upload, first cell sort, neighbour search) while the batch
keeps stepping; a scene joins the batch at the first step after it is ready.  A
request whose build fails, or whose solve fails on its own (a non-finite residual:
the batch reports it per member), is retired as failed with its error and does not
disturb the others.

The reference has no serving layer (one scene per process, source/main.cpp:313-521);
this module is the host-side scheduler around the C ABI.
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from .abi import OPERATOR_ASSEMBLED, PRECISION_FP64, PRECISION_MIXED, Params
from .scene import Batch, MphError, Scene


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
    error: str = ""         # why the request failed ("" if it did not)
    built: float = -1.0     # when its scene was ready to join the batch

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
                 keep_fields: tuple = (), lazy_mixed: bool = False, build_ahead: bool = True):
        self.device = device
        self.budget = int(budget_bytes)
        self.max_active = max_active
        self.keep_fields = keep_fields
        self.lazy_mixed = lazy_mixed  # precision assumed for requests not yet built
        self.build_ahead = build_ahead  # build admitted scenes on the worker thread
        self.batch = Batch(device)
        self.model = FootprintModel()
        self.pending: list[Request] = []
        self.active: list[Request] = []
        self.finished: list[Request] = []
        self.building: list[Request] = []  # admitted, scene being built on the worker
        self.failed: list[Request] = []
        self.particle_steps = 0
        self.fluid_steps = 0
        self.peak_bytes = 0
        self.batch_steps = 0
        self._todo: queue.Queue = queue.Queue()
        self._ready: queue.Queue = queue.Queue()
        self._worker = threading.Thread(target=self._build_loop, name="mph-scene-builder", daemon=True)
        self._worker.start()

    # -- accounting ---------------------------------------------------------
    def used_bytes(self) -> int:
        return sum(r.est_bytes for r in self.active) + sum(r.est_bytes for r in self.building)

    def _build_loop(self) -> None:
        """Worker thread: build admitted requests' scenes in admission order."""
        while True:
            r = self._todo.get()
            if r is None:
                return
            self._build(r)
            self._ready.put(r)

    def _build(self, r: Request) -> None:
        try:
            if r.params is None:
                r.params, r.grid = r.make()
            r.scene = Scene(r.params, r.grid, self.device)
        except Exception as e:  # a bad request fails alone
            r.error = "build: %s" % e
            r.scene = None

    def _admit(self, now: float) -> None:
        while self.pending and len(self.active) + len(self.building) < self.max_active:
            r = self.pending[0]
            np_ = r.grid.np if r.grid is not None else r.np
            mixed = r.params.precision == PRECISION_MIXED if r.params is not None else self.lazy_mixed
            est = self.model.estimate(r.dim, r.nf, np_, mixed)
            if (self.active or self.building) and self.used_bytes() + est > self.budget:
                break  # wait for memory (an empty server admits anything)
            self.pending.pop(0)
            r.est_bytes = est
            r.admitted = now
            self.building.append(r)
            if self.build_ahead:
                self._todo.put(r)
            else:  # build here; the scene joins before the next step
                self._build(r)
                self._ready.put(r)

    def _join_ready(self, now: float, wait: bool) -> None:
        """Move built scenes into the batch (waiting for one if `wait`)."""
        while self.building:
            try:
                r = self._ready.get(block=wait, timeout=None)
            except queue.Empty:
                return
            wait = False
            self.building.remove(r)
            r.built = now
            if r.scene is None:
                self._fail(r, now, r.error)
                continue
            self.batch.add(r.scene)
            self.active.append(r)

    def _fail(self, r: Request, now: float, why: str) -> None:
        r.error = why or "failed"
        if r.scene is not None:
            if r in self.active:
                self.batch.remove(r.scene)
                self.active.remove(r)
            r.scene.close()
            r.scene = None
        if r.make is not None:
            r.params = r.grid = None
        r.done = now
        self.failed.append(r)

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
            if st.status != 0:
                self._fail(r, now, "solve: %s" % MphError(st.status, "non-finite CR residual"))
                continue
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
            # join what the worker has built; with nothing to step, wait for it
            self._join_ready(now, wait=not self.active and bool(self.building))
            if self.active:
                self.step(clock)
                if log and self.batch_steps % 10 == 0:
                    log("[serve] t=%.1fs active %d building %d pending %d done %d failed %d" %
                        (clock(), len(self.active), len(self.building), len(self.pending), len(self.finished),
                         len(self.failed)))
            elif self.building:
                continue
            elif nxt < len(reqs):
                time.sleep(max(0.0, min(0.05, reqs[nxt].arrival - clock())))
            else:
                break
            if clock() > duration_s:
                break
        wall = clock()
        lat = np.array([r.latency for r in self.finished]) if self.finished else np.zeros(1)
        # offered load: the fluid-particle-steps of the requests that arrived in the run
        offered = sum(r.nf * r.steps for r in reqs[:nxt])
        errors: dict[str, int] = {}
        for r in self.failed:
            errors[r.error] = errors.get(r.error, 0) + 1
        return {
            "wall_s": wall,
            "requests_arrived": nxt,
            "requests_done": len(self.finished),
            "requests_failed": len(self.failed),
            "failures": errors,
            "requests_active_at_end": len(self.active) + len(self.building),
            "requests_pending_at_end": len(self.pending),
            "offered_fluid_particle_steps_per_s": offered / wall if wall > 0 else 0.0,
            "fluid_particle_steps_per_s": self.fluid_steps / wall if wall > 0 else 0.0,
            "particle_steps_per_s": self.particle_steps / wall if wall > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "batch_steps": self.batch_steps,
            "peak_bytes": self.peak_bytes,
        }

    def close(self) -> None:
        self._todo.put(None)
        self._worker.join()
        while self.building:  # built but never joined
            r = self._ready.get()
            self.building.remove(r)
            if r.scene is not None:
                r.scene.close()
        for r in list(self.active):
            self.batch.remove(r.scene)
            r.scene.close()
        self.active = []
        self.batch.close()


def make_requests(stream, precision=PRECISION_FP64, tol=1e-16, operator_form=OPERATOR_ASSEMBLED) -> list[Request]:
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
