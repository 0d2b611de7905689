"""Serving: continuous batching of shear_fdtd solve requests on one GPU.

A request (one solve) arrives at some time; the server admits it as soon as
its HBM footprint fits (fields, state, force and, during its projection, the
CG vectors: phi in float64, r and p in float32) and a slot is free. Admission (upload, coefficients, and
one batched projection of every request admitted together) runs on a second
stream, so the solves already on the GPU keep stepping meanwhile. Admitted
solves join the running batch at the next tick: every tick is `tick_passes`
global passes of ONE persistent launch per pass over all active solves, each
at its own pass (two steps in one read of its state, or one step). Finished solves are exported (u, v in the reference's
AoS layout) into pinned host memory; latency is arrival -> result on the host.

A solve's arithmetic is the same whichever solves share its ticks (per-solve
reductions in a fixed order; the step is pointwise), so every result is
bit-identical alone and in the mix.
"""

from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi
from . import engine as E
from ._abi import check, lib
from .batch import Batch, SolveSpec, _Solve, solve_bytes, step_items


@dataclass
class Request:
    rid: int
    spec: SolveSpec
    arrival: float  # seconds after the server starts
    admitted: float = math.nan
    completed: float = math.nan
    start_pass: int = -1
    n_passes: int = 0
    cg: dict = field(default_factory=dict)
    u: np.ndarray = None
    v: np.ndarray = None

    @property
    def latency(self):
        return self.completed - self.arrival


class Server:
    """Continuous-batching solve server (float32 storage, increment form)."""

    def __init__(self, tick_passes=10, max_active=128, budget_bytes=None, tuner=None, tile_y=8, two_step=True):
        self.tick_passes = int(tick_passes)
        self.two_step = bool(two_step)
        self.max_active = int(max_active)
        self.tile_y = int(tile_y)
        self.tuner = tuner
        dev = E.device()
        if budget_bytes is None:
            free, _total = torch.cuda.mem_get_info(dev)
            budget_bytes = int(0.85 * free)
        self.budget = int(budget_bytes)
        self.main = torch.cuda.current_stream(dev)
        self.admit_stream = torch.cuda.Stream(dev)  # default priority
        self.gpass_dev = torch.zeros(1, dtype=torch.int64, device=dev)
        self.gpass = 0
        self._helper = Batch([], tuner=tuner)  # descriptor / projection helpers

    # ------------------------------------------------------------ admission
    def _admit(self, reqs, t0):
        """Upload and project `reqs` on the admission stream; returns their device solves."""
        with torch.cuda.stream(self.admit_stream):
            solves = []
            for r in reqs:
                s = _Solve(r.spec, self.two_step)
                s.xchunk = self.tuner.xchunk(r.spec.shape) if self.tuner else min(r.spec.shape[0], 64)
                solves.append(s)
            info = self._helper.project(solves)
        ev = torch.cuda.Event()
        ev.record(self.admit_stream)
        self.main.wait_event(ev)
        now = time.perf_counter() - t0
        for r, s, i in zip(reqs, solves, info):
            for t in (s.m, s.e, s.inv_rho, s.u, s.d, s.f, s.force, s.env, s.plan):
                if t is not None:
                    t.record_stream(self.main)
            s.start = self.gpass
            r.start_pass = self.gpass
            r.n_passes = s.n_passes
            r.admitted = now
            r.cg = i
        return solves

    # ------------------------------------------------------------------ run
    def run(self, requests, *, verbose=False):
        """Serve a trace of requests (sorted by arrival); returns (requests, summary)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        nxt = 0
        pending = collections.deque()
        active = []  # (request, solve)
        retiring = []  # (request, solve, event, pinned u, pinned v)
        used = 0
        tick_events = collections.deque()
        desc = items_t = None
        n_items = 0
        changed = False
        t0 = time.perf_counter()
        done = 0

        while done < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                pending.append(reqs[nxt])
                nxt += 1
            # admit in arrival order while memory and slots allow
            batch = []
            while pending and len(active) + len(batch) < self.max_active:
                need = solve_bytes(pending[0].spec, with_cg=True)
                if used + need > self.budget and (active or batch):
                    break
                r = pending.popleft()
                used += solve_bytes(r.spec, with_cg=False)
                batch.append(r)
            if batch:
                for r, s in zip(batch, self._admit(batch, t0)):
                    active.append((r, s))
                changed = True
            if active:
                if changed:
                    sol = [s for _, s in active]
                    desc = self._helper._desc_array(sol)
                    items = step_items(sol, self.tile_y)
                    items_t = torch.from_numpy(items.reshape(-1).copy()).to(E.device())
                    n_items = items.shape[0]
                    changed = False
                # at most two ticks in flight
                while len(tick_events) >= 2:
                    tick_events.popleft().synchronize()
                check(lib().sw_batch_step(desc.data_ptr(), items_t.data_ptr(), n_items, self.gpass_dev.data_ptr(),
                                          self.tile_y, 0, self.tick_passes, self.main.cuda_stream))
                self.gpass += self.tick_passes
                ev = torch.cuda.Event()
                ev.record(self.main)
                tick_events.append(ev)
                still = []
                for r, s in active:
                    if s.start + s.n_passes <= self.gpass:
                        pu, pv = self._export(s)
                        e2 = torch.cuda.Event()
                        e2.record(self.main)
                        retiring.append((r, s, e2, pu, pv))
                        used -= solve_bytes(r.spec, with_cg=False)
                        changed = True
                    else:
                        still.append((r, s))
                active = still
            keep = []
            for r, s, e2, pu, pv in retiring:
                if e2.query():
                    r.completed = time.perf_counter() - t0
                    r.u, r.v = pu.numpy(), pv.numpy()
                    done += 1
                else:
                    keep.append((r, s, e2, pu, pv))
            retiring = keep
            if not active and not retiring and not pending and nxt < len(reqs):
                wait = reqs[nxt].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.05))
            elif not active and retiring:
                time.sleep(0.0002)
        wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in reqs])
        summary = {
            "requests": len(reqs),
            "wall_s": wall,
            "solves_per_s": len(reqs) / wall,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()),
            "global_passes": self.gpass,
            "voxel_steps": int(sum(r.spec.voxels * r.spec.n_steps for r in reqs)),
        }
        return reqs, summary

    def _export(self, s):
        """u, v of a finished solve into pinned host buffers (queued on the main stream)."""
        par = s.n_passes & 1
        u, d = s.u[par], s.d[par]
        aos = s.shape + (3,)
        ua = torch.empty(aos, dtype=torch.float32, device=u.device)
        va = torch.empty(aos, dtype=torch.float32, device=u.device)
        check(lib().sw_export_state(_abi.SW_F32, _abi.SW_FORM_INCREMENT, _abi.ptrs(u[0], u[1], u[2]),
                                    _abi.ptrs(d[0], d[1], d[2]), 1.0 / float(s.spec.dt), ua.data_ptr(), va.data_ptr(),
                                    int(np.prod(s.shape)), self.main.cuda_stream))
        pu = torch.empty(aos, dtype=torch.float32, pin_memory=True)
        pv = torch.empty(aos, dtype=torch.float32, pin_memory=True)
        pu.copy_(ua, non_blocking=True)
        pv.copy_(va, non_blocking=True)
        return pu, pv


def poisson_trace(rate, n_requests, seed=4, sizes=(64, 256), steps=(300, 1500), h=0.5e-3):
    """Seeded Poisson arrivals at `rate` req/s of solves with random grids, inclusion maps and pushes."""
    from . import workloads as W

    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n_requests):
        t += rng.exponential(1.0 / rate)
        shape = tuple(int(8 * round(rng.integers(sizes[0], sizes[1] + 1) / 8)) for _ in range(3))
        rho = (1000.0 * (1.0 + rng.uniform(-0.01, 0.01, shape))).astype(np.float32)
        mu = W.inclusion_map(shape, h, rng, int(rng.integers(3, 9)), 2000.0, mu_range=(2e3, 20e3)).astype(np.float32)
        eta = rng.uniform(0.5, 3.0, shape).astype(np.float32)
        n_steps = int(rng.integers(steps[0], steps[1] + 1))
        dt, dt_cfl = W.stable_dt(h, float(rho.min()), float(mu.max()), float(eta.max()))
        centre = [rng.uniform(0.25, 0.75) * (n - 1) * h for n in shape]
        bz = W.gaussian_push(shape, h, 2e-3, centre=centre).astype(np.float32)
        env = (np.arange(n_steps) * dt < 100e-6).astype(np.float64)
        spec = SolveSpec(shape=shape, h=(h, h, h), dt=dt, rho=rho, mu=mu, eta=eta, force=(None, None, bz), env=env,
                         n_steps=n_steps, tol=1e-8, max_iters=20000, meta={"index": k})
        out.append(Request(rid=k, spec=spec, arrival=t))
    return out
