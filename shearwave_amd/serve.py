"""Serving: continuous batching of shear_fdtd solve requests on one GPU.

A request (one solve) arrives at some time; the server hands it to admission
as soon as its HBM footprint fits (fields, state, force and, during its
projection, the CG vectors: phi in float64, r and p in float32) and a slot is
free. Admission (upload, coefficients, and one batched projection of every
request waiting for it) runs on a worker thread with its own stream, so the tick loop keeps the GPU busy meanwhile: the C
calls release the GIL and the worker only waits on its own stream. Admitted
solves join the running batch at the next tick: every tick is `tick_passes`
global passes of ONE persistent launch per pass over all active solves, each
at its own pass (two steps in one read of its state, or one step), replayed
from a hipGraph captured whenever the set of active solves changes; up to
`ticks_in_flight` ticks are queued on the device. Finished solves are
exported (u, v in the reference's AoS layout) into pinned host memory;
latency is arrival -> result on the host.

A solve's arithmetic is the same whichever solves share its ticks (per-solve
reductions in a fixed order; the step is pointwise; a solve's axis order in
HBM depends only on its shape), so every result is bit-identical alone and in
the mix.
"""

from __future__ import annotations

import collections
import ctypes as C
import math
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import engine as E
from ._abi import check, lib
from .batch import PASS_STATE, STEP_BY, Batch, SolveSpec, _Solve, fill_xchunks, n_compute_units, solve_bytes, \
    step_items, xchunk_for


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


class _TickGraph:
    """A tick's passes over one set of active solves, captured once and replayed (kept alive, with the
    descriptors and items it points at, until its last replay has finished)."""

    def __init__(self, desc, items_t, n_items, gpass_dev, tile_y, n_passes):
        self.desc, self.items_t = desc, items_t
        self.handle = C.c_void_p()
        check(lib().sw_batch_graph_create(desc.data_ptr(), items_t.data_ptr(), n_items, gpass_dev.data_ptr(), tile_y,
                                          0, n_passes, C.byref(self.handle)))
        self.last = None

    def launch(self, stream):
        check(lib().sw_graph_launch(self.handle, stream.cuda_stream))

    def destroy(self):
        if self.handle:
            lib().sw_graph_destroy(self.handle)
            self.handle = C.c_void_p()


class Server:
    """Continuous-batching solve server (float32 storage, increment form)."""

    def __init__(self, tick_passes=10, max_active=128, budget_bytes=None, tuner=None, tile_y=None, two_step=True,
                 ticks_in_flight=3, graph=True, admit_voxels=96 << 20, idle_admit_voxels=None):
        self.tick_passes = int(tick_passes)
        self.two_step = bool(two_step)
        self.max_active = int(max_active)
        self.tile_y = int(tile_y) if tile_y else (tuner.tile_y() if tuner else STEP_BY)
        self.tuner = tuner
        self.ticks_in_flight = max(1, int(ticks_in_flight))
        self.admit_voxels = int(admit_voxels)  # at most this many voxels per batched projection of admissions
        self.idle_admit_voxels = self.admit_voxels if idle_admit_voxels is None else int(idle_admit_voxels)
        self._stepping_voxels = 0  # voxels of the active solves (set by the tick loop, read by admission)
        self.graph = bool(graph)
        self.dev = E.device()
        if budget_bytes is None:
            free, _total = torch.cuda.mem_get_info(self.dev)
            budget_bytes = int(0.85 * free)
        self.budget = int(budget_bytes)
        self.main = torch.cuda.current_stream(self.dev)
        self.n_cu = n_compute_units()
        self.admit_stream = torch.cuda.Stream(self.dev)  # default priority
        self.gpass_dev = torch.zeros(PASS_STATE, dtype=torch.int64, device=self.dev)
        self.gpass = 0
        self._helper = Batch([], tuner=tuner, tile_y=self.tile_y)  # descriptor / projection helpers
        self.admission_s = []  # wall time of each admission (upload + batched projection)

    # ------------------------------------------------------------ admission
    def _admit_loop(self, todo, ready):
        """Worker thread: upload and project handed-over requests on the admission stream.

        Each projection takes every request waiting at that moment (in arrival
        order, up to `admit_voxels` voxels): a batched CG costs about the same
        number of iterations whatever it holds, so under load the groups grow
        and admission keeps up; when requests trickle in, each joins alone."""
        torch.cuda.set_device(self.dev)
        with torch.cuda.stream(self.admit_stream):
            stop = False
            while not stop:
                r = todo.get()
                if r is None:
                    return
                reqs, vox = [r], r.spec.voxels
                # with nothing stepping, a smaller group joins sooner and starts the stepping
                cap = self.admit_voxels if self._stepping_voxels > 0 else self.idle_admit_voxels
                while vox < cap:
                    try:
                        r = todo.get_nowait()
                    except queue.Empty:
                        break
                    if r is None:
                        stop = True
                        break
                    reqs.append(r)
                    vox += r.spec.voxels
                try:
                    t = time.perf_counter()
                    solves = []
                    for r in reqs:
                        s = _Solve(r.spec, self.two_step, "auto", self.tile_y)
                        s.xchunk = s.base_xchunk = xchunk_for(s, self.tuner)
                        solves.append(s)
                    info = self._helper.project(solves)  # waits on the admission stream only
                    ev = torch.cuda.Event()
                    ev.record(self.admit_stream)
                    self.admission_s.append(time.perf_counter() - t)
                    ready.put((reqs, solves, info, ev, None))
                except BaseException as e:  # noqa: BLE001 - handed to the tick loop, which raises it
                    ready.put((reqs, None, None, None, e))

    # ------------------------------------------------------------------ run
    def run(self, requests, *, verbose=False):
        """Serve a trace of requests (sorted by arrival); returns (requests, summary)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        todo, ready = queue.Queue(), queue.Queue()
        worker = threading.Thread(target=self._admit_loop, args=(todo, ready), daemon=True)
        worker.start()
        try:
            return self._serve(reqs, todo, ready)
        finally:
            todo.put(None)
            worker.join()

    def _serve(self, reqs, todo, ready):
        nxt = 0
        pending = collections.deque()
        admitting = 0  # requests handed to the worker, not yet back
        active = []  # (request, solve)
        retiring = []  # (request, event, pinned u, pinned v)
        used = 0
        tick_events = collections.deque()
        graphs = []  # tick graphs not yet destroyed
        tick = None  # the current active set's graph (or None: direct launches)
        desc = items_t = None
        n_items = 0
        changed = False
        t0 = time.perf_counter()
        done = 0
        ticks = []  # per tick: (time, active solves, active voxels, handed to admission, waiting)
        idle_s = 0.0  # time with no active solve while requests were waiting or being admitted

        while done < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                pending.append(reqs[nxt])
                nxt += 1
            # hand requests to admission, in arrival order, while memory and slots allow
            while pending and len(active) + admitting < self.max_active:
                need = solve_bytes(pending[0].spec, with_cg=True)
                if used + need > self.budget and (active or admitting):
                    break
                r = pending.popleft()
                used += solve_bytes(r.spec, with_cg=False)
                todo.put(r)
                admitting += 1
            # admitted solves join at the next tick
            while True:
                try:
                    got, solves, info, ev, err = ready.get_nowait()
                except queue.Empty:
                    break
                if err is not None:
                    raise err
                self.main.wait_event(ev)
                now = time.perf_counter() - t0
                for r, s, i in zip(got, solves, info):
                    for t in (s.m, s.e, s.inv_rho, s.u, s.d, s.f, s.force, s.env, s.plan):
                        if t is not None:
                            t.record_stream(self.main)
                    s.start = self.gpass
                    r.start_pass = self.gpass
                    r.n_passes = s.n_passes
                    r.admitted = now
                    r.cg = i
                    active.append((r, s))
                admitting -= len(got)
                changed = True
            if active:
                if changed:
                    sol = [s for _, s in active]
                    for s in sol:
                        s.xchunk = s.base_xchunk
                    fill_xchunks(sol, self.n_cu, by=self.tile_y)
                    desc = self._helper._desc_array(sol)
                    items = step_items(sol, self.tile_y)
                    items_t = E.upload_async(items.reshape(-1))
                    n_items = items.shape[0]
                    if self.graph:
                        tick = _TickGraph(desc, items_t, n_items, self.gpass_dev, self.tile_y, self.tick_passes)
                        graphs.append(tick)
                    changed = False
                while len(tick_events) >= self.ticks_in_flight:
                    tick_events.popleft().synchronize()
                if tick is not None:
                    tick.launch(self.main)
                else:
                    check(lib().sw_batch_step(desc.data_ptr(), items_t.data_ptr(), n_items, self.gpass_dev.data_ptr(),
                                              self.tile_y, 0, self.tick_passes, self.main.cuda_stream))
                self.gpass += self.tick_passes
                self._stepping_voxels = sum(s_.spec.voxels for _, s_ in active)
                ticks.append((time.perf_counter() - t0, len(active), self._stepping_voxels, admitting, len(pending)))
                ev = torch.cuda.Event()
                ev.record(self.main)
                tick_events.append(ev)
                if tick is not None:
                    tick.last = ev
                still = []
                for r, s in active:
                    if s.start + s.n_passes <= self.gpass:
                        pu, pv = self._export(s)
                        e2 = torch.cuda.Event()
                        e2.record(self.main)
                        retiring.append((r, e2, pu, pv))
                        used -= solve_bytes(r.spec, with_cg=False)
                        changed = True
                    else:
                        still.append((r, s))
                active = still
                if not active:
                    self._stepping_voxels = 0
                # graphs of earlier active sets, once their last replay is done
                keep = []
                for g in graphs:
                    if g is not tick and (g.last is None or g.last.query()):
                        g.destroy()
                    else:
                        keep.append(g)
                graphs = keep
            keep = []
            for r, e2, pu, pv in retiring:
                if e2.query():
                    r.completed = time.perf_counter() - t0
                    r.u, r.v = pu.numpy(), pv.numpy()
                    done += 1
                else:
                    keep.append((r, e2, pu, pv))
            retiring = keep
            if not active:
                if admitting or pending:
                    idle_s += 0.0005
                if retiring:
                    time.sleep(0.0002)
                elif admitting:
                    time.sleep(0.0005)  # the worker is uploading or projecting
                elif not pending and nxt < len(reqs):
                    wait = reqs[nxt].arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
        torch.cuda.synchronize(self.dev)
        for g in graphs:
            g.destroy()
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
            "admission_s_mean": float(np.mean(self.admission_s)) if self.admission_s else 0.0,
            "admission_groups": len(self.admission_s),
            "ticks": len(ticks),
            "active_solves_mean": float(np.mean([t[1] for t in ticks])) if ticks else 0.0,
            "active_solves_max": int(max((t[1] for t in ticks), default=0)),
            "active_mvox_mean": float(np.mean([t[2] for t in ticks]) / 1e6) if ticks else 0.0,
            "no_active_while_waiting_s": round(idle_s, 3),
        }
        self.tick_log = ticks
        return reqs, summary

    def _export(self, s):
        """u, v of a finished solve into pinned host buffers (queued on the main stream)."""
        ua, va = s.export_aos(s.n_passes & 1)
        pu = torch.empty(ua.shape, dtype=torch.float32, pin_memory=True)
        pv = torch.empty(va.shape, dtype=torch.float32, pin_memory=True)
        pu.copy_(ua, non_blocking=True)
        pv.copy_(va, non_blocking=True)
        return pu, pv


def request_specs(n_requests, seed=4, sizes=(64, 256), steps=(300, 1500), h=0.5e-3):
    """Seeded solve requests: grids with sides U{sizes} (multiples of 8), random inclusion maps, a Gaussian push at
    a random position, nT U{steps}, CG tol 1e-8."""
    from . import workloads as W

    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_requests):
        shape = tuple(int(8 * round(rng.integers(sizes[0], sizes[1] + 1) / 8)) for _ in range(3))
        rho = (1000.0 * (1.0 + rng.uniform(-0.01, 0.01, shape))).astype(np.float32)
        mu = W.inclusion_map(shape, h, rng, int(rng.integers(3, 9)), 2000.0, mu_range=(2e3, 20e3)).astype(np.float32)
        eta = rng.uniform(0.5, 3.0, shape).astype(np.float32)
        n_steps = int(rng.integers(steps[0], steps[1] + 1))
        dt, dt_cfl = W.stable_dt(h, float(rho.min()), float(mu.max()), float(eta.max()))
        centre = [rng.uniform(0.25, 0.75) * (n - 1) * h for n in shape]
        bz = W.gaussian_push(shape, h, 2e-3, centre=centre).astype(np.float32)
        env = (np.arange(n_steps) * dt < 100e-6).astype(np.float64)
        out.append(SolveSpec(shape=shape, h=(h, h, h), dt=dt, rho=rho, mu=mu, eta=eta, force=(None, None, bz),
                             env=env, n_steps=n_steps, tol=1e-8, max_iters=20000, meta={"index": k}))
    return out


def with_arrivals(specs, rate, seed=5):
    """Requests for `specs` arriving as a Poisson process of `rate` req/s (seeded exponential gaps)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k, sp in enumerate(specs):
        t += rng.exponential(1.0 / rate)
        out.append(Request(rid=k, spec=sp, arrival=t))
    return out


def poisson_trace(rate, n_requests, seed=4, sizes=(64, 256), steps=(300, 1500), h=0.5e-3):
    """Seeded Poisson arrivals at `rate` req/s of solves with random grids, inclusion maps and pushes."""
    return with_arrivals(request_specs(n_requests, seed, sizes, steps, h), rate, seed + 1)
