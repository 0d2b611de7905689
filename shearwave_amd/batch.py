"""Batched solves: many shear_fdtd problems of uneven sizes in one persistent launch per pass.

:class:`Batch` keeps every solve's fields resident in HBM (float32 storage,
increment form, each solve in the axis order that fills the tiles), projects
each force with the spectral-preconditioned CG where its tolerance admits it
and the rest with one batched CG (``sw_batch_project``), and advances all
active solves with one persistent kernel launch per pass (``sw_batch_step``),
replayed from a captured hipGraph (``sw_batch_graph_create``). A solve's pass is two time
steps in one read of its state (the two-step tile), or one step, as its pass
plan (:func:`pass_plan`) says. Work items are (solve, z tile,
y tile, x chunk) quadruples; when solves finish, the item list is rebuilt so
finished solves cost nothing. A solve's arithmetic is the single-solve tile
code with per-solve reductions in a fixed order, so its result does not
depend on the other solves in the batch.
"""

from __future__ import annotations

import ctypes as C
import functools
import heapq
import itertools
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi
from . import engine as E
from ._abi import check, lib
from .autotune import DEFAULT_CAP, balanced_chunk

STEP_BY = 8  # output rows of a batched step tile (blocks of 64 x 11 threads)
PASS_STATE = 16  # int64 of the batch's pass state: the global pass and the work-claim counters
MAX_PLANE = 1 << 28  # the kernels address a y-z plane with 32-bit offsets
CG_BY = 16  # tile rows of the batched CG (fixed by the kernels)
CG_SUB = 48  # x planes per partial sum of the batched CG (fixed by the kernels)
CG_XCHUNKS = (192, 96, 48)  # x planes per batched CG work item, longest first
CG_PROLOGUE = 6  # x planes a CG work item loads before its first output plane


def cg_xchunk(shapes, n_cu):
    """x planes per CG work item for one batched projection of grids `shapes`.

    Longer items pay fewer march prologues (about CG_PROLOGUE planes each),
    shorter ones balance the blocks: each candidate of CG_XCHUNKS is priced
    by the time its items take when the blocks (one per CU) claim them
    longest first, each item costing its planes plus a prologue, and the
    cheapest is taken (the longer on a tie). The kernels sum every dot
    product per CG_SUB-plane sub-chunk, so the choice does not change any
    result.
    """
    best = None
    for xc in CG_XCHUNKS:
        costs = []
        for nx, ny, nz in shapes:
            xe = min(nx, xc)
            tiles = math.ceil(ny / CG_BY) * math.ceil(nz / 64)
            for c in range(math.ceil(nx / xe)):
                costs += [min(xe, nx - c * xe) + CG_PROLOGUE] * tiles
        span = _claimed_makespan(costs, n_cu)
        if best is None or span < best[0]:
            best = (span, xc)
    return best[1]


def _claimed_makespan(costs, n_blocks):
    """Finish time of items claimed longest first by n_blocks blocks, each taking the next item when free."""
    costs = sorted(costs, reverse=True)
    if len(costs) <= n_blocks:
        return costs[0] if costs else 0
    free = [0] * n_blocks
    heapq.heapify(free)
    for c in costs:
        heapq.heappush(free, heapq.heappop(free) + c)
    return max(free)


@dataclass
class SolveSpec:
    """One solve: grid, spacing, time step, materials, force pattern and envelope."""

    shape: tuple
    h: tuple
    dt: float
    rho: object
    mu: object
    eta: object = None
    force: tuple = (None, None, None)
    env: np.ndarray = None
    n_steps: int = 0
    vel_damp: float = 0.0
    tol: float = 1e-8
    max_iters: int = 20000
    xchunk: int = 0  # 0: from the autotuner / default
    meta: dict = field(default_factory=dict)

    @property
    def voxels(self):
        return int(np.prod(self.shape))


def lane_fill(shape, by=STEP_BY):
    """Fraction of the step tiles' (by x 64) lanes that cover voxels of a (nx, ny, nz) grid."""
    _, ny, nz = shape
    return ny * nz / (math.ceil(nz / 64) * 64 * math.ceil(ny / by) * by)


def axis_order(shape, by=STEP_BY):
    """The order in which a solve's axes are stored in HBM: stored axis i is given axis order[i].

    The stencil treats the three axes alike (each with its own spacing), so a
    solve may be stored with any axis fastest. The tiles are 64 voxels wide
    along the fastest axis and `by` rows along the middle one, so the order
    that fills the most lanes is taken (ties: the given order, then the
    longest marching axis). A 128 x 192 x 200 grid stored as given fills 78%
    of its lanes, stored as (200, 128, 192) all of them.
    """
    shape = tuple(int(v) for v in shape)

    def key(p):
        s = tuple(shape[i] for i in p)
        return (round(lane_fill(s, by), 9), p == (0, 1, 2), s[0])

    return max(itertools.permutations(range(3)), key=key)


def vec_to_given(v, perm):
    """A (3, n0, n1, n2) SoA vector field stored in axis order `perm` back in the given order.

    Stored axis i is given axis perm[i] and stored component i is the
    component along it: stored = [given[perm[i]].permute(perm) for i]."""
    if tuple(perm) == (0, 1, 2):
        return v
    inv = [int(i) for i in np.argsort(perm)]
    return torch.stack([v[inv[m]].permute(*inv) for m in range(3)])


def aos_to_given(a, perm):
    """An (n0, n1, n2, 3) AoS vector field stored in axis order `perm` back in the given order."""
    if tuple(perm) == (0, 1, 2):
        return a
    inv = [int(i) for i in np.argsort(perm)]
    return a.permute(*inv, 3)[..., inv].contiguous()


class _Solve:
    """Device buffers of one solve in a batch, stored in the axis order `perm` (given axis perm[i] first)."""

    def __init__(self, spec: SolveSpec, two_step=True, axes="auto", by=STEP_BY):
        dev = E.device()
        self.spec = spec
        given = tuple(int(s) for s in spec.shape)
        if len(given) != 3 or min(given) < 1:
            raise ValueError(f"solve shape {given} must be three positive sizes")
        self.perm = axis_order(given, by) if axes == "auto" else (0, 1, 2)
        self.inv = tuple(int(i) for i in np.argsort(self.perm))
        shape = tuple(given[i] for i in self.perm)
        if shape[1] * shape[2] >= MAX_PLANE:
            raise ValueError(f"solve shape {given}: the y-z plane must hold fewer than {MAX_PLANE} voxels")
        self.given_shape = given
        self.shape = shape
        self.h = tuple(float(spec.h[i]) for i in self.perm)
        st = E.stream_ptr()

        def up(x, dtype):
            if np.isscalar(x):
                return E.to_device(x, dtype, shape)
            t = E.to_device(x, dtype, given)
            return t if self.perm == (0, 1, 2) else t.permute(*self.perm).contiguous()

        rho64 = up(spec.rho, "float64")
        mu64 = up(spec.mu, "float64")
        eta64 = None if spec.eta is None else up(spec.eta, "float64")
        # every per-voxel field in one allocation, staggered across the L2 channels (engine.FieldArena):
        # mu/rho, eta/rho, 1/rho, the two (u, d) parities, the force pattern and the loop force
        arena = E.FieldArena(shape, torch.float32, 3 + 12 + 6)
        self.m = arena.take()
        self.e = None if eta64 is None else arena.take()
        self.inv_rho = arena.take()
        check(lib().sw_prepare_coefficients(
            _abi.SW_F32, rho64.data_ptr(), mu64.data_ptr(), None if eta64 is None else eta64.data_ptr(),
            self.m.data_ptr(), None if self.e is None else self.e.data_ptr(), self.inv_rho.data_ptr(),
            int(np.prod(shape)), st))
        del rho64, mu64, eta64
        self.u = arena.take((2, 3))
        self.d = arena.take((2, 3))
        self.f = arena.take((3,))
        for c in range(3):
            comp = spec.force[self.perm[c]]  # the component along stored axis c
            if comp is None:
                self.f[c].zero_()
            else:
                self.f[c].copy_(up(comp, "float32"))
        self.force = arena.take((3,))
        env = np.zeros(spec.n_steps) if spec.env is None else np.asarray(spec.env, dtype=np.float64)
        if env.shape[0] != spec.n_steps:
            raise ValueError("env length must equal n_steps")
        self.env = torch.from_numpy(np.ascontiguousarray(env)).to(dev)
        plan = pass_plan(spec.n_steps, two_step)
        self.n_passes = int(plan.shape[0])
        self.plan = torch.from_numpy(plan).to(dev)
        self.start = 0  # global pass of the first pass

    def footprint(self):
        n = int(np.prod(self.shape))
        return n * 4 * (3 + 12 + 6) + self.spec.n_steps * 8

    def given_vec(self, v):
        """A (3, ...) vector field of this solve's storage order in the given axis order."""
        return vec_to_given(v, self.perm)

    def export_aos(self, par):
        """(u, v) of state parity `par` as AoS (nx, ny, nz, 3) device tensors in the given axis order."""
        u, d = self.u[par], self.d[par]
        aos = self.shape + (3,)
        ua = torch.empty(aos, dtype=torch.float32, device=u.device)
        va = torch.empty(aos, dtype=torch.float32, device=u.device)
        check(lib().sw_export_state(_abi.SW_F32, _abi.SW_FORM_INCREMENT, _abi.ptrs(u[0], u[1], u[2]),
                                    _abi.ptrs(d[0], d[1], d[2]), 1.0 / float(self.spec.dt), ua.data_ptr(),
                                    va.data_ptr(), int(np.prod(self.shape)), E.stream_ptr()))
        return aos_to_given(ua, self.perm), aos_to_given(va, self.perm)


def pass_plan(n_steps, two_step=True):
    """Passes covering steps 0 .. n_steps-1: entry (first step << 1) | 1 for a two-step pass.

    With ``two_step`` consecutive steps are paired (the pass scales the force
    by each step's own env value) and an odd last step is a single pass;
    without it every pass is one step.
    """
    plan = []
    t = 0
    while t < n_steps:
        if two_step and t + 1 < n_steps:
            plan.append((t << 1) | 1)
            t += 2
        else:
            plan.append(t << 1)
            t += 1
    return np.asarray(plan, dtype=np.int32)


def solve_bytes(spec: SolveSpec, with_cg=True) -> int:
    """HBM footprint of one solve: fields, state, force, and (during projection) the CG
    vectors (phi in float64; r and two p buffers in float32)."""
    n = spec.voxels
    b = n * 4 * (3 + 12 + 6) + spec.n_steps * 8
    if with_cg:
        b += n * (8 + 3 * 4)
    return b


def item_cost(x_planes, z_width):
    """Relative time of a step work item: its planes plus the two-step prologue (3), each costing a fixed
    part and a part that grows with the lanes in use."""
    return (x_planes + 3) * (32 + z_width)


@functools.lru_cache(maxsize=4096)
def _tiles(nx, ny, nz, xc, by):
    """(z tile, y tile, x chunk) of one grid, z fastest, then y, then x chunk (int32, read-only)."""
    tx, ty, tz = np.meshgrid(np.arange(math.ceil(nx / xc)), np.arange(math.ceil(ny / by)),
                             np.arange(math.ceil(nz / 64)), indexing="ij")
    t = np.stack([tz.ravel(), ty.ravel(), tx.ravel()], axis=1).astype(np.int32)
    t.flags.writeable = False
    return t


def step_items(solves, by=STEP_BY, order="spatial"):
    """Work items (solve, z tile, y tile, x chunk).

    ``"spatial"``: each solve's items with z fastest, then y, then x chunk,
    so that the step kernel's per-XCD ranges hold neighbouring tiles;
    ``"longest"``: longest first (short ones fill in at the end of a pass);
    ``"spatial_lpt"``: the spatial list cut into the kernel's 8 per-XCD
    ranges, each range longest first (a stable sort, so equal items of a
    solve keep their spatial order)."""
    tiles = [_tiles(*s.shape, int(s.xchunk), int(by)) for s in solves]
    if not tiles:
        return np.zeros((0, 4), dtype=np.int32)
    items = np.empty((sum(t.shape[0] for t in tiles), 4), dtype=np.int32)
    items[:, 0] = np.repeat(np.arange(len(tiles), dtype=np.int32), [t.shape[0] for t in tiles])
    items[:, 1:] = np.concatenate(tiles)
    if order == "spatial":
        return items
    xc = np.repeat(np.asarray([int(s.xchunk) for s in solves]), [t.shape[0] for t in tiles])
    nx = np.repeat(np.asarray([s.shape[0] for s in solves]), [t.shape[0] for t in tiles])
    nz = np.repeat(np.asarray([s.shape[2] for s in solves]), [t.shape[0] for t in tiles])
    cost = item_cost(np.minimum(xc, nx - items[:, 3] * xc), np.minimum(64, nz - 64 * items[:, 1])).astype(np.int64)
    if order == "spatial_lpt":
        n = items.shape[0]
        out = np.empty_like(items)
        for j in range(8):
            lo, hi = j * n // 8, (j + 1) * n // 8
            out[lo:hi] = items[lo + np.argsort(-cost[lo:hi], kind="stable")]
        return out
    return items[np.argsort(-cost, kind="stable")]


def default_xchunk(shape):
    """x planes per batched work item without a tuner: balanced chunks under the default cap."""
    return balanced_chunk(shape[0], DEFAULT_CAP)


def n_step_items(solves, by=STEP_BY):
    return sum(math.ceil(s.shape[0] / s.xchunk) * math.ceil(s.shape[1] / by) * math.ceil(s.shape[2] / 64)
               for s in solves)


def fill_xchunks(solves, n_blocks, per_block=2, floor=48, by=STEP_BY):
    """Shorten x-chunks until the work items give each of `n_blocks` persistent blocks `per_block` of them.

    Long chunks pay fewer prologues (on the 48-solve workload whole columns
    beat 64-plane chunks by 4%), but a set of few solves (a serving tick, the
    tail of a batch) needs shorter items to occupy every CU: the longest
    chunk is halved, down to `floor` planes, until there are enough.
    """
    while n_step_items(solves, by) < per_block * n_blocks:
        s = max(solves, key=lambda s: min(s.xchunk, s.shape[0]), default=None)
        if s is None or min(s.xchunk, s.shape[0]) <= floor:
            return
        s.xchunk = max(floor, min(s.xchunk, s.shape[0]) // 2)


def n_compute_units():
    return int(torch.cuda.get_device_properties(E.device()).multi_processor_count)


def xchunk_for(s, tuner=None):
    """x planes per step work item of a stored solve: its spec's choice, else balanced chunks under the
    tuner's cap, else under the default cap."""
    if s.spec.xchunk:
        return int(s.spec.xchunk)
    return tuner.xchunk(s.shape) if tuner else default_xchunk(s.shape)


@dataclass
class ProjectionJob:
    """One force field of a batched projection: f (3, nx, ny, nz) float32 on the device, projected in place;
    out (3, nx, ny, nz) float32 receives scale * f_S * inv_rho."""

    shape: tuple
    h: tuple
    f: object
    out: object
    inv_rho: object
    scale: float
    tol: float
    max_iters: int


def batched_projection(jobs, xchunk=None):
    """One batched plain CG over several force fields (sw_batch_project): r and p in float32, phi, the dot
    products and the residual norms in float64, each field's CG stopping on its own (poisson_cg's rules,
    solver.py:1010-1050). `xchunk` (a multiple of CG_SUB) overrides :func:`cg_xchunk`; it changes the speed
    only. Returns one CG info dict per job; the stream is synchronized."""
    dev = E.device()
    cgs = (_abi.BatchCg * len(jobs))()
    items = []
    keep = []
    if xchunk is None:
        xchunk = cg_xchunk([j.shape for j in jobs], torch.cuda.get_device_properties(dev).multi_processor_count)
    elif xchunk < 1 or xchunk % CG_SUB:
        raise ValueError(f"xchunk must be a positive multiple of {CG_SUB}, got {xchunk}")
    cost = []
    for k, job in enumerate(jobs):
        nx, ny, nz = job.shape
        xc = min(nx, xchunk)
        n_items0 = len(items)
        for tx in range(math.ceil(nx / xc)):
            for ty in range(math.ceil(ny / CG_BY)):
                for tz in range(math.ceil(nz / 64)):
                    items.append((k, tz, ty, tx))
                    cost.append(item_cost(min(xc, nx - tx * xc), min(64, nz - 64 * tz)))
        phi = torch.empty(job.shape, dtype=torch.float64, device=dev)
        vec = torch.empty((3,) + tuple(job.shape), dtype=torch.float32, device=dev)  # r, p0, p1
        n_slots = math.ceil(nx / CG_SUB) * math.ceil(ny / CG_BY) * math.ceil(nz / 64)
        part = torch.empty(2 * n_slots, dtype=torch.float64, device=dev)
        keep += [phi, vec, part]
        c = cgs[k]
        c.nx, c.ny, c.nz = job.shape
        c.xchunk = xc
        c.dx, c.dy, c.dz = job.h
        c.tol = float(job.tol)
        c.max_iters = int(job.max_iters)
        for j in range(3):
            c.f[j] = job.f[j].data_ptr()
            c.out[j] = job.out[j].data_ptr()
        c.inv_rho = job.inv_rho.data_ptr()
        c.scale = float(job.scale)
        c.phi = phi.data_ptr()
        c.r, c.p0, c.p1 = (vec[j].data_ptr() for j in range(3))
        c.part = part.data_ptr()
        c.item0 = n_items0
        c.n_items = len(items) - n_items0
    # longest first: the dir and step kernels' blocks claim items in list order
    order = np.argsort(-np.asarray(cost, dtype=np.int64), kind="stable")
    items_t = torch.from_numpy(np.asarray(items, dtype=np.int32).reshape(-1, 4)[order].reshape(-1)).to(dev)
    desc = torch.frombuffer(bytearray(bytes(cgs)), dtype=torch.uint8).to(dev)
    # a CG state per job, then the kernels' work-claim counters
    states = torch.empty((len(jobs) + 1) * int(lib().sw_cg_state_bytes()), dtype=torch.uint8, device=dev)
    info = (_abi.CgInfo * len(jobs))()
    check(lib().sw_batch_project(desc.data_ptr(), len(jobs), items_t.data_ptr(), len(items), states.data_ptr(),
                                 info, E.stream_ptr()))
    torch.cuda.current_stream().synchronize()
    del keep
    return [{"iters": int(i.iters), "relres": float(i.relres), "max_div": float(i.max_div),
             "reason": _abi.CG_REASONS.get(int(i.reason), str(i.reason))} for i in info]


class Batch:
    """A set of solves advanced together on one GPU (float32 storage, increment form)."""

    def __init__(self, specs, *, tile_y=None, tuner=None, graph_steps=16, two_step=True, axes="auto"):
        self.specs = list(specs)
        # tile rows of the step launch: given, else the tuner's, else the default
        self.tile_y = int(tile_y) if tile_y else (tuner.tile_y() if tuner else STEP_BY)
        self.graph_steps = int(graph_steps)  # passes per captured graph, replayed (0: direct launches only)
        self.solves = []
        for sp in self.specs:
            s = _Solve(sp, two_step, axes, self.tile_y)
            s.xchunk = s.base_xchunk = xchunk_for(s, tuner)
            self.solves.append(s)
        # the global pass and the step kernel's work-claim counters (SW_BATCH_PASS_STATE int64)
        self.gpass = torch.zeros(PASS_STATE, dtype=torch.int64, device=E.device())
        self.cg_info = None

    # --------------------------------------------------------- descriptors
    def _desc_array(self, solves):
        return E.upload_async(self._desc_bytes(solves))

    def _desc_bytes(self, solves):
        """The sw_batch_solve descriptors of `solves` as a host byte array."""
        arr = (_abi.BatchSolve * len(solves))()
        for k, s in enumerate(solves):
            sp = s.spec
            d = arr[k]
            d.nx, d.ny, d.nz = s.shape
            d.xchunk = s.xchunk
            hx, hy, hz = s.h
            dt = float(sp.dt)
            d.cx, d.cy, d.cz = dt * dt / (hx * hx), dt * dt / (hy * hy), dt * dt / (hz * hz)
            d.vx, d.vy, d.vz = dt / (hx * hx), dt / (hy * hy), dt / (hz * hz)
            d.damp = 1.0 - float(sp.vel_damp)
            d.start = s.start
            d.n_steps = sp.n_steps
            d.mu_over_rho = s.m.data_ptr()
            d.eta_over_rho = None if s.e is None else s.e.data_ptr()
            for c in range(3):
                d.force[c] = s.force[c].data_ptr()
                for par in range(2):
                    d.u[par][c] = s.u[par, c].data_ptr()
                    d.d[par][c] = s.d[par, c].data_ptr()
            d.env = s.env.data_ptr()
            d.plan = s.plan.data_ptr()
            d.n_passes = s.n_passes
        return np.frombuffer(bytes(arr), dtype=np.uint8)

    # ---------------------------------------------------------- projection
    def project(self, solves=None, xchunk=None, method="auto"):
        """Project every solve's force; keeps dT^2 f_S / rho for the loop.

        ``method``: ``"each"`` projects the solves one after another with the
        single-solve projection (the spectral-preconditioned CG converges in a
        few iterations); ``"batched"`` runs one batched plain CG over all of
        them; ``"auto"`` takes ``"each"`` when every solve's projection would be
        preconditioned (``sw_project_preconditioned``). `xchunk` (a multiple of
        CG_SUB) overrides :func:`cg_xchunk` for the batched CG; it changes the
        speed only.
        """
        solves = self.solves if solves is None else solves
        if method == "auto":
            method = "each" if solves and all(
                lib().sw_project_preconditioned(_abi.SW_F32, *s.shape, float(s.spec.tol), int(s.spec.max_iters))
                for s in solves) else "batched"
        if method == "each":
            return self._project_each(solves)
        if method != "batched":
            raise ValueError(f"unknown projection method {method!r}")
        jobs = [ProjectionJob(shape=s.shape, h=s.h, f=s.f, out=s.force, inv_rho=s.inv_rho,
                              scale=float(s.spec.dt) ** 2, tol=float(s.spec.tol), max_iters=int(s.spec.max_iters))
                for s in solves]
        self.cg_info = batched_projection(jobs, xchunk)
        return self.cg_info

    def _project_each(self, solves):
        """The single-solve projection of each solve in turn, with one workspace sized for the largest."""
        nbytes = max(int(lib().sw_project_workspace_bytes(*s.shape)) for s in solves)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=E.device())
        st = E.stream_ptr()
        info = []
        for s in solves:
            ci = _abi.CgInfo()
            check(lib().sw_project_vector_field(
                _abi.SW_F32, _abi.SW_F32, s.f[0].data_ptr(), s.f[1].data_ptr(), s.f[2].data_ptr(), *s.shape, *s.h,
                float(s.spec.tol), int(s.spec.max_iters), ws.data_ptr(), ws.numel(), _abi.ptrs(*s.force),
                s.inv_rho.data_ptr(), float(s.spec.dt) ** 2, C.byref(ci), st))
            info.append({"iters": int(ci.iters), "relres": float(ci.relres), "max_div": float(ci.max_div),
                         "reason": _abi.CG_REASONS.get(int(ci.reason), str(ci.reason)),
                         "true_relres": float(ci.true_relres)})
        self.cg_info = info
        return info

    # ---------------------------------------------------------------- loop
    def _graph(self, desc_ptr, items_ptr, n_items):
        """A captured graph of graph_steps passes over the items (the caller destroys it)."""
        g = C.c_void_p()
        check(lib().sw_batch_graph_create(desc_ptr, items_ptr, n_items, self.gpass.data_ptr(), self.tile_y, 0,
                                          self.graph_steps, C.byref(g)))
        return g

    def _queue(self, desc_ptr, items_ptr, n_items, n_passes, g=None):
        """Queue n_passes global passes over the items: replays of graph `g` (graph_steps passes each), then
        direct launches for the rest."""
        rest = n_passes
        if g is not None:
            reps = n_passes // self.graph_steps
            for _ in range(reps):
                check(lib().sw_graph_launch(g, E.stream_ptr()))
            rest -= reps * self.graph_steps
        if rest:
            check(lib().sw_batch_step(desc_ptr, items_ptr, n_items, self.gpass.data_ptr(), self.tile_y, 0, rest,
                                      E.stream_ptr()))

    def _advance(self, desc, items_t, n_items, n_passes):
        """Queue and finish n_passes global passes over the items (descriptor and item tensors)."""
        if n_passes <= 0:
            return
        g = self._graph(desc.data_ptr(), items_t.data_ptr(), n_items) \
            if self.graph_steps and n_passes >= self.graph_steps else None
        try:
            self._queue(desc.data_ptr(), items_t.data_ptr(), n_items, n_passes, g)
        finally:
            torch.cuda.current_stream().synchronize()
            if g is not None:
                lib().sw_graph_destroy(g)

    def run(self, collect=False):
        """All solves start together; the item list shrinks as solves finish (x-chunks shorten when the
        remaining solves are too few to occupy every CU with long ones).

        Every segment's descriptors, item list and graph are made before the
        first pass is queued, in one upload: made segment by segment while
        the persistent step kernel ran, these uploads and graph instantiations
        slowed single passes 3.5x.

        With ``collect`` every solve's (u, v) is exported after the last pass
        and copied to pinned host memory in one stream of copies with a single
        synchronisation; :meth:`result` then returns those arrays. The export
        buffers (one device and one pinned allocation) are made while the
        passes run.
        """
        self._collected = None
        t0 = time.perf_counter()
        t_buffers = 0.0
        for s in self.solves:
            s.start = 0
        self.gpass.zero_()
        ends = sorted(set(s.n_passes for s in self.solves))
        n_cu = n_compute_units()
        done = 0
        segs, parts, off = [], [], 0  # (passes, descriptor offset, item offset, items) per segment

        def put(a):
            nonlocal off
            b = np.frombuffer(np.ascontiguousarray(a).tobytes(), dtype=np.uint8)
            pad = (-b.size) % 256
            parts.append(np.concatenate([b, np.zeros(pad, np.uint8)]) if pad else b)
            o, off = off, off + b.size + pad
            return o

        # the descriptors differ between segments only in their x chunks
        base_desc = self._desc_bytes(self.solves)
        size = C.sizeof(_abi.BatchSolve)
        xoff = _abi.BatchSolve.xchunk.offset
        for end in ends:
            active = [k for k, s in enumerate(self.solves) if s.n_passes > done]
            act = [self.solves[k] for k in active]
            for s in act:
                s.xchunk = s.base_xchunk
            fill_xchunks(act, n_cu, by=self.tile_y)
            items = step_items(act, self.tile_y)
            # map local solve index back to the full descriptor array
            items[:, 0] = np.asarray(active, dtype=np.int32)[items[:, 0]]
            desc = base_desc.copy()
            xc = desc.reshape(-1, size)[:, xoff:xoff + 4].view(np.int32)
            xc[:, 0] = [s.xchunk for s in self.solves]
            segs.append((end - done, put(desc), put(items.reshape(-1)), items.shape[0]))
            done = end
        blob = E.upload_async(np.concatenate(parts))
        base = blob.data_ptr()
        graphs = [self._graph(base + d, base + i, n) if self.graph_steps and p >= self.graph_steps else None
                  for p, d, i, n in segs]
        main = torch.cuda.current_stream()
        t_start, t_end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t_start.record(main)
        try:
            for (p, d, i, n), g in zip(segs, graphs):
                self._queue(base + d, base + i, n, p, g)
            t_end.record(main)
            t_queued = time.perf_counter() - t0
            if collect:
                # allocated while the passes run (the host would otherwise idle): pinning host memory
                # beside the persistent kernel measured no slower passes
                exports = self._export_buffers()
                t_buffers = time.perf_counter() - t0 - t_queued
                # after the last pass: copies to the host running beside the persistent kernel slowed its
                # passes up to 2.8x, more than they save
                for k in range(len(self.solves)):
                    self._export_into(k, exports[k])
                for buf in exports:
                    self._copy_out(buf)
        finally:
            main.synchronize()
            for g in graphs:
                if g is not None:
                    lib().sw_graph_destroy(g)
            del blob
        self.last_loop_s = t_start.elapsed_time(t_end) / 1e3  # the passes on the main stream
        self.last_run_timing = {"export_buffers_s": t_buffers, "queued_s": t_queued,
                                "total_s": time.perf_counter() - t0}
        if collect:
            self._collected = [(e[4].numpy(), e[5].numpy()) for e in exports]

    def _export_buffers(self):
        """Per solve: stored-order AoS u and v (device), given-order copies when the solve is stored
        permuted (device, else None), and pinned host arrays for the results; carved from one device and
        one pinned allocation."""
        sizes = [3 * int(np.prod(s.shape)) for s in self.solves]
        n_dev = sum(2 * n * (2 if tuple(s.perm) != (0, 1, 2) else 1) for n, s in zip(sizes, self.solves))
        dev_all = torch.empty(n_dev, dtype=torch.float32, device=E.device())
        host_all = torch.empty(2 * sum(sizes), dtype=torch.float32, pin_memory=True)
        out, od, oh = [], 0, 0
        for n, s in zip(sizes, self.solves):
            aos = s.shape + (3,)
            given = tuple(s.given_shape) + (3,)

            def take(buf, off, shape):
                return buf[off:off + n].view(shape), off + n

            ua, od = take(dev_all, od, aos)
            va, od = take(dev_all, od, aos)
            gu = gv = None
            if tuple(s.perm) != (0, 1, 2):
                gu, od = take(dev_all, od, given)
                gv, od = take(dev_all, od, given)
            hu, oh = take(host_all, oh, given)
            hv, oh = take(host_all, oh, given)
            out.append((ua, va, gu, gv, hu, hv))
        return out

    @staticmethod
    def _copy_out(buf):
        """Queue the copies of an exported (u, v) (of _export_buffers) to its pinned host arrays."""
        ua, va, gu, gv, hu, hv = buf
        if gu is not None:
            ua, va = gu, gv
        hu.copy_(ua, non_blocking=True)
        hv.copy_(va, non_blocking=True)

    def _export_into(self, k, buf):
        """Queue the export of solve k's final (u, v) into the device arrays of `buf` (of _export_buffers)."""
        s = self.solves[k]
        ua, va, gu, gv, hu, hv = buf
        par = s.n_passes & 1
        u, d = s.u[par], s.d[par]
        check(lib().sw_export_state(_abi.SW_F32, _abi.SW_FORM_INCREMENT, _abi.ptrs(u[0], u[1], u[2]),
                                    _abi.ptrs(d[0], d[1], d[2]), 1.0 / float(s.spec.dt), ua.data_ptr(),
                                    va.data_ptr(), int(np.prod(s.shape)), E.stream_ptr()))
        if gu is not None:
            inv = [int(i) for i in np.argsort(s.perm)]
            for src, dst in ((ua, gu), (va, gv)):
                view = src.permute(*inv, 3)
                for m in range(3):
                    dst[..., m].copy_(view[..., inv[m]])

    # ------------------------------------------------------------- results
    def state(self, k):
        """(u, d) of solve k after its n_steps steps, as (3, nx, ny, nz) device tensors (given axis order)."""
        s = self.solves[k]
        par = s.n_passes & 1
        return s.given_vec(s.u[par]), s.given_vec(s.d[par])

    def result(self, k, to_host=True):
        """(u, v) of solve k in the reference's AoS (nx, ny, nz, 3) layout (host arrays collected by
        ``run(collect=True)`` when there are any)."""
        if to_host and getattr(self, "_collected", None) is not None:
            return self._collected[k]
        s = self.solves[k]
        ua, va = s.export_aos(s.n_passes & 1)
        if to_host:
            return E.to_host(ua), E.to_host(va)
        return ua, va
