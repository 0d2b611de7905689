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
import itertools
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi
from . import engine as E
from ._abi import check, lib

STEP_BY = 8  # output rows of a batched step tile (blocks of 64 x 11 threads)
PASS_STATE = 16  # int64 of the batch's pass state: the global pass and the work-claim counters
MAX_PLANE = 1 << 28  # the kernels address a y-z plane with 32-bit offsets
CG_BY = 16  # tile rows of the batched CG (fixed by the kernels)
CG_SUB = 48  # x planes per partial sum of the batched CG (fixed by the kernels)
CG_XCHUNKS = (192, 96, 48)  # x planes per batched CG work item, longest first
CG_ITEMS_PER_CU = 4  # a longer item is taken only while the call keeps this many items per CU


def cg_xchunk(shapes, n_cu):
    """x planes per CG work item for one batched projection of grids `shapes`.

    Longer items pay fewer march prologues, but a call with few solves needs
    short items to occupy every CU: the longest of CG_XCHUNKS that still gives
    CG_ITEMS_PER_CU items per CU, else the shortest. The kernels sum every
    dot product per CG_SUB-plane sub-chunk, so the choice does not change any
    result.
    """
    def n_items(xc):
        return sum(math.ceil(nx / xc) * math.ceil(ny / CG_BY) * math.ceil(nz / 64) for nx, ny, nz in shapes)

    for xc in CG_XCHUNKS:
        if n_items(xc) >= CG_ITEMS_PER_CU * n_cu:
            return xc
    return CG_XCHUNKS[-1]


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

    def __init__(self, spec: SolveSpec, two_step=True, axes="auto"):
        dev = E.device()
        self.spec = spec
        given = tuple(int(s) for s in spec.shape)
        if len(given) != 3 or min(given) < 1:
            raise ValueError(f"solve shape {given} must be three positive sizes")
        self.perm = axis_order(given) if axes == "auto" else (0, 1, 2)
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
        self.m = torch.empty(shape, dtype=torch.float32, device=dev)
        self.e = None if eta64 is None else torch.empty(shape, dtype=torch.float32, device=dev)
        self.inv_rho = torch.empty(shape, dtype=torch.float32, device=dev)
        check(lib().sw_prepare_coefficients(
            _abi.SW_F32, rho64.data_ptr(), mu64.data_ptr(), None if eta64 is None else eta64.data_ptr(),
            self.m.data_ptr(), None if self.e is None else self.e.data_ptr(), self.inv_rho.data_ptr(),
            int(np.prod(shape)), st))
        del rho64, mu64, eta64
        self.u = torch.zeros((2, 3) + shape, dtype=torch.float32, device=dev)
        self.d = torch.zeros((2, 3) + shape, dtype=torch.float32, device=dev)
        self.f = torch.empty((3,) + shape, dtype=torch.float32, device=dev)
        for c in range(3):
            comp = spec.force[self.perm[c]]  # the component along stored axis c
            if comp is None:
                self.f[c].zero_()
            else:
                self.f[c].copy_(up(comp, "float32"))
        self.force = torch.empty_like(self.f)
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


def step_items(solves, by=STEP_BY, order="spatial"):
    """Work items (solve, z tile, y tile, x chunk).

    ``"spatial"``: each solve's items with z fastest, then y, then x chunk,
    so that the step kernel's per-XCD ranges hold neighbouring tiles;
    ``"longest"``: longest first (short ones fill in at the end of a pass)."""
    items, cost = [], []
    for s_idx, s in enumerate(solves):
        nx, ny, nz = s.shape
        xc = s.xchunk
        tx, ty, tz = np.meshgrid(np.arange(math.ceil(nx / xc)), np.arange(math.ceil(ny / by)),
                                 np.arange(math.ceil(nz / 64)), indexing="ij")
        tx, ty, tz = tx.ravel(), ty.ravel(), tz.ravel()
        items.append(np.stack([np.full_like(tx, s_idx), tz, ty, tx], axis=1))
        cost.append(item_cost(np.minimum(xc, nx - tx * xc), np.minimum(64, nz - 64 * tz)))
    if not items:
        return np.zeros((0, 4), dtype=np.int32)
    items = np.concatenate(items).astype(np.int32)
    if order == "spatial":
        return items
    return items[np.argsort(-np.concatenate(cost).astype(np.int64), kind="stable")]


def default_xchunk(shape):
    """x planes per batched work item when the autotuner has no entry for the grid class."""
    return int(min(shape[0], 64))


def n_step_items(solves, by=STEP_BY):
    return sum(math.ceil(s.shape[0] / s.xchunk) * math.ceil(s.shape[1] / by) * math.ceil(s.shape[2] / 64)
               for s in solves)


def fill_xchunks(solves, n_blocks, per_block=2, floor=48):
    """Shorten x-chunks until the work items give each of `n_blocks` persistent blocks `per_block` of them.

    Long chunks pay fewer prologues (on the 48-solve workload whole columns
    beat 64-plane chunks by 4%), but a set of few solves (a serving tick, the
    tail of a batch) needs shorter items to occupy every CU: the longest
    chunk is halved, down to `floor` planes, until there are enough.
    """
    while n_step_items(solves) < per_block * n_blocks:
        s = max(solves, key=lambda s: min(s.xchunk, s.shape[0]), default=None)
        if s is None or min(s.xchunk, s.shape[0]) <= floor:
            return
        s.xchunk = max(floor, min(s.xchunk, s.shape[0]) // 2)


def n_compute_units():
    return int(torch.cuda.get_device_properties(E.device()).multi_processor_count)


def xchunk_for(s, tuner=None):
    """x planes per step work item of a stored solve: its spec's choice, else the tuner's for its stored
    shape, else the default."""
    if s.spec.xchunk:
        return int(s.spec.xchunk)
    return tuner.xchunk(s.shape) if tuner else default_xchunk(s.shape)


class Batch:
    """A set of solves advanced together on one GPU (float32 storage, increment form)."""

    def __init__(self, specs, *, tile_y=STEP_BY, tuner=None, graph_steps=16, two_step=True, axes="auto"):
        self.specs = list(specs)
        self.tile_y = int(tile_y)
        self.graph_steps = int(graph_steps)  # passes per captured graph, replayed (0: direct launches only)
        self.solves = []
        for sp in self.specs:
            s = _Solve(sp, two_step, axes)
            s.xchunk = s.base_xchunk = xchunk_for(s, tuner)
            self.solves.append(s)
        # the global pass and the step kernel's work-claim counters (SW_BATCH_PASS_STATE int64)
        self.gpass = torch.zeros(PASS_STATE, dtype=torch.int64, device=E.device())
        self.cg_info = None

    # --------------------------------------------------------- descriptors
    def _desc_array(self, solves):
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
        return E.upload_async(np.frombuffer(bytes(arr), dtype=np.uint8))

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
        dev = E.device()
        cgs = (_abi.BatchCg * len(solves))()
        items = []
        keep = []
        if xchunk is None:
            xchunk = cg_xchunk([s.shape for s in solves], torch.cuda.get_device_properties(dev).multi_processor_count)
        elif xchunk < 1 or xchunk % CG_SUB:
            raise ValueError(f"xchunk must be a positive multiple of {CG_SUB}, got {xchunk}")
        cost = []
        for k, s in enumerate(solves):
            sp = s.spec
            nx, ny, nz = s.shape
            xc = min(nx, xchunk)
            n_items0 = len(items)
            for tx in range(math.ceil(nx / xc)):
                for ty in range(math.ceil(ny / CG_BY)):
                    for tz in range(math.ceil(nz / 64)):
                        items.append((k, tz, ty, tx))
                        cost.append(item_cost(min(xc, nx - tx * xc), min(64, nz - 64 * tz)))
            phi = torch.empty(s.shape, dtype=torch.float64, device=dev)
            vec = torch.empty((3,) + s.shape, dtype=torch.float32, device=dev)  # r, p0, p1
            n_slots = math.ceil(nx / CG_SUB) * math.ceil(ny / CG_BY) * math.ceil(nz / 64)
            part = torch.empty(2 * n_slots, dtype=torch.float64, device=dev)
            keep += [phi, vec, part]
            c = cgs[k]
            c.nx, c.ny, c.nz = s.shape
            c.xchunk = xc
            c.dx, c.dy, c.dz = s.h
            c.tol = float(sp.tol)
            c.max_iters = int(sp.max_iters)
            for j in range(3):
                c.f[j] = s.f[j].data_ptr()
                c.out[j] = s.force[j].data_ptr()
            c.inv_rho = s.inv_rho.data_ptr()
            c.scale = float(sp.dt) ** 2
            c.phi = phi.data_ptr()
            c.r, c.p0, c.p1 = (vec[j].data_ptr() for j in range(3))
            c.part = part.data_ptr()
            c.item0 = n_items0
            c.n_items = len(items) - n_items0
        # longest first: the dir and step kernels' blocks claim items in list order
        order = np.argsort(-np.asarray(cost, dtype=np.int64), kind="stable")
        items_t = torch.from_numpy(np.asarray(items, dtype=np.int32).reshape(-1, 4)[order].reshape(-1)).to(dev)
        desc = torch.frombuffer(bytearray(bytes(cgs)), dtype=torch.uint8).to(dev)
        # a CG state per solve, then the kernels' work-claim counters
        states = torch.empty((len(solves) + 1) * int(lib().sw_cg_state_bytes()), dtype=torch.uint8, device=dev)
        info = (_abi.CgInfo * len(solves))()
        check(lib().sw_batch_project(desc.data_ptr(), len(solves), items_t.data_ptr(), len(items), states.data_ptr(),
                                     info, E.stream_ptr()))
        torch.cuda.current_stream().synchronize()
        del keep
        self.cg_info = [{"iters": int(i.iters), "relres": float(i.relres), "max_div": float(i.max_div),
                         "reason": _abi.CG_REASONS.get(int(i.reason), str(i.reason))} for i in info]
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
    def _advance(self, desc, items_t, n_items, n_passes, keep=None):
        """Queue n_passes global passes over the items: whole graphs of graph_steps passes, replayed, then
        direct launches for the rest. A graph is destroyed after a synchronisation: at once, or, when a list
        `keep` is given, by the caller (which also keeps desc and items alive until then)."""
        if n_passes <= 0:
            return
        if self.graph_steps and n_passes >= self.graph_steps:
            g = C.c_void_p()
            check(lib().sw_batch_graph_create(desc.data_ptr(), items_t.data_ptr(), n_items, self.gpass.data_ptr(),
                                              self.tile_y, 0, self.graph_steps, C.byref(g)))
            try:
                reps = n_passes // self.graph_steps
                for _ in range(reps):
                    check(lib().sw_graph_launch(g, E.stream_ptr()))
                rest = n_passes - reps * self.graph_steps
            finally:
                if keep is None:
                    torch.cuda.current_stream().synchronize()
                    lib().sw_graph_destroy(g)
                else:
                    keep.append((g, desc, items_t))
        else:
            rest = n_passes
        if rest:
            check(lib().sw_batch_step(desc.data_ptr(), items_t.data_ptr(), n_items, self.gpass.data_ptr(), self.tile_y,
                                      0, rest, E.stream_ptr()))

    def run(self):
        """All solves start together; the item list shrinks as solves finish (x-chunks shorten when the
        remaining solves are too few to occupy every CU with long ones)."""
        for s in self.solves:
            s.start = 0
        self.gpass.zero_()
        ends = sorted(set(s.n_passes for s in self.solves))
        n_cu = n_compute_units()
        done = 0
        graphs = []  # (graph, desc, items) of the segments, destroyed after the final synchronisation
        for end in ends:
            active = [k for k, s in enumerate(self.solves) if s.n_passes > done]
            act = [self.solves[k] for k in active]
            for s in act:
                s.xchunk = s.base_xchunk
            fill_xchunks(act, n_cu)
            desc = self._desc_array(self.solves)
            items = step_items(act, self.tile_y)
            # map local solve index back to the full descriptor array
            items[:, 0] = np.asarray(active, dtype=np.int32)[items[:, 0]]
            items_t = E.upload_async(items.reshape(-1))
            self._advance(desc, items_t, items.shape[0], end - done, keep=graphs)
            graphs.append((None, desc, items_t))  # direct launches of the segment read them too
            done = end
        torch.cuda.current_stream().synchronize()
        for g, _, _ in graphs:
            if g is not None:
                lib().sw_graph_destroy(g)

    # ------------------------------------------------------------- results
    def state(self, k):
        """(u, d) of solve k after its n_steps steps, as (3, nx, ny, nz) device tensors (given axis order)."""
        s = self.solves[k]
        par = s.n_passes & 1
        return s.given_vec(s.u[par]), s.given_vec(s.d[par])

    def result(self, k, to_host=True):
        """(u, v) of solve k in the reference's AoS (nx, ny, nz, 3) layout."""
        s = self.solves[k]
        ua, va = s.export_aos(s.n_passes & 1)
        if to_host:
            return E.to_host(ua), E.to_host(va)
        return ua, va
