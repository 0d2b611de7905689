"""Batched solves: many shear_fdtd problems of uneven sizes in one persistent launch per pass.

:class:`Batch` keeps every solve's fields resident in HBM (float32 storage,
increment form), projects all their forces with one batched CG
(``sw_batch_project``), and advances all active solves with one persistent
kernel launch per pass (``sw_batch_step``), optionally replayed from a
captured hipGraph (``sw_batch_graph_create``). A solve's pass is two time
steps in one read of its state (the two-step tile), or one step, as its pass
plan (:func:`pass_plan`) says. Work items are (solve, z tile,
y tile, x chunk) quadruples; when solves finish, the item list is rebuilt so
finished solves cost nothing. A solve's arithmetic is the single-solve tile
code with per-solve reductions in a fixed order, so its result does not
depend on the other solves in the batch.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi
from . import engine as E
from ._abi import check, lib

STEP_BY = 8  # output rows of a batched step tile (blocks of 64 x 11 threads)
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


class _Solve:
    """Device buffers of one solve in a batch."""

    def __init__(self, spec: SolveSpec, two_step=True):
        dev = E.device()
        self.spec = spec
        shape = tuple(int(s) for s in spec.shape)
        if len(shape) != 3 or min(shape) < 1:
            raise ValueError(f"solve shape {shape} must be three positive sizes")
        if shape[1] * shape[2] >= MAX_PLANE:
            raise ValueError(f"solve shape {shape}: the y-z plane must hold fewer than {MAX_PLANE} voxels")
        self.shape = shape
        st = E.stream_ptr()
        rho64 = E.to_device(spec.rho, "float64", shape)
        mu64 = E.to_device(spec.mu, "float64", shape)
        eta64 = None if spec.eta is None else E.to_device(spec.eta, "float64", shape)
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
        for c, comp in enumerate(spec.force):
            if comp is None:
                self.f[c].zero_()
            else:
                self.f[c].copy_(E.to_device(comp, "float32", shape))
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


def step_items(solves, by=STEP_BY):
    """Work items (solve, z tile, y tile, x chunk): z fastest, each solve's items contiguous."""
    items = []
    for s_idx, s in enumerate(solves):
        nx, ny, nz = s.shape
        xc = s.xchunk
        for tx in range(math.ceil(nx / xc)):
            for ty in range(math.ceil(ny / by)):
                for tz in range(math.ceil(nz / 64)):
                    items.append((s_idx, tz, ty, tx))
    return np.asarray(items, dtype=np.int32).reshape(-1, 4)


def default_xchunk(shape):
    """x planes per batched work item when the autotuner has no entry for the grid class."""
    return int(min(shape[0], 64))


class Batch:
    """A set of solves advanced together on one GPU (float32 storage, increment form)."""

    def __init__(self, specs, *, tile_y=STEP_BY, tuner=None, graph_steps=0, two_step=True):
        self.specs = list(specs)
        self.tile_y = int(tile_y)
        self.graph_steps = int(graph_steps)  # passes per captured graph (0: no graph)
        self.solves = []
        for sp in self.specs:
            s = _Solve(sp, two_step)
            s.xchunk = int(sp.xchunk) if sp.xchunk else (tuner.xchunk(sp.shape) if tuner else default_xchunk(sp.shape))
            self.solves.append(s)
        self.gpass = torch.zeros(1, dtype=torch.int64, device=E.device())
        self.cg_info = None

    # --------------------------------------------------------- descriptors
    def _desc_array(self, solves):
        arr = (_abi.BatchSolve * len(solves))()
        for k, s in enumerate(solves):
            sp = s.spec
            d = arr[k]
            d.nx, d.ny, d.nz = s.shape
            d.xchunk = s.xchunk
            hx, hy, hz = (float(v) for v in sp.h)
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
        raw = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(E.device())
        return raw

    # ---------------------------------------------------------- projection
    def project(self, solves=None, xchunk=None):
        """Project every solve's force with one batched CG; keeps dT^2 f_S / rho for the loop.

        `xchunk` (a multiple of CG_SUB) overrides :func:`cg_xchunk`; it changes
        the speed only.
        """
        solves = self.solves if solves is None else solves
        dev = E.device()
        cgs = (_abi.BatchCg * len(solves))()
        items = []
        keep = []
        if xchunk is None:
            xchunk = cg_xchunk([s.shape for s in solves], torch.cuda.get_device_properties(dev).multi_processor_count)
        elif xchunk < 1 or xchunk % CG_SUB:
            raise ValueError(f"xchunk must be a positive multiple of {CG_SUB}, got {xchunk}")
        for k, s in enumerate(solves):
            sp = s.spec
            nx, ny, nz = s.shape
            xc = min(nx, xchunk)
            n_items0 = len(items)
            for tx in range(math.ceil(nx / xc)):
                for ty in range(math.ceil(ny / CG_BY)):
                    for tz in range(math.ceil(nz / 64)):
                        items.append((k, tz, ty, tx))
            phi = torch.empty(s.shape, dtype=torch.float64, device=dev)
            vec = torch.empty((3,) + s.shape, dtype=torch.float32, device=dev)  # r, p0, p1
            n_slots = math.ceil(nx / CG_SUB) * math.ceil(ny / CG_BY) * math.ceil(nz / 64)
            part = torch.empty(2 * n_slots, dtype=torch.float64, device=dev)
            keep += [phi, vec, part]
            c = cgs[k]
            c.nx, c.ny, c.nz = s.shape
            c.xchunk = xc
            c.dx, c.dy, c.dz = (float(v) for v in sp.h)
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
        items_t = torch.from_numpy(np.asarray(items, dtype=np.int32).reshape(-1)).to(dev)
        desc = torch.frombuffer(bytearray(bytes(cgs)), dtype=torch.uint8).to(dev)
        states = torch.empty(len(solves) * int(lib().sw_cg_state_bytes()), dtype=torch.uint8, device=dev)
        info = (_abi.CgInfo * len(solves))()
        check(lib().sw_batch_project(desc.data_ptr(), len(solves), items_t.data_ptr(), len(items), states.data_ptr(),
                                     info, E.stream_ptr()))
        torch.cuda.current_stream().synchronize()
        del keep
        self.cg_info = [{"iters": int(i.iters), "relres": float(i.relres), "max_div": float(i.max_div),
                         "reason": _abi.CG_REASONS.get(int(i.reason), str(i.reason))} for i in info]
        return self.cg_info

    # ---------------------------------------------------------------- loop
    def _advance(self, desc, items_t, n_items, n_passes):
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
                torch.cuda.current_stream().synchronize()
                lib().sw_graph_destroy(g)
        else:
            rest = n_passes
        if rest:
            check(lib().sw_batch_step(desc.data_ptr(), items_t.data_ptr(), n_items, self.gpass.data_ptr(), self.tile_y,
                                      0, rest, E.stream_ptr()))

    def run(self):
        """All solves start together; the item list shrinks as solves finish."""
        for s in self.solves:
            s.start = 0
        self.gpass.zero_()
        desc = self._desc_array(self.solves)
        ends = sorted(set(s.n_passes for s in self.solves))
        done = 0
        for end in ends:
            active = [k for k, s in enumerate(self.solves) if s.n_passes > done]
            items = step_items([self.solves[k] for k in active], self.tile_y)
            # map local solve index back to the full descriptor array
            items[:, 0] = np.asarray(active, dtype=np.int32)[items[:, 0]]
            items_t = torch.from_numpy(items.reshape(-1).copy()).to(E.device())
            self._advance(desc, items_t, items.shape[0], end - done)
            done = end
        torch.cuda.current_stream().synchronize()

    # ------------------------------------------------------------- results
    def state(self, k):
        """(u, d) of solve k after its n_steps steps, as (3, nx, ny, nz) device tensors."""
        s = self.solves[k]
        par = s.n_passes & 1
        return s.u[par], s.d[par]

    def result(self, k, to_host=True):
        """(u, v) of solve k in the reference's AoS (nx, ny, nz, 3) layout."""
        s = self.solves[k]
        u, d = self.state(k)
        n = int(np.prod(s.shape))
        aos = s.shape + (3,)
        ua = torch.empty(aos, dtype=torch.float32, device=u.device)
        va = torch.empty(aos, dtype=torch.float32, device=u.device)
        check(lib().sw_export_state(_abi.SW_F32, _abi.SW_FORM_INCREMENT, _abi.ptrs(u[0], u[1], u[2]),
                                    _abi.ptrs(d[0], d[1], d[2]), 1.0 / float(s.spec.dt), ua.data_ptr(), va.data_ptr(),
                                    n, E.stream_ptr()))
        if to_host:
            return E.to_host(ua), E.to_host(va)
        return ua, va
