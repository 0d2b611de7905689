"""Autotuner for the launch shape of the time step: tile rows and x-chunk depth.

Both step paths march x-chunks of (rows x 64) tiles of (y, z) columns; a
tile's row count trades its halo (two rows above and below for a two-step
pass) against the registers and LDS of one block, and the chunk depth trades
the chunk's prologue (three planes per two-step chunk) against the balance
of the grid of blocks. The tuner measures both, on the device, once per
device and workload class, and caches the fastest in a JSON file
(``$SW_AUTOTUNE_CACHE`` or ``~/.cache/shearwave_amd/autotune.json``).

Batched solves (``Tuner.batch_shape``): one persistent launch per pass runs
every solve's items, so the tile rows are one per batch, and a pass ends
when its last item does, so the chunk depth is a property of the whole mix,
not of one grid (timed on copies of one grid at a time, per-grid depths
picked whole columns and made the 48-solve mix run at 110 Gvoxel-steps/s,
against 128 with every solve cut under one cap). So the tuner times the
batched pass on a fixed synthetic mix of uneven grids (`TUNE_MIX`) for every
pair of tile rows in `BATCH_ROWS` and cap in `CAPS`; a batch then uses the
fastest rows and cuts each solve's marching axis into the fewest equal
chunks no longer than the cap (`balanced_chunk`). The tile's orientation
(which grid axis is stored fastest, ``batch.axis_order``) is fixed by the
lanes it fills, and a batch with too few items for the GPU shortens its
chunks further (``batch.fill_xchunks``).

A single drop-in solve (``Tuner.pass_shape``): the two-step pass of one
grid, timed on the solve's own buffers before its loop (``sw_fdtd_trial_passes``
reads the state and writes the next-step buffers without advancing the
state) for every tile-row count the storage type has (`PASS_ROWS`) and every
depth that cuts the marching axis into `PASS_SPLITS` equal chunks, and 0
(the depth that makes the grid one wave of resident blocks), cached per
storage type, viscosity and grid shape. Equal chunks only: a depth that
leaves a short last chunk leaves the last wave of blocks part idle (at 256^3,
96-plane chunks ran 35% slower than 128-plane ones). Every candidate computes
each voxel with the same arithmetic, so the choice changes the time and never
the result.
"""

from __future__ import annotations

import json
import math
import os
import threading

CAPS = (64, 80, 96, 128, 160, 192, 256)
BATCH_ROWS = (4, 8, 12)
DEFAULT_CAP = 96  # the measured best on the 48-solve mix; used without a tuner
DEFAULT_ROWS = 8

PASS_ROWS = {"float32": (8, 12), "float64": (4,)}
PASS_SPLITS = (1, 2, 3, 4, 6, 8, 12, 16)  # chunks along the marching axis
PASS_MIN_XCHUNK = 8

# The tuning mix: 24 grids with stored extents drawn from {64, 72, .., 256} (seed 11).
TUNE_MIX_SEED = 11
TUNE_MIX_SOLVES = 24


def balanced_chunk(nx, cap):
    """x planes per work item: the fewest chunks no longer than `cap`, of equal length."""
    nx = int(nx)
    n = max(1, math.ceil(nx / max(1, int(cap))))
    return math.ceil(nx / n)


def tune_mix_shapes(seed=TUNE_MIX_SEED, n=TUNE_MIX_SOLVES):
    """Stored shapes of the tuning mix (deterministic)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    return [tuple(int(8 * v) for v in rng.integers(8, 33, 3)) for _ in range(n)]


def batch_candidates():
    """(tile rows, cap) pairs the batch tuner times."""
    return [(r, c) for r in BATCH_ROWS for c in CAPS]


def pass_candidates(dtype, shape):
    """(tile rows, x chunk) pairs the drop-in tuner times for a grid: 0 (one wave of resident blocks) and
    the depths of equal splits of the marching axis, at least PASS_MIN_XCHUNK planes, each once."""
    nx = int(shape[0])
    xcs = [0]
    for k in PASS_SPLITS:
        v = math.ceil(nx / k)
        if v >= min(nx, PASS_MIN_XCHUNK) and v not in xcs:
            xcs.append(v)
    return [(r, xc) for r in PASS_ROWS[dtype] for xc in xcs]


def _cache_path():
    p = os.environ.get("SW_AUTOTUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "shearwave_amd", "autotune.json")


def _label(c):
    return "x".join(str(int(v)) for v in c)


class Tuner:
    """Launch shapes of the batched step (per device) and of a drop-in solve's two-step pass (per device and
    grid class), measured on first use and cached.

    ``timer(candidates)`` (batch) and ``pass_timer(solve, candidates)``
    (drop-in) return seconds per voxel-step for each candidate; the defaults
    time the device.
    """

    def __init__(self, path=None, timer=None, arch=None, pass_timer=None):
        self.path = path or _cache_path()
        self._timer = timer
        self._pass_timer = pass_timer
        self._arch = arch
        self._lock = threading.Lock()
        self._table = {}
        try:
            with open(self.path) as f:
                self._table = json.load(f)
        except (OSError, ValueError):
            self._table = {}

    @property
    def arch(self):
        if self._arch is None:
            import ctypes as C

            from ._abi import lib

            buf = C.create_string_buffer(64)
            cu = C.c_int()
            lib().sw_device_info(C.byref(cu), buf, 64)
            self._arch = f"{buf.value.decode()}:{cu.value}"
        return self._arch

    # ------------------------------------------------------------- batches
    def key(self):
        return f"{self.arch}|batch_shape"

    def batch_shape(self):
        """(tile rows, x-chunk cap) of the batched step on this device."""
        k = self.key()
        with self._lock:
            if k not in self._table:
                self._table[k] = self._tune_batch()
                self._save()
            e = self._table[k]
            return int(e["tile_y"]), int(e["cap"])

    def tile_y(self):
        return self.batch_shape()[0]

    def cap(self):
        return self.batch_shape()[1]

    def xchunk(self, shape):
        """x planes per work item for a solve of stored `shape`."""
        return balanced_chunk(shape[0], self.cap())

    def _tune_batch(self):
        cands = batch_candidates()
        times = (self._timer or _time_mix)(cands)
        best = min(times, key=times.get)
        return {"tile_y": int(best[0]), "cap": int(best[1]), "mix": f"{TUNE_MIX_SOLVES} grids, seed {TUNE_MIX_SEED}",
                "ps_per_voxel_step": {_label(c): round(v * 1e12, 3) for c, v in times.items()}}

    # -------------------------------------------------------------- drop-in
    def pass_key(self, dtype, viscous, shape):
        return f"{self.arch}|pass|{dtype}|{'viscous' if viscous else 'inviscid'}|{_label(shape)}"

    def pass_shape(self, solve):
        """(tile rows, x chunk) of the two-step pass for an ``engine.Solve`` (timed on its buffers once per
        grid class)."""
        k = self.pass_key(solve.dtype, solve.e is not None, solve.shape)
        with self._lock:
            if k not in self._table:
                cands = pass_candidates(solve.dtype, solve.shape)
                times = (self._pass_timer or _time_pass)(solve, cands)
                best = min(times, key=times.get)
                self._table[k] = {"tile_y": int(best[0]), "x_chunk": int(best[1]),
                                  "ps_per_voxel_step": {_label(c): round(v * 1e12, 3) for c, v in times.items()}}
                self._save()
            e = self._table[k]
            return int(e["tile_y"]), int(e["x_chunk"])

    def entry(self, key):
        """The cached record of a key (None if not tuned)."""
        return self._table.get(key)

    def _save(self):
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            tmp = f"{self.path}.{os.getpid()}.tmp"  # processes sharing the file each write their own
            with open(tmp, "w") as f:
                json.dump(self._table, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass  # a read-only home keeps the table in memory


def _time_mix(cands, n_passes=6, rounds=2):
    """Seconds per voxel-step of force-free batched two-step passes over the tuning mix, per (tile rows, cap).

    Every candidate is timed `rounds` times, interleaved, each after one
    untimed pass, and its best time kept. The passes' results do not depend
    on the tiling (a voxel's arithmetic is the same whichever item computes
    it), so the mix is advanced, not reset, between candidates.
    """
    import numpy as np
    import torch

    from .batch import Batch, SolveSpec, step_items

    h = (0.5e-3,) * 3
    n_steps = 2 * (rounds * len(cands) * (n_passes + 1) + 2)
    specs = [SolveSpec(shape=shape, h=h, dt=1e-6, rho=1000.0, mu=3000.0, eta=1.0, force=(None, None, None),
                       env=np.zeros(n_steps), n_steps=n_steps) for shape in tune_mix_shapes()]
    b = Batch(specs, axes="given", tile_y=DEFAULT_ROWS)  # the shapes are stored orders already
    vox = sum(int(np.prod(s.shape)) for s in b.solves)
    best = {}
    for _ in range(rounds):
        for rows, cap in cands:
            for s in b.solves:
                s.xchunk = balanced_chunk(s.shape[0], cap)
            b.tile_y = rows
            desc = b._desc_array(b.solves)
            items = step_items(b.solves, rows)
            items_t = torch.from_numpy(items.reshape(-1).copy()).to(desc.device)
            b._advance(desc, items_t, items.shape[0], 1)
            start = torch.cuda.Event(enable_timing=True)
            stop = torch.cuda.Event(enable_timing=True)
            start.record()
            b._advance(desc, items_t, items.shape[0], n_passes)
            stop.record()
            stop.synchronize()
            t = start.elapsed_time(stop) / 1e3 / (2 * n_passes) / vox
            best[(rows, cap)] = min(best.get((rows, cap), t), t)
    del b
    torch.cuda.empty_cache()
    return best


def _time_pass(solve, cands, rounds=1, target_s=1e-3):
    """Seconds per voxel-step of the two-step pass on `solve`'s buffers, per (tile rows, x chunk).

    After a few untimed passes (clocks up), each candidate runs one untimed
    pass, then enough timed passes for about `target_s` (at least two),
    `rounds` times interleaved; the best time is kept. The passes read the
    solve's state and write its next-step buffers only
    (``sw_fdtd_trial_passes``), so the solve is unchanged. About 25 ms at
    256^3 for the 18 float32 candidates.
    """
    import ctypes as C

    import torch

    from . import _abi
    from . import engine as E

    lib = _abi.lib()
    vox = int(solve.n)
    n_passes = max(2, min(50, int(target_s * 150e9 / (2 * vox))))  # at about 150 Gvoxel-steps/s
    st = E.stream_ptr()
    s0 = _abi.FdtdState.from_buffer_copy(solve.state)
    _abi.check(lib.sw_fdtd_trial_passes(C.byref(s0), 3, st))
    best = {}
    for _ in range(rounds):
        for rows, xc in cands:
            s = _abi.FdtdState.from_buffer_copy(solve.state)
            s.tile_y, s.x_chunk = int(rows), int(xc)
            _abi.check(lib.sw_fdtd_trial_passes(C.byref(s), 1, st))
            start = torch.cuda.Event(enable_timing=True)
            stop = torch.cuda.Event(enable_timing=True)
            start.record()
            _abi.check(lib.sw_fdtd_trial_passes(C.byref(s), n_passes, st))
            stop.record()
            stop.synchronize()
            t = start.elapsed_time(stop) / 1e3 / (2 * n_passes) / vox
            best[(rows, xc)] = min(best.get((rows, xc), t), t)
    return best


_DEFAULT = None


def default_tuner():
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Tuner()
    return _DEFAULT
