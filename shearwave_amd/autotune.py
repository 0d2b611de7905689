"""Autotuner for the batched time-step launch shape.

The batched step kernel marches x-chunks of a (8 x 64) tile of (y, z)
columns; the chunk depth trades the chunk's prologue and halo re-reads
(three planes per chunk for a two-step pass) and the per-item fixed cost
against the number of work items available to balance the persistent grid.
The tuner times the step kernel on a batch of copies of a grid class (in its
stored axis order) with every candidate chunk depth, keeps the fastest, and
caches the choice per device architecture in a JSON file
(``$SW_AUTOTUNE_CACHE`` or ``~/.cache/shearwave_amd/autotune.json``). The
tile's orientation (which grid axis is stored fastest, see
``batch.axis_order``) is fixed by the lanes it fills, and a batch with too
few items for the GPU shortens its chunks further (``batch.fill_xchunks``).
"""

from __future__ import annotations

import json
import math
import os
import threading

CANDIDATES = (48, 64, 96, 128, 192, 256)


def _pow2(v):
    return 1 << max(0, math.ceil(math.log2(max(int(v), 1))))


def grid_class(shape):
    """Bucket a grid by its x length (the march) and its (y, z) tiles per plane, both to powers of two."""
    nx, ny, nz = (int(s) for s in shape)
    return (_pow2(nx), _pow2(math.ceil(ny / 8) * math.ceil(nz / 64)))


def _cache_path():
    p = os.environ.get("SW_AUTOTUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "shearwave_amd", "autotune.json")


class Tuner:
    """Per-class x-chunk choice, measured on first use and cached."""

    def __init__(self, path=None, timer=None, arch=None):
        self.path = path or _cache_path()
        self._timer = timer
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

    def key(self, shape):
        return f"{self.arch}|{'x'.join(str(v) for v in grid_class(shape))}"

    def xchunk(self, shape):
        k = self.key(shape)
        with self._lock:
            if k not in self._table:
                self._table[k] = self._tune(shape)
                self._save()
            return min(int(self._table[k]["xchunk"]), int(shape[0]))

    def _save(self):
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            tmp = self.path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self._table, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass  # a read-only home keeps the table in memory

    def _tune(self, shape):
        timer = self._timer or _time_step_kernel
        cls = grid_class(shape)
        rep = (min(cls[0], 512), 8 * cls[1], 64)  # same x length and tiles per plane
        times = {}
        for xc in [c for c in CANDIDATES if c <= rep[0]] or [rep[0]]:  # a short grid: whole columns
            times[xc] = float(timer(rep, xc))
        best = min(times, key=times.get)
        return {"xchunk": best, "shape": list(rep), "ps_per_voxel_step": {str(k): round(v * 1e12, 3) for k, v in times.items()}}


def _time_step_kernel(shape, xchunk, n_passes=6):
    """Seconds per voxel-step of the batched two-step passes on copies of one `shape` solve.

    Enough copies are batched that the work list is about twice the
    persistent grid, so the chunk depth is judged as it runs in a full batch
    (halo and prologue cost against balance), not on a lone solve.
    """
    import numpy as np
    import torch

    from .batch import STEP_BY, Batch, SolveSpec, step_items

    h = (0.5e-3,) * 3
    per_copy = math.ceil(shape[0] / xchunk) * math.ceil(shape[1] / STEP_BY) * math.ceil(shape[2] / 64)
    copies = max(1, min(16, math.ceil(2048 / per_copy)))
    n_steps = 2 * (n_passes + 3)  # force-free: every pass is two steps
    specs = [SolveSpec(shape=shape, h=h, dt=1e-6, rho=1000.0, mu=3000.0, eta=1.0, force=(None, None, None),
                       env=np.zeros(n_steps), n_steps=n_steps, xchunk=xchunk) for _ in range(copies)]
    b = Batch(specs, axes="given")  # `shape` is already a stored axis order
    desc = b._desc_array(b.solves)
    items = step_items(b.solves)
    items_t = torch.from_numpy(items.reshape(-1).copy()).to(desc.device)
    b._advance(desc, items_t, items.shape[0], 3)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    b._advance(desc, items_t, items.shape[0], n_passes)
    stop.record()
    stop.synchronize()
    vox = copies * int(np.prod(shape))
    return start.elapsed_time(stop) / 1e3 / (2 * n_passes) / vox


_DEFAULT = None


def default_tuner():
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Tuner()
    return _DEFAULT
