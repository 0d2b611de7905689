"""Autotuner for the batched time-step launch shape.

The batched step kernel marches x-chunks of (8 x 64) tiles of (y, z)
columns; the chunk depth trades the chunk's prologue and halo re-reads
(three planes per chunk for a two-step pass) against the balance of the
persistent grid: a pass ends when its last item does, so a list of few long
items leaves blocks idle at the end of every pass. That balance is a
property of the whole batch, not of one grid: timed on copies of one grid at
a time, per-grid depths picked whole columns and made the 48-solve mix run at
110 Gvoxel-steps/s, against 128 with every solve cut under one cap.

So the tuner times the batched pass on a fixed synthetic mix of uneven grids
(`TUNE_MIX`) once per device, for every cap in `CAPS`, and caches the fastest
cap per device architecture in a JSON file (``$SW_AUTOTUNE_CACHE`` or
``~/.cache/shearwave_amd/autotune.json``). A batch then cuts each solve's
marching axis into the fewest equal chunks no longer than the cap
(`balanced_chunk`). The tile's orientation (which grid axis is stored
fastest, ``batch.axis_order``) is fixed by the lanes it fills, and a batch
with too few items for the GPU shortens its chunks further
(``batch.fill_xchunks``).
"""

from __future__ import annotations

import json
import math
import os
import threading

CAPS = (64, 80, 96, 128, 160, 192, 256)
DEFAULT_CAP = 96  # the measured best on the 48-solve mix; used without a tuner

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


def _cache_path():
    p = os.environ.get("SW_AUTOTUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "shearwave_amd", "autotune.json")


class Tuner:
    """The batched step's x-chunk cap per device architecture, measured on first use and cached."""

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

    def key(self):
        return f"{self.arch}|batch_cap"

    def cap(self):
        k = self.key()
        with self._lock:
            if k not in self._table:
                self._table[k] = self._tune()
                self._save()
            return int(self._table[k]["cap"])

    def xchunk(self, shape):
        """x planes per work item for a solve of stored `shape`."""
        return balanced_chunk(shape[0], self.cap())

    def _save(self):
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            tmp = self.path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self._table, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass  # a read-only home keeps the table in memory

    def _tune(self):
        timer = self._timer or _time_mix
        times = timer(CAPS)
        best = min(times, key=times.get)
        return {"cap": int(best), "mix": f"{TUNE_MIX_SOLVES} grids, seed {TUNE_MIX_SEED}",
                "ps_per_voxel_step": {str(k): round(v * 1e12, 3) for k, v in times.items()}}


def _time_mix(caps, n_passes=6, rounds=2):
    """Seconds per voxel-step of force-free batched two-step passes over the tuning mix, per cap.

    Every cap is timed `rounds` times, interleaved, and its best time kept.
    The passes' results do not depend on the chunking (a voxel's arithmetic
    is the same whichever item computes it), so the mix is advanced, not
    reset, between candidates.
    """
    import numpy as np
    import torch

    from .batch import Batch, SolveSpec, step_items

    h = (0.5e-3,) * 3
    n_steps = 2 * (rounds * len(caps) * (n_passes + 1) + 2)
    specs = [SolveSpec(shape=shape, h=h, dt=1e-6, rho=1000.0, mu=3000.0, eta=1.0, force=(None, None, None),
                       env=np.zeros(n_steps), n_steps=n_steps) for shape in tune_mix_shapes()]
    b = Batch(specs, axes="given")  # the shapes are stored orders already
    vox = sum(int(np.prod(s.shape)) for s in b.solves)
    best = {}
    warm = True
    for _ in range(rounds):
        for cap in caps:
            for s in b.solves:
                s.xchunk = balanced_chunk(s.shape[0], cap)
            desc = b._desc_array(b.solves)
            items = step_items(b.solves)
            items_t = torch.from_numpy(items.reshape(-1).copy()).to(desc.device)
            b._advance(desc, items_t, items.shape[0], 2 if warm else 1)
            warm = False
            start = torch.cuda.Event(enable_timing=True)
            stop = torch.cuda.Event(enable_timing=True)
            start.record()
            b._advance(desc, items_t, items.shape[0], n_passes)
            stop.record()
            stop.synchronize()
            t = start.elapsed_time(stop) / 1e3 / (2 * n_passes) / vox
            best[cap] = min(best.get(cap, t), t)
    del b
    torch.cuda.empty_cache()
    return best


_DEFAULT = None


def default_tuner():
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Tuner()
    return _DEFAULT
