"""Batch layout rules on the CPU: the per-solve axis order, its inverse, and the x-chunk fill rule."""

import itertools
import math
from types import SimpleNamespace

import numpy as np
import torch

from shearwave_amd.batch import aos_to_given, axis_order, fill_xchunks, lane_fill, n_step_items, step_items, \
    vec_to_given


def test_lane_fill_and_axis_order():
    assert lane_fill((10, 64, 128)) == 1.0
    assert math.isclose(lane_fill((10, 20, 70)), (20 * 70) / (24 * 128))
    # the order that fills the (8 x 64) tiles best; ties keep the given order
    assert axis_order((128, 192, 200)) == (2, 0, 1)  # stored (200, 128, 192): every lane used
    assert lane_fill((200, 128, 192)) == 1.0
    assert axis_order((64, 64, 64)) == (0, 1, 2)
    assert axis_order((24, 20, 70)) == (1, 0, 2)  # stored (20, 24, 70)
    for shape in [(24, 20, 70), (80, 176, 256), (130, 72, 200), (9, 33, 20)]:
        p = axis_order(shape)
        best = max(lane_fill(tuple(shape[i] for i in q)) for q in itertools.permutations(range(3)))
        assert lane_fill(tuple(shape[i] for i in p)) == best


def _stored(v_given, perm):
    """The layout a solve stores: stored axis i is given axis perm[i], stored component i along it."""
    return torch.stack([v_given[perm[i]].permute(*perm) for i in range(3)])


def test_vector_and_aos_round_trips():
    rng = np.random.default_rng(0)
    g = torch.from_numpy(rng.standard_normal((3, 4, 5, 6)))
    for perm in itertools.permutations(range(3)):
        s = _stored(g, perm)
        assert tuple(s.shape[1:]) == tuple(g.shape[1 + i] for i in perm)
        assert torch.equal(vec_to_given(s, perm), g)
        aos_s = s.permute(1, 2, 3, 0)  # (n0', n1', n2', 3) of the stored layout
        assert torch.equal(aos_to_given(aos_s.contiguous(), perm), g.permute(1, 2, 3, 0))


def test_fill_xchunks_shortens_only_when_needed():
    def solve(shape, xc):
        return SimpleNamespace(shape=shape, xchunk=xc)

    many = [solve((128, 128, 128), 128) for _ in range(48)]
    fill_xchunks(many, 256)
    assert all(s.xchunk == 128 for s in many)  # 48 * 16 * 2 = 1536 items >= 512
    few = [solve((192, 64, 64), 192)]  # 8 items as whole columns
    fill_xchunks(few, 256)
    assert few[0].xchunk == 48 and n_step_items(few) == 32  # halved down to the floor
    two = [solve((192, 128, 128), 192), solve((64, 128, 128), 64)]
    fill_xchunks(two, 64)
    assert n_step_items(two) >= 128 and max(s.xchunk for s in two) <= 96


def test_step_items_orders():
    s = [SimpleNamespace(shape=(100, 16, 70), xchunk=64), SimpleNamespace(shape=(30, 8, 64), xchunk=64)]
    # spatial (the default): per solve, z tile fastest, then y tile, then x chunk
    sp = step_items(s)
    assert [tuple(r) for r in sp[:4]] == [(0, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 1, 1, 0)]
    assert tuple(sp[-1]) == (1, 0, 0, 0)
    items = step_items(s, order="longest")
    assert sorted(map(tuple, items)) == sorted(map(tuple, sp))
    assert items.shape == (2 * 2 * 2 + 1, 4)
    # the first item is a full 64-plane, 64-wide tile; the last a short, narrow one
    assert tuple(items[0][[0, 3]]) == (0, 0) and items[0][1] == 0
    assert tuple(items[-1][[0, 1, 3]]) == (0, 1, 1)
