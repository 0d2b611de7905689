"""Pass plans of the batched loop (host logic, CPU)."""

import numpy as np
import pytest

from shearwave_amd.batch import pass_plan


def _steps(plan):
    out = []
    for e in plan:
        t, two = int(e) >> 1, int(e) & 1
        out += [t] + ([t + 1] if two else [])
    return out


@pytest.mark.parametrize("n", [0, 1, 2, 7, 10])
@pytest.mark.parametrize("two_step", [True, False])
def test_plan_covers_every_step_once_in_order(n, two_step):
    plan = pass_plan(n, two_step)
    assert _steps(plan) == list(range(n))
    assert plan.dtype == np.int32
    assert len(plan) == ((n + 1) // 2 if two_step else n)


def test_plan_pairs_steps():
    assert [(int(e) >> 1, int(e) & 1) for e in pass_plan(5)] == [(0, 1), (2, 1), (4, 0)]
    assert all(int(e) & 1 == 0 for e in pass_plan(9, two_step=False))


def test_cg_xchunk_rule():
    """The batched CG item length: of 192, 96 and 48 planes, the one whose items (planes plus a 6-plane
    prologue each) finish first when one block per CU claims them longest first (host logic only)."""
    from shearwave_amd.batch import CG_SUB, CG_XCHUNKS, _claimed_makespan, cg_xchunk

    assert all(xc % CG_SUB == 0 for xc in CG_XCHUNKS)
    assert _claimed_makespan([5, 3, 3, 2, 2], 2) == 8  # (5, 3) and (3, 2, 2): longest first to the free block
    assert _claimed_makespan([4, 4], 8) == 4
    one = [(128, 128, 160)]  # 8 y tiles x 3 z tiles: 24 items at 192 planes, 72 at 48
    assert cg_xchunk(one, 256) == 48  # every item at once: the shortest items finish first
    assert cg_xchunk(one, 6) == 192  # 4 whole-axis items per block: the fewest prologues
    assert cg_xchunk(one, 9) == 96  # 24 whole-axis items leave 3 of 9 blocks idle in the last round
    assert cg_xchunk(one * 48, 256) == 96  # 1152 whole-axis items: 4.5 rounds of 256
    assert cg_xchunk([(128, 128, 128)] * 16, 256) == 192  # 256 whole-axis items, one per CU
    assert cg_xchunk([(8, 8, 8)], 256) == 192  # one item whatever the length: the longest on the tie
