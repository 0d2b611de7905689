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
