"""Pass plans of the batched loop (host logic, CPU)."""

import numpy as np
import pytest

from shearwave_amd.batch import pass_plan


def _steps(plan):
    out = []
    for e in plan:
        t, two = int(e) >> 1, int(e) & 1
        out += [(t, two)] + ([(t + 1, two)] if two else [])
    return out


@pytest.mark.parametrize("n", [0, 1, 2, 7, 10])
def test_plan_covers_every_step_once_in_order(n):
    env = np.zeros(n)
    env[: n // 3] = 1.0
    plan = pass_plan(env, n)
    steps = _steps(plan)
    assert [t for t, _ in steps] == list(range(n))
    for t, two in steps:
        if two:
            assert env[t] == 0.0  # forced steps are never paired
    assert plan.dtype == np.int32


def test_plan_pairs_force_free_steps():
    env = np.array([1.0, 0.5, 0, 0, 0, 0, 0, 1.0, 0, 0, 0])
    plan = pass_plan(env, len(env))
    assert [(int(e) >> 1, int(e) & 1) for e in plan] == [(0, 0), (1, 0), (2, 1), (4, 1), (6, 0), (7, 0), (8, 1), (10, 0)]
    assert len(pass_plan(None, 9)) == 5
    assert all(int(e) & 1 == 0 for e in pass_plan(None, 9, two_step=False))
