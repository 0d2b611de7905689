"""The three-step march pass (three force-free steps per read of the state) against single steps.

Each step of a pass uses the single-step kernel's operations (contraction
off, fused multiply-adds written out), so the increments are bit-identical
to those of single steps, and the displacement is stored as the two-step pass
stores it, u_{n+3} = u_{n+1} + (d_{n+2} + d_{n+3}). The reference runs single
steps with that rule applied at the end of each pass, so the comparison is
exact: on grids that leave partial tiles in y and z, with and without
viscosity, with x chunks of one plane and up, and in runs where outputs and
forced steps fall between passes (those steps go to two-step passes and
single steps). Float64 storage runs two-step passes under the same kernel
choice.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from tests.test_gpu_parity import _problem, _run_gpu, _run_oracle  # noqa: E402
from tests.test_gpu_two_step import _assert_equal_states, _plan, _run_rule, _solve  # noqa: E402
from tests.test_gpu_two_step import run_with_outputs_against_rule  # noqa: E402


@pytest.mark.parametrize("shape", [(40, 37, 70), (24, 9, 130), (33, 20, 64), (9, 8, 32), (30, 15, 66)])
@pytest.mark.parametrize("viscous", [True, False])
@pytest.mark.parametrize("n_steps", [7, 8, 9])
def test_three_step_pass_rule(sw, shape, viscous, n_steps):
    ref = _solve(sw, shape, 5, "march", viscous)
    plan = _plan(n_steps, three=True)
    assert plan[0] == 3
    _run_rule(ref, plan)
    got = _solve(sw, shape, 5, "march3", viscous)
    got.run(n_steps)
    torch.cuda.synchronize()
    _assert_equal_states(got, ref)
    assert float(ref.u_components()[0].abs().max()) > 0


@pytest.mark.parametrize("x_chunk", [1, 2, 3, 5, 17])
def test_three_step_x_chunks(sw, x_chunk):
    shape = (37, 29, 140)
    ref = _solve(sw, shape, 6, "march", True)
    _run_rule(ref, _plan(9, three=True))
    got = _solve(sw, shape, 6, "march3", True, x_chunk=x_chunk)
    got.run(9)
    torch.cuda.synchronize()
    _assert_equal_states(got, ref)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_three_step_around_force_and_outputs(sw, dtype):
    """Forced steps, snapshots every 3 and strain every 5 steps break the triples."""
    run_with_outputs_against_rule(sw, "march3", dtype, 3, 5, three=True)


def test_three_step_long_run_matches_oracle(sw):
    """A force-free stretch of 300 steps in three-step passes stays within 1e-5 of the float64 oracle."""
    shape = (30, 26, 70)
    p = _problem(shape, seed=22, n_steps=300)
    got = _run_gpu(sw, p, shape, {"sw_kernel": "march3"})
    want = _run_oracle(p)
    for a, b in zip(got, want):
        err = float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))
        assert err < 1e-5, err
