"""The three-step march pass (three force-free steps per read of the state) against single steps.

Each step of a pass uses the single-step kernel's operations (contraction
off, fused multiply-adds written out), so float32 results must be
bit-identical to the single-step march run: on grids that leave partial
tiles in y and z, with and without viscosity, with x chunks of one plane and
up, and in runs where outputs and forced steps fall between passes (those
steps go to two-step passes and single steps). Float64 storage runs two-step
passes under the same kernel choice.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from tests.test_gpu_parity import _problem, _run_gpu, _run_oracle  # noqa: E402
from tests.test_gpu_two_step import _flat, _solve, _state  # noqa: E402


@pytest.mark.parametrize("shape", [(40, 37, 70), (24, 9, 130), (33, 20, 64), (9, 8, 32), (30, 15, 66)])
@pytest.mark.parametrize("viscous", [True, False])
@pytest.mark.parametrize("n_steps", [7, 8, 9])
def test_three_step_pass_bit_identical(sw, shape, viscous, n_steps):
    ref = _solve(sw, shape, 5, "march", viscous)
    ref.run(n_steps)
    got = _solve(sw, shape, 5, "march3", viscous)
    got.run(n_steps)
    torch.cuda.synchronize()
    for a, b in zip(_state(got), _state(ref)):
        assert torch.equal(a, b), (a - b).abs().max().item()
    assert float(ref.u_components()[0].abs().max()) > 0


@pytest.mark.parametrize("x_chunk", [1, 2, 3, 5, 17])
def test_three_step_x_chunks(sw, x_chunk):
    shape = (37, 29, 140)
    ref = _solve(sw, shape, 6, "march", True)
    ref.run(9)
    got = _solve(sw, shape, 6, "march3", True, x_chunk=x_chunk)
    got.run(9)
    torch.cuda.synchronize()
    for a, b in zip(_state(got), _state(ref)):
        assert torch.equal(a, b)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_three_step_around_force_and_outputs(sw, dtype):
    """Forced steps, snapshots every 3 and strain every 5 steps break the triples."""
    shape = (18, 16, 70)
    p = _problem(shape, seed=21, n_steps=60)  # a multiple of the snapshot stride (solver.py:469-507)
    kw = dict(snapshot_stride=3, return_strain=True)
    o = {"sw_strain_stride": 5, "sw_dtype": dtype}
    ref = _run_gpu(sw, p, shape, {**o, "sw_kernel": "march"}, **kw)
    got = _run_gpu(sw, p, shape, {**o, "sw_kernel": "march3"}, **kw)
    assert len(ref) == 4  # u, v, snapshots, strain dict
    for a, b in zip(_flat(got), _flat(ref)):
        np.testing.assert_array_equal(a, b)


def test_three_step_long_run_matches_oracle(sw):
    """A force-free stretch of 300 steps in three-step passes stays within 1e-5 of the float64 oracle."""
    shape = (30, 26, 70)
    p = _problem(shape, seed=22, n_steps=300)
    got = _run_gpu(sw, p, shape, {"sw_kernel": "march3"})
    want = _run_oracle(p)
    for a, b in zip(got, want):
        err = float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))
        assert err < 1e-5, err
