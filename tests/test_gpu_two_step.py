"""The two-step march pass (two steps per read of the state) against single steps.

A pass evaluates each step with the same operations as the single-step
kernel (contraction off, fused multiply-adds written out), so float32 and
float64 results must be bit-identical to the single-step march run, including
on grids that leave partial tiles in y and z, with and without viscosity,
around outputs (snapshots, strain) and around forced steps. Parity of the
pass against the float64 oracle is covered in test_gpu_parity (kernel
"march2": float64 within 1e-10, float32 within 1e-5).
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from tests.test_gpu_parity import _problem, _run_gpu  # noqa: E402


def _random_state(shape, seed):
    rng = np.random.default_rng(seed)
    u0 = 1e-6 * rng.standard_normal(shape + (3,))
    v0 = 1e-3 * rng.standard_normal(shape + (3,))
    return u0, v0


def _solve(sw, shape, seed, kernel, viscous, tile_y=0, x_chunk=0, dtype="float32"):
    from shearwave_amd import engine as E

    rng = np.random.default_rng(seed + 1)
    rho = 1000.0 * (1 + rng.uniform(-0.01, 0.01, shape))
    mu = rng.uniform(2e3, 2e4, shape)
    eta = rng.uniform(0.5, 3.0, shape) if viscous else None
    h = 0.5e-3
    dt = 0.2 * h / np.sqrt(2e4 / 990.0 * 3)
    s = E.Solve(shape, (h, h, h), dt, rho, mu, eta, dtype=dtype, kernel=kernel, tile_y=tile_y,
                x_chunk=x_chunk, vel_damp=0.003)
    u0, v0 = _random_state(shape, seed)
    s.set_initial(u0, v0)
    return s


def _flat(res):
    out = []
    for x in res:
        out.extend(x.values() if isinstance(x, dict) else [x])
    return [np.asarray(v) for v in out]


def _state(s):
    return [t.clone() for t in s.u_components()] + [t.clone() for t in s.b_components()]


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("shape", [(40, 37, 70), (24, 9, 130), (33, 20, 64), (9, 8, 32)])
@pytest.mark.parametrize("viscous", [True, False])
@pytest.mark.parametrize("n_steps", [7, 10])
def test_two_step_pass_bit_identical(sw, shape, viscous, n_steps, dtype):
    ref = _solve(sw, shape, 5, "march", viscous, dtype=dtype)
    ref.run(n_steps)
    got = _solve(sw, shape, 5, "march2", viscous, dtype=dtype)
    got.run(n_steps)
    torch.cuda.synchronize()
    for a, b in zip(_state(got), _state(ref)):
        assert torch.equal(a, b), (a - b).abs().max().item()
    assert float(ref.u_components()[0].abs().max()) > 0


@pytest.mark.parametrize("tile_y,x_chunk", [(12, 0), (8, 5), (8, 17), (12, 3)])
def test_two_step_tile_shapes(sw, tile_y, x_chunk):
    shape = (37, 29, 140)
    ref = _solve(sw, shape, 6, "march", True)
    ref.run(8)
    got = _solve(sw, shape, 6, "march2", True, tile_y=tile_y, x_chunk=x_chunk)
    got.run(8)
    torch.cuda.synchronize()
    for a, b in zip(_state(got), _state(ref)):
        assert torch.equal(a, b)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_two_step_around_force_and_outputs(sw, dtype):
    """Passes over forced and ramped steps; snapshots every 3 and strain every 4 steps break the pairing."""
    shape = (18, 16, 70)
    p = _problem(shape, seed=21, n_steps=48)  # a multiple of the snapshot stride (solver.py:469-507)
    kw = dict(snapshot_stride=3, return_strain=True)
    o = {"sw_strain_stride": 4, "sw_dtype": dtype}
    ref = _run_gpu(sw, p, shape, {**o, "sw_kernel": "march"}, **kw)
    got = _run_gpu(sw, p, shape, {**o, "sw_kernel": "march2"}, **kw)
    auto = _run_gpu(sw, p, shape, o, **kw)
    assert len(ref) == 4  # u, v, snapshots, strain dict
    for a, b, c in zip(_flat(got), _flat(ref), _flat(auto)):
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(c, b)
