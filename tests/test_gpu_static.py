"""GPU parity of elastic_solve_static (solver.py:569-625) against the oracle.

Gate: bit-identical. The kernels run the reference's float32 operations in
its order and sum every dot product in NumPy's float32 order, so u, the
iteration count, the stop reason and every (r.r, rel) of the history must
equal the oracle's exactly. Shapes cover a grid smaller than one 8192-element
reduction buffer, whole buffers plus a partial one, whole buffers only, small
and large z extents, grids without an interior, and the reference's
quasi-static validation case (validate_quasistatic.py:24-33, 83-85).
"""

import numpy as np
import pytest

from oracle import static_solve as S

pytestmark = pytest.mark.gpu


def _force(shape, seed):
    rng = np.random.default_rng(seed)
    taper = np.ones(shape)
    for ax, n in enumerate(shape):
        w = np.sin(np.pi * (np.arange(n) + 0.5) / n)
        taper = taper * w.reshape([-1 if a == ax else 1 for a in range(3)])
    return (rng.standard_normal(shape) * taper * 0.02).astype(np.float32)


def _run_both(sw, bz, h, mu, tol, max_iters):
    hist_o, info_o, info_g = [], {}, {}
    want = S.elastic_solve_static(bz, *h, mu, tol=tol, max_iters=max_iters, history=hist_o, info=info_o)
    got = sw.elastic_solve_static(bz, *h, mu, tol=tol, max_iters=max_iters, info=info_g)
    return got, want, info_g, info_o, hist_o


CASES = [
    ((20, 20, 20), 1e-6, 400),      # 8000 voxels: one partial buffer
    ((24, 24, 24), 1e-6, 400),      # one whole buffer and a partial one
    ((32, 32, 32), 1e-6, 400),      # four whole buffers
    ((5, 300, 7), 1e-5, 300),       # z extent far below the 512-voxel stride
    ((4, 5, 700), 1e-5, 300),       # z extent above it
    ((9, 17, 33), 1e-7, 60),        # stops at max_iters
]


@pytest.mark.parametrize("shape,tol,max_iters", CASES)
@pytest.mark.parametrize("field_mu", [False, True])
def test_bit_identical_to_oracle(sw, shape, tol, max_iters, field_mu):
    bz = _force(shape, sum(shape))
    h = (2e-4, 2.5e-4, 1.5e-4)
    mu = np.random.default_rng(1).uniform(2000, 6000, shape).astype(np.float32) if field_mu else 4000.0
    got, want, info_g, info_o, _ = _run_both(sw, bz, h, mu, tol, max_iters)
    assert got.dtype == np.float32 and got.shape == shape
    assert info_g == info_o
    assert info_o["iters"] > 1
    np.testing.assert_array_equal(got, want)


def test_validation_case(sw):
    """The reference's quasi-static check: 72x48x72, mu = rho cs^2, tol 1e-6, 1500 iterations."""
    shape, d = (72, 48, 72), 2.0e-4
    xg, yg, zg = np.meshgrid(*(np.arange(n) - n // 2 for n in shape), indexing="ij")
    G = np.exp(-0.5 * ((xg / 8.0) ** 2 + (yg / 8.0) ** 2 + (zg / 6.0) ** 2)).astype(np.float32)
    G /= np.max(G)
    bz = (0.02 * G * _force(shape, 0).astype(np.float64) / 0.02 + 0.02 * G).astype(np.float32)
    mu = 1000.0 * 2.0**2
    got, want, info_g, info_o, _ = _run_both(sw, bz, (d, d, d), mu, 1e-6, 1500)
    assert info_g == info_o and info_o["reason"] == "converged"
    np.testing.assert_array_equal(got, want)


def test_reference_manufactured_solution(sw):
    """test_projection_static.py:86-105 on the GPU path."""
    n = 32
    d = 1.0 / n
    ax = np.linspace(0.0, 1.0, n, dtype=np.float32)
    x, y, z = ax[:, None, None], ax[None, :, None], ax[None, None, :]
    mask = np.zeros((n, n, n), np.float32)
    mask[1:-1, 1:-1, 1:-1] = 1.0
    mu = 1000.0
    u_true = (np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)).astype(np.float32) * np.ones((n, n, n),
                                                                                                   np.float32)
    lap = np.zeros_like(u_true)
    c = u_true[1:-1, 1:-1, 1:-1]
    lap[1:-1, 1:-1, 1:-1] = (
        (u_true[2:, 1:-1, 1:-1] - 2 * c + u_true[:-2, 1:-1, 1:-1]) / d**2
        + (u_true[1:-1, 2:, 1:-1] - 2 * c + u_true[1:-1, :-2, 1:-1]) / d**2
        + (u_true[1:-1, 1:-1, 2:] - 2 * c + u_true[1:-1, 1:-1, :-2]) / d**2
    )
    bz = (-mu * lap).astype(np.float32)
    u = sw.elastic_solve_static(bz, d, d, d, mu, tol=1e-8, max_iters=5000)
    assert np.max(np.abs(u)) > 0.0
    rel = np.max(np.abs((u - u_true) * mask)) / np.max(np.abs(u_true * mask))
    assert rel < 1e-3
    np.testing.assert_array_equal(u, S.elastic_solve_static(bz, d, d, d, mu, tol=1e-8, max_iters=5000))


def test_history_and_verbose(sw, capsys):
    shape = (12, 10, 14)
    bz = _force(shape, 4)
    hist = []
    S.elastic_solve_static(bz, 1e-3, 1e-3, 1e-3, 2500.0, tol=1e-5, max_iters=500, history=hist)
    sw.elastic_solve_static(bz, 1e-3, 1e-3, 1e-3, 2500.0, tol=1e-5, max_iters=500, verbose=True)
    lines = capsys.readouterr().out.strip().splitlines()
    want = [
        f"CG (GPU): it={it + 1}/500, |r|={np.sqrt(rs):.3e}, rel={rel:.3e}"
        for it, (rs, rel) in enumerate(hist)
        if it == 0 or (it + 1) % 25 == 0 or rel < 1e-5
    ]
    assert lines == want and len(want) >= 2


@pytest.mark.parametrize("max_iters", [0, 1, 7])
def test_iteration_caps(sw, max_iters):
    bz = _force((14, 15, 16), 5)
    got, want, info_g, info_o, _ = _run_both(sw, bz, (1e-3, 1e-3, 1e-3), 2500.0, 1e-6, max_iters)
    assert info_g == info_o
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("shape", [(1, 10, 10), (3, 3, 3), (2, 40, 40)])
def test_degenerate_grids(sw, shape):
    bz = _force(shape, 6) + 1.0
    got, want, info_g, info_o, _ = _run_both(sw, bz, (1e-3, 1e-3, 1e-3), 2500.0, 1e-6, 50)
    assert info_g == info_o
    np.testing.assert_array_equal(got, want)


def test_zero_force_and_float64_tol(sw):
    shape = (10, 11, 12)
    got, want, info_g, info_o, _ = _run_both(sw, np.zeros(shape, np.float32), (1e-3,) * 3, 2500.0, 1e-6, 50)
    assert info_g == info_o == {"iters": 0, "relres": None, "reason": "denom"} and not got.any()
    bz = _force(shape, 7)
    got, want, info_g, info_o, _ = _run_both(sw, bz, (1e-3,) * 3, 2500.0, np.float64(3e-6), 500)
    assert info_g == info_o
    np.testing.assert_array_equal(got, want)


def test_device_tensors(sw):
    import torch

    shape = (16, 18, 20)
    bz = _force(shape, 8)
    mu = np.random.default_rng(2).uniform(2000, 6000, shape)
    tb, tm = torch.from_numpy(bz).cuda(), torch.from_numpy(mu).cuda()  # float64 mu, cast to float32
    u = sw.elastic_solve_static(tb, 1e-3, 1e-3, 1e-3, tm, tol=1e-6, max_iters=300)
    assert isinstance(u, torch.Tensor) and u.is_cuda and u.dtype == torch.float32
    np.testing.assert_array_equal(u.cpu().numpy(),
                                  S.elastic_solve_static(bz, 1e-3, 1e-3, 1e-3, mu, tol=1e-6, max_iters=300))
