"""Pins of the static-solve oracle (oracle/static_solve.py).

* NumPy's float32 summation order: the restatement and the pairwise plan
  against ``np.sum`` itself, bit for bit, over sizes that cross every branch
  (leaf sizes, the 128 split, the 8192 buffer).
* The operator against an explicit sparse matrix built from its closed form.
* The solve: the reference's manufactured-solution test
  (test_projection_static.py:86-105), a sparse direct solve of the same
  discrete system, and a non-zero result (the sign convention of
  solver.py:581-588).
"""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import static_solve as S


def _bits(x):
    return np.float32(x).view(np.uint32)


SIZES = list(range(1, 20)) + [63, 64, 127, 128, 129, 136, 255, 256, 257, 1000, 4095, 4096, 4097, 8191, 8192, 8193,
                              16384, 20000, 24577]
SHAPES = [(7, 9, 11), (33, 40, 50), (64, 64, 64), (72, 48, 72), (3, 300, 301), (20, 20, 20), (24, 24, 24)]


@pytest.mark.parametrize("n", SIZES)
def test_numpy_sum_order_1d(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n).astype(np.float32)
    for a in (x, x * x):
        assert _bits(S.numpy_sum_f32(a)) == _bits(np.sum(a))


@pytest.mark.parametrize("shape", SHAPES)
def test_numpy_sum_order_3d(shape):
    rng = np.random.default_rng(sum(shape))
    x = (rng.standard_normal(shape) * rng.uniform(0.1, 1e3, shape)).astype(np.float32)
    y = rng.standard_normal(shape).astype(np.float32)
    for a in (x, x * x, x * y):
        assert _bits(S.numpy_sum_f32(a)) == _bits(np.sum(a))


def test_sum_order_discriminates():
    """A plain running sum and a single pairwise pass both differ from np.sum here."""
    x = np.random.default_rng(7).standard_normal(64**3).astype(np.float32)
    running = np.float32(0.0)
    for v in x[:20000]:
        running = np.float32(running + v)
    assert _bits(running) != _bits(np.sum(x[:20000]))
    assert _bits(S.pairwise_sum_f32(x)) != _bits(np.sum(x))
    assert _bits(S.numpy_sum_f32(x)) == _bits(np.sum(x))


@pytest.mark.parametrize("m", list(range(1, 20)) + [100, 128, 129, 200, 1000, 1031, 4097, 5000, 6143, 8191, 8192])
def test_plan_is_the_pairwise_tree(m):
    leaves, levels = S.pairwise_plan(m)
    # leaves tile [0, m) in order, start on multiples of 8 and hold at most 128
    pos = 0
    for s, n in leaves:
        assert s == pos and s % 8 == 0 and 1 <= n <= 128
        pos += n
    assert pos == m
    assert len(leaves) <= 65 and len(levels) <= 7
    assert sum(len(lv) for lv in levels) == len(leaves) - 1
    seen = set(range(len(leaves)))
    nid = len(leaves)
    for lv in levels:
        for a, b in lv:
            assert a in seen and b in seen
        for _ in lv:
            seen.add(nid)
            nid += 1
    x = (np.random.default_rng(m).standard_normal(m) * 10).astype(np.float32)
    assert _bits(S.eval_plan(x, (leaves, levels))) == _bits(np.sum(x))


def _operator_matrix(shape, mu, h):
    """-mu lap as a sparse matrix over all voxels (rows of boundary voxels are zero)."""
    nx, ny, nz = shape
    idx = np.arange(nx * ny * nz).reshape(shape)
    rows, cols, vals = [], [], []
    for i, j, k in np.ndindex(*shape):
        if not (0 < i < nx - 1 and 0 < j < ny - 1 and 0 < k < nz - 1):
            continue
        c = idx[i, j, k]
        diag = 0.0
        for (di, dj, dk), hh in (((1, 0, 0), h[0]), ((0, 1, 0), h[1]), ((0, 0, 1), h[2])):
            w = -mu[i, j, k] / hh**2
            for s in (1, -1):
                rows.append(c)
                cols.append(idx[i + s * di, j + s * dj, k + s * dk])
                vals.append(w)
            diag -= 2 * w
        rows.append(c)
        cols.append(c)
        vals.append(diag)
    n = nx * ny * nz
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def test_operator_matches_matrix():
    shape, h = (5, 6, 7), (0.3, 0.2, 0.25)
    rng = np.random.default_rng(1)
    mu = rng.uniform(500, 3000, shape)
    x = rng.standard_normal(shape)
    got = S.apply_static_operator(x, mu, *h)
    want = (_operator_matrix(shape, mu, h) @ x.ravel()).reshape(shape)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_reference_manufactured_solution():
    """test_projection_static.py:86-105, restated on the oracle."""
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
    u = S.elastic_solve_static(bz, d, d, d, mu, tol=1e-8, max_iters=5000)
    assert np.max(np.abs(u)) > 0.0
    rel = np.max(np.abs((u - u_true) * mask)) / np.max(np.abs(u_true * mask))
    assert rel < 1e-3


def test_matches_direct_solve():
    shape, h = (10, 9, 11), (2e-4, 2.5e-4, 1.8e-4)
    rng = np.random.default_rng(3)
    mu = rng.uniform(1000, 9000, shape).astype(np.float32)
    bz = rng.standard_normal(shape).astype(np.float32) * 50
    info = {}
    u = S.elastic_solve_static(bz, *h, mu, tol=1e-7, max_iters=2000, info=info)
    assert info["reason"] == "converged" and u.dtype == np.float32
    A = _operator_matrix(shape, mu.astype(np.float64), h)
    inner = np.zeros(shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    sel = np.flatnonzero(inner.ravel())
    want = np.zeros(np.prod(shape))
    want[sel] = spla.spsolve(A[sel][:, sel].tocsc(), bz.astype(np.float64).ravel()[sel])
    want = want.reshape(shape)
    assert np.all(u[~inner] == 0)
    assert np.abs(u - want).max() < 1e-5 * np.abs(want).max()


def test_history_and_stop_rules():
    shape = (12, 10, 14)
    bz = np.random.default_rng(4).standard_normal(shape).astype(np.float32)
    hist, info = [], {}
    S.elastic_solve_static(bz, 1e-3, 1e-3, 1e-3, 2500.0, tol=1e-5, max_iters=500, history=hist, info=info)
    assert info["reason"] == "converged" and info["iters"] == len(hist)
    assert hist[-1][1] < np.float32(1e-5) <= min(r for _, r in hist[:-1])
    hist2, info2 = [], {}
    S.elastic_solve_static(bz, 1e-3, 1e-3, 1e-3, 2500.0, tol=1e-5, max_iters=7, history=hist2, info=info2)
    assert info2 == {"iters": 7, "relres": float(hist2[-1][1]), "reason": "max_iters"}
    assert hist2 == hist[:7]
    # zero force: r0 = 0, p.Ap = 0 <= 0 stops before any update
    info3 = {}
    u = S.elastic_solve_static(np.zeros(shape, np.float32), 1e-3, 1e-3, 1e-3, 2500.0, info=info3)
    assert info3["reason"] == "denom" and info3["iters"] == 0 and not u.any()
