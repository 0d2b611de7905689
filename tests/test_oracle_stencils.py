"""Pins for the oracle's stencil primitives and the projection operator.

Each check is chosen so that a plausible slip (a swapped spacing, a wrong
shift, a missing factor 2, a wrong wall row) fails it: closed forms exact on
low-degree polynomials, brute-force loops on tiny grids, and the Poisson
operator against a matrix assembled entry by entry from its closed form.
"""

import itertools

import numpy as np
import pytest

from oracle import solver as O

DX, DY, DZ = 0.7e-3, 1.1e-3, 0.9e-3


def _coords(shape):
    x = np.arange(shape[0])[:, None, None] * DX
    y = np.arange(shape[1])[None, :, None] * DY
    z = np.arange(shape[2])[None, None, :] * DZ
    return x, y, z


def test_laplacian_exact_on_quadratic():
    shape = (9, 7, 8)
    x, y, z = _coords(shape)
    f = 3.0 * x**2 - 5.0 * y**2 + 4.5 * z**2 + 7.0 * x * y - 4.0 * y * z + x - 2.0 * z + 0.3
    f = f * np.ones(shape)
    lap = O.laplacian_center(f, DX, DY, DZ)
    np.testing.assert_allclose(lap[1:-1, 1:-1, 1:-1], 2 * 3.0 - 2 * 5.0 + 2 * 4.5, rtol=1e-7)
    m = O.interior_mask(shape, np.float64)
    assert np.all(lap[m == 0] == 0.0)


def test_divergence_and_gradient_exact_on_linear():
    shape = (6, 9, 7)
    x, y, z = _coords(shape)
    one = np.ones(shape)
    fx = (2.0 * x + 3.0 * y) * one
    fy = (-4.0 * y + z) * one
    fz = (5.0 * z - x) * one
    div = O.divergence_center(fx, fy, fz, DX, DY, DZ)
    np.testing.assert_allclose(div[1:-1, 1:-1, 1:-1], 2.0 - 4.0 + 5.0, rtol=1e-9)
    phi = (1.5 * x - 2.5 * y + 0.5 * z + 4.0) * one
    gx, gy, gz = O.gradient_center(phi, DX, DY, DZ)
    for g, want in ((gx, 1.5), (gy, -2.5), (gz, 0.5)):
        np.testing.assert_allclose(g[1:-1, 1:-1, 1:-1], want, rtol=1e-9)
        m = O.interior_mask(shape, np.float64)
        assert np.all(g[m == 0] == 0.0)


def _brute_lap(f, h):
    n = f.shape
    out = np.zeros_like(f)
    for i, j, k in itertools.product(range(1, n[0] - 1), range(1, n[1] - 1), range(1, n[2] - 1)):
        s = (f[i + 1, j, k] - 2 * f[i, j, k] + f[i - 1, j, k]) / h[0] ** 2
        s += (f[i, j + 1, k] - 2 * f[i, j, k] + f[i, j - 1, k]) / h[1] ** 2
        s += (f[i, j, k + 1] - 2 * f[i, j, k] + f[i, j, k - 1]) / h[2] ** 2
        out[i, j, k] = s
    return out


def _brute_div(fx, fy, fz, h):
    n = fx.shape
    out = np.zeros_like(fx)
    for i, j, k in itertools.product(range(1, n[0] - 1), range(1, n[1] - 1), range(1, n[2] - 1)):
        out[i, j, k] = (
            (fx[i + 1, j, k] - fx[i - 1, j, k]) / (2 * h[0])
            + (fy[i, j + 1, k] - fy[i, j - 1, k]) / (2 * h[1])
            + (fz[i, j, k + 1] - fz[i, j, k - 1]) / (2 * h[2])
        )
    return out


def test_stencils_match_brute_force():
    rng = np.random.default_rng(11)
    shape = (5, 6, 7)
    h = (DX, DY, DZ)
    f = rng.standard_normal(shape)
    np.testing.assert_allclose(O.laplacian_center(f, *h), _brute_lap(f, h), rtol=1e-12, atol=1e-6)
    fx, fy, fz = (rng.standard_normal(shape) for _ in range(3))
    np.testing.assert_allclose(O.divergence_center(fx, fy, fz, *h), _brute_div(fx, fy, fz, h), rtol=1e-12, atol=1e-9)


def _assemble_poisson(shape, h):
    """Explicit matrix of -div_c grad_c on the interior, row by row from the closed form.

    Per axis at interior index i (interior is 1..n-2): deep rows read
    -(x[i+2] - 2x[i] + x[i-2]) / (4h^2); the gradient vanishes on the
    boundary layer, so the term G(i-1) drops at i = 1 and G(i+1) drops at
    i = n-2, leaving -(x[i+2] - x[i]) / (4h^2) and +(x[i] - x[i-2]) / (4h^2).
    Entries that land on the boundary layer multiply zeros and are skipped.
    """
    n = int(np.prod(shape))
    a = np.zeros((n, n))
    flat = np.arange(n).reshape(shape)
    for i, j, k in itertools.product(*(range(1, s - 1) for s in shape)):
        row = flat[i, j, k]
        idx = (i, j, k)
        for ax in range(3):
            c = 1.0 / (4.0 * h[ax] ** 2)
            nn = shape[ax]
            p = idx[ax]
            terms = []
            if p + 1 <= nn - 2:  # G(p+1) present: +(x[p+2] - x[p]) / (2h) ... times -1/(2h)
                terms += [(p + 2, -c), (p, +c)]
            if p - 1 >= 1:  # G(p-1) present: -(x[p] - x[p-2]) / (2h) ... times -1/(2h)
                terms += [(p, +c), (p - 2, -c)]
            for q, w in terms:
                if 1 <= q <= nn - 2:
                    t = list(idx)
                    t[ax] = q
                    a[row, flat[tuple(t)]] += w
    return a


def test_poisson_operator_matches_assembled_matrix():
    shape = (6, 7, 8)
    h = (DX, DY, DZ)
    rng = np.random.default_rng(3)
    m = O.interior_mask(shape, np.float64)
    x = rng.standard_normal(shape) * m
    a = _assemble_poisson(shape, h)
    want = (a @ x.ravel()).reshape(shape)
    got = O.poisson_operator(x, *h) * m
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-6 * np.abs(want).max())
    # symmetric positive definite on the interior
    inner = m.ravel() > 0
    ai = a[np.ix_(inner, inner)]
    np.testing.assert_allclose(ai, ai.T, rtol=0, atol=1e-9 * np.abs(ai).max())
    assert np.linalg.eigvalsh(ai).min() > 0


@pytest.mark.parametrize("shape", [(6, 6, 6), (7, 9, 8)])
def test_poisson_cg_matches_direct_solve(shape):
    h = (DX, DY, DZ)
    rng = np.random.default_rng(5)
    m = O.interior_mask(shape, np.float64)
    rhs = rng.standard_normal(shape) * m
    a = _assemble_poisson(shape, h)
    inner = m.ravel() > 0
    phi_ref = np.zeros(int(np.prod(shape)))
    phi_ref[inner] = np.linalg.solve(a[np.ix_(inner, inner)], rhs.ravel()[inner])
    info = {}
    phi = O.poisson_cg(rhs, *h, m, tol=1e-13, max_iters=5000, info=info)
    assert info["reason"] == "converged"
    np.testing.assert_allclose(phi.ravel(), phi_ref, rtol=0, atol=1e-9 * np.abs(phi_ref).max())


def test_projection_residual_identity():
    """|div_c(f - grad_c phi)| equals the CG residual |r_k| (r = -div of the output)."""
    shape = (14, 12, 16)
    rng = np.random.default_rng(7)
    taper = O.interior_mask(shape, np.float64)
    f = [rng.standard_normal(shape) * taper for _ in range(3)]
    info = {}
    out = O.project_vector_field(*f, DX, DY, DZ, 1e-6, 20, dtype=np.float64, info=info)
    d_in = np.linalg.norm(O.divergence_center(*f, DX, DY, DZ))
    d_out = np.linalg.norm(O.divergence_center(*out, DX, DY, DZ))
    assert info["iters"] == 20
    np.testing.assert_allclose(d_out / d_in, info["relres"], rtol=1e-6)
