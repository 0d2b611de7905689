"""Pins for the oracle's time loop (restating solver.py:27-243).

* Dirichlet sine eigenmodes of the 7-point Laplacian evolve by a scalar
  two-term recurrence whose characteristic roots are known in closed form;
  the oracle must follow them to round-off (elastic, Kelvin-Voigt, damped,
  with and without ``initial_v``).
* A brute-force loop implementation with heterogeneous coefficients.
* The reference's divergence-free invariant (tests/test_divergence_free.py).
* Argument checking and error messages (solver.py:88-171).
"""

import itertools
import math

import numpy as np
import pytest

from oracle import solver as O


def _mode(shape, k):
    idx = [np.arange(n) for n in shape]
    s = [np.sin(math.pi * kk * i / (n - 1)) for kk, i, n in zip(k, idx, shape)]
    return s[0][:, None, None] * s[1][None, :, None] * s[2][None, None, :]


def _eig(shape, k, h):
    return sum(4.0 / hh**2 * math.sin(math.pi * kk / (2 * (n - 1))) ** 2 for kk, n, hh in zip(k, shape, h))


def _closed_form(n_steps, a, b, g, a_prev):
    """A_n of A_{n+1} = (1 + g(1-a-b)) A_n - g(1-b) A_{n-1}, A_0 = 1, A_{-1} = a_prev."""
    tr = 1.0 + g * (1.0 - a - b)
    det = g * (1.0 - b)
    disc = complex(tr * tr - 4.0 * det)
    z1 = (tr + disc**0.5) / 2.0
    z2 = (tr - disc**0.5) / 2.0
    # A_0 = c1 + c2 = 1 ; A_{-1} = c1/z1 + c2/z2 = a_prev
    m = np.array([[1.0, 1.0], [1.0 / z1, 1.0 / z2]], dtype=complex)
    c1, c2 = np.linalg.solve(m, np.array([1.0, a_prev], dtype=complex))
    an = c1 * z1**n_steps + c2 * z2**n_steps
    an1 = c1 * z1 ** (n_steps - 1) + c2 * z2 ** (n_steps - 1)
    return an.real, an1.real


@pytest.mark.parametrize(
    "eta,vel_damp,with_v0,k",
    [
        (0.0, 0.0, True, (1, 1, 1)),
        (0.8, 0.0, True, (2, 1, 3)),
        (0.5, 0.02, True, (1, 3, 2)),
        (0.0, 0.0, False, (1, 2, 1)),
        (1.2, 0.05, False, (3, 2, 2)),
    ],
)
def test_eigenmode_follows_closed_form(eta, vel_damp, with_v0, k):
    shape = (12, 10, 14)
    h = (0.5e-3, 0.6e-3, 0.45e-3)
    rho, mu = 1100.0, 3500.0
    cs = math.sqrt(mu / rho)
    dt = 0.2 * min(h) / (cs * math.sqrt(3.0))
    if eta:
        dt = min(dt, 0.4 * rho * min(h) ** 2 / (6 * eta))
    n_steps = 37
    comp = 1
    mode = _mode(shape, k)
    u0 = np.zeros(shape + (3,))
    u0[..., comp] = mode
    opts = {"eta": eta, "vel_damp": vel_damp, "initial_u": u0, "env_t": np.zeros(n_steps)}
    if with_v0:
        opts["initial_v"] = np.zeros_like(u0)
    zero = np.zeros(shape)
    u, v = O.shear_fdtd(zero, None, None, rho, mu, *h, dt, n_steps, opts=opts, dtype=np.float64)
    lam = _eig(shape, k, h)
    a = dt * dt * (mu / rho) * lam
    b = dt * (eta / rho) * lam
    an, an1 = _closed_form(n_steps, a, b, 1.0 - vel_damp, 1.0 if with_v0 else 0.0)
    np.testing.assert_allclose(u[..., comp], an * mode, rtol=0, atol=1e-11)
    np.testing.assert_allclose(v[..., comp], (an - an1) / dt * mode, rtol=0, atol=1e-11 / dt)
    assert np.all(u[..., [0, 2]] == 0.0)


def _brute_step_loop(u0, f0, env, m, e, inv_rho, h, dt, damp, n_steps):
    """Pure-Python loops over voxels: the reference scheme written out per point."""
    shape = u0.shape[:3]
    up = u0.copy()
    uc = u0.copy()
    inner = list(itertools.product(*(range(1, s - 1) for s in shape)))
    for it in range(n_steps):
        un = np.zeros_like(uc)
        for c in range(3):
            for i, j, k in inner:

                def lap(a):
                    return (
                        (a(i + 1, j, k) - 2 * a(i, j, k) + a(i - 1, j, k)) / h[0] ** 2
                        + (a(i, j + 1, k) - 2 * a(i, j, k) + a(i, j - 1, k)) / h[1] ** 2
                        + (a(i, j, k + 1) - 2 * a(i, j, k) + a(i, j, k - 1)) / h[2] ** 2
                    )

                lu = lap(lambda p, q, r: uc[p, q, r, c])
                lv = lap(lambda p, q, r: (uc[p, q, r, c] - up[p, q, r, c]) / dt)
                acc = m[i, j, k] * lu + e[i, j, k] * lv + f0[c][i, j, k] * env[it] * inv_rho[i, j, k]
                raw = 2 * uc[i, j, k, c] - up[i, j, k, c] + dt * dt * acc
                un[i, j, k, c] = uc[i, j, k, c] + (1 - damp) * (raw - uc[i, j, k, c])
        up, uc = uc, un
    return uc, (uc - up) / dt


def test_heterogeneous_loop_matches_brute_force():
    rng = np.random.default_rng(21)
    shape = (5, 6, 7)
    h = (0.5e-3, 0.4e-3, 0.6e-3)
    rho = 1000.0 * (1 + 0.1 * rng.uniform(-1, 1, shape))
    mu = 3000.0 * (1 + 0.5 * rng.uniform(0, 1, shape))
    eta = rng.uniform(0.5, 2.0, shape)
    dt = 2e-6
    n_steps = 4
    m = O.interior_mask(shape, np.float64)
    # a force that is already discretely divergence-free passes the projection unchanged
    f0 = O.project_vector_field(*(rng.standard_normal(shape) for _ in range(3)), *h, 1e-13, 5000, dtype=np.float64)
    env = np.array([1.0, 0.5, -0.25, 0.0])
    u0 = rng.standard_normal(shape + (3,)) * 1e-6 * m[..., None]
    opts = {"eta": eta, "vel_damp": 0.03, "initial_u": u0, "initial_v": np.zeros_like(u0), "env_t": env,
            "poisson_tol": 1e-13, "poisson_max_iters": 5000}
    u, v = O.shear_fdtd(*f0, rho, mu, *h, dt, n_steps, opts=opts, dtype=np.float64)
    ub, vb = _brute_step_loop(u0, f0, env, mu / rho, eta / rho, 1.0 / rho, h, dt, 0.03, n_steps)
    scale = np.abs(ub).max()
    np.testing.assert_allclose(u, ub, rtol=0, atol=1e-9 * scale)
    np.testing.assert_allclose(v, vb, rtol=0, atol=1e-7 * np.abs(vb).max())


# --- the reference's divergence-free invariant (tests/test_divergence_free.py:72-111) ---


def _div_curl_ratio(u, d):
    j = [np.gradient(u[..., i].astype(np.float64), d, d, d, edge_order=1) for i in range(3)]
    div = j[0][0] + j[1][1] + j[2][2]
    curl = np.sqrt((j[2][1] - j[1][2]) ** 2 + (j[0][2] - j[2][0]) ** 2 + (j[1][0] - j[0][1]) ** 2)
    n = u.shape[0]
    q = n // 4
    sl = (slice(q, n - q),) * 3
    return float(np.sqrt(np.mean(div[sl] ** 2)) / np.sqrt(np.mean(curl[sl] ** 2)))


def _push_setup(n=32, d=1.0e-3, mu=4580.0, rho=1000.0):
    ax = (np.arange(n) - (n - 1) / 2) * d
    x, y, z = np.meshgrid(ax, ax, ax, indexing="ij")
    bx = (300.0 * np.exp(-(x**2 + y**2 + z**2) / (2 * (2.5 * d) ** 2))).astype(np.float32)
    cs = math.sqrt(mu / rho)
    dt = 0.1 * d / (cs * math.sqrt(3.0))
    return bx, mu, rho, d, dt, d / (cs * dt)


def _ratios(n_steps, opts, dtype, n=32, every=20):
    bx, mu, rho, d, dt, _ = _push_setup(n)
    env = np.zeros(n_steps, np.float32)
    env[: max(1, n_steps // 3)] = 1.0
    out = []

    def cb(it, u):
        if it % every == every - 1:
            out.append(_div_curl_ratio(u, d))

    O.shear_fdtd(bx, None, None, rho, mu, d, d, d, dt, n_steps, opts={**opts, "env_t": env}, callback=cb, dtype=dtype)
    return out


@pytest.mark.parametrize("damped", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_push_stays_solenoidal_before_wall_arrival(damped, dtype):
    spv = _push_setup()[-1]
    n_steps = int(0.5 * 8 * spv)
    r = _ratios(n_steps, {"vel_damp": 0.02 if damped else 0.0, "eta": 0.5 if damped else 0.0}, dtype)
    assert r and max(r) < 1e-4


@pytest.mark.slow
def test_reproject_every_bounds_divergence():
    spv = _push_setup()[-1]
    n_steps = int(3.0 * 8 * spv)
    r = _ratios(n_steps, {"vel_damp": 0.02, "eta": 0.5, "reproject_every": 20}, np.float32)
    assert max(r) < 1e-4


def test_wall_injects_divergence_without_reprojection():
    spv = _push_setup()[-1]
    n_steps = int(1.5 * 8 * spv)
    r = _ratios(n_steps, {"vel_damp": 0.02, "eta": 0.5}, np.float32)
    assert max(r) > 1e-2


# --- argument checking (solver.py:88-171, 1060-1079) ---


def test_errors():
    b = np.zeros((6, 6, 6), np.float32)
    with pytest.raises(ValueError, match="At least one body-force"):
        O.shear_fdtd(None, None, None, 1000.0, 3000.0, 1e-3, 1e-3, 1e-3, 1e-6, 3)
    with pytest.raises(ValueError, match="reproject_every"):
        O.shear_fdtd(b, None, None, 1000.0, 3000.0, 1e-3, 1e-3, 1e-3, 1e-6, 3, opts={"reproject_every": -1})
    with pytest.raises(ValueError, match=r"env_t length \(2\) must equal n_steps \(3\)"):
        O.shear_fdtd(b, None, None, 1000.0, 3000.0, 1e-3, 1e-3, 1e-3, 1e-6, 3, opts={"env_t": [1, 1]})
    with pytest.raises(ValueError, match="Material field has incompatible shape"):
        O.shear_fdtd(b, None, None, np.ones((5, 6, 6)), 3000.0, 1e-3, 1e-3, 1e-3, 1e-6, 3, opts={"env_t": [1, 1, 1]})
    with pytest.raises(ValueError, match="Body-force component has incompatible shape"):
        O.shear_fdtd(b, np.zeros((6, 6, 5)), None, 1000.0, 3000.0, 1e-3, 1e-3, 1e-3, 1e-6, 3)
    with pytest.raises(ValueError, match=r"initial_u must match shape"):
        O.shear_fdtd(b, None, None, 1e3, 3e3, 1e-3, 1e-3, 1e-3, 1e-6, 3, opts={"env_t": [1, 1, 1], "initial_u": np.zeros((6, 6, 6))})
    with pytest.raises(ValueError, match=r"initial_v must match shape"):
        O.shear_fdtd(b, None, None, 1e3, 3e3, 1e-3, 1e-3, 1e-3, 1e-6, 3, opts={"env_t": [1, 1, 1], "initial_v": np.zeros((6, 6, 6))})


def test_snapshot_semantics_match_callback():
    rng = np.random.default_rng(2)
    shape = (8, 7, 9)
    b = rng.standard_normal(shape).astype(np.float32)
    env = np.sin(np.arange(12) * 0.5).astype(np.float32)
    opts = {"env_t": env, "eta": 0.3}
    u, v, snaps = O.shear_fdtd_snapshots(b, None, None, 1000.0, 3000.0, 1e-3, 1e-3, 1e-3, 2e-6, 12, opts, 4)
    assert snaps.shape == (3,) + shape + (3,)
    seen = {}

    def cb(it, uu):
        seen[it] = uu.copy()

    O.shear_fdtd(b, None, None, 1000.0, 3000.0, 1e-3, 1e-3, 1e-3, 2e-6, 12, opts=opts, callback=cb)
    for k in range(3):
        np.testing.assert_array_equal(snaps[k], seen[(k + 1) * 4 - 1])
    np.testing.assert_array_equal(snaps[-1], u)
