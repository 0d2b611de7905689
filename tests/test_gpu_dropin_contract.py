"""GPU tests of the drop-in's remaining contract, against the NumPy oracle.

Covers: ``return_device`` (solver.py:556-566), ``project_body_force_to_shear``
with its defaults (solver.py:628-641), the float32 form of ``reproject_every``
over three wall-arrival times (the scenario of the reference's
tests/test_divergence_free.py:96-102) and of 4-D forcing (solver.py:118-122,
199), the reference's CPU-runnable 64^3 workload at full size, a 96^3 float64
solve with the automatic launch geometry, the CG's true residual, the
non-finite check and the bounds checks of the trace kernels.

Gates: float64 storage within 1e-10 relative L2 of the float64 oracle;
float32 storage within 1e-5 on u and v; true CG residual below tol.
"""

import math
import warnings

import numpy as np
import pytest
import torch

from oracle import solver as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def taper(shape):
    s = [np.sin(np.pi * np.arange(n) / (n - 1)) for n in shape]
    return s[0][:, None, None] * s[1][None, :, None] * s[2][None, None, :]


def gauss(shape, sigma, centre=None):
    c = [(n - 1) / 2 for n in shape] if centre is None else centre
    g = np.meshgrid(*(np.arange(n) - cc for n, cc in zip(shape, c)), indexing="ij")
    return np.exp(-(g[0] ** 2 + g[1] ** 2 + g[2] ** 2) / (2 * sigma**2))


def _problem(shape=(34, 32, 36), n_steps=60, seed=21):
    rng = np.random.default_rng(seed)
    h = (0.3e-3, 0.3e-3, 0.3e-3)
    rho = 1000.0 * (1 + rng.uniform(-0.01, 0.01, shape))
    mu = 3000.0 + 4000.0 * rng.uniform(0, 1, shape)
    eta = rng.uniform(0.5, 2.0, shape)
    bz = 300.0 * gauss(shape, 3.0)
    env = np.zeros(n_steps)
    env[:12] = np.linspace(1.0, 0.2, 12)
    dt = 2e-6
    opts = {"env_t": env, "eta": eta, "vel_damp": 0.01, "poisson_tol": 1e-12, "poisson_max_iters": 5000}
    return dict(shape=shape, h=h, rho=rho, mu=mu, bz=bz, dt=dt, n_steps=n_steps, opts=opts)


# ------------------------------------------------------------- return_device


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_return_device_equals_host_arrays(sw, dtype):
    """return_device=True (solver.py:556-566): the same values as device tensors, plain, with snapshots and traces."""
    p = _problem()
    args = (None, None, p["bz"], p["rho"], p["mu"], *p["h"], p["dt"], p["n_steps"])
    opts = {**p["opts"], "sw_dtype": dtype}
    u, v = sw.shear_fdtd(*args, opts)
    ud, vd = sw.shear_fdtd(*args, opts, return_device=True)
    for a, b in ((ud, u), (vd, v)):
        assert isinstance(a, torch.Tensor) and a.is_cuda and a.dtype == getattr(torch, dtype)
        assert tuple(a.shape) == p["shape"] + (3,)
        np.testing.assert_array_equal(a.cpu().numpy(), b)
    # snapshots
    _, _, s = sw.shear_fdtd(*args, opts, snapshot_stride=10)
    ud, vd, sd = sw.shear_fdtd(*args, opts, snapshot_stride=10, return_device=True)
    assert isinstance(sd, torch.Tensor) and tuple(sd.shape) == (p["n_steps"] // 10,) + p["shape"] + (3,)
    np.testing.assert_array_equal(sd.cpu().numpy(), s)
    np.testing.assert_array_equal(ud.cpu().numpy(), u)
    # traces
    idx = (np.array([3, 17, 30]), np.array([5, 16, 2]), np.array([18, 18, 33]))
    # (the trace runs are compared with each other: traces of a two-step pass are gathered after it)
    u_t, v_t, tr = sw.shear_fdtd_jax(*args, opts, trace_indices=idx, return_traces=True)
    ud, vd, trd = sw.shear_fdtd_jax(*args, opts, trace_indices=idx, return_traces=True, return_device=True)
    assert isinstance(trd, torch.Tensor) and tuple(trd.shape) == (3, p["n_steps"])
    np.testing.assert_array_equal(trd.cpu().numpy(), tr)
    np.testing.assert_array_equal(ud.cpu().numpy(), u_t)
    np.testing.assert_array_equal(vd.cpu().numpy(), v_t)
    # against the oracle's own callbacks: trace k at step t is u_z[idx_k] after step t
    want = []
    O.shear_fdtd(*args, opts=p["opts"], callback=lambda it, uu: want.append(uu[idx[0], idx[1], idx[2], 2].copy()),
                 dtype=np.float64)
    tol = 1e-10 if dtype == "float64" else 1e-5
    assert rel(tr, np.stack(want, axis=1)) < tol


# ------------------------------------------------- project_body_force_to_shear


def test_project_body_force_to_shear_defaults(sw):
    """solver.py:628-641: project_vector_field(0, 0, bz) with tol 5e-5 and 400 iterations."""
    shape = (40, 36, 44)
    h = (0.3e-3, 0.25e-3, 0.35e-3)
    rng = np.random.default_rng(8)
    bz = (300.0 * gauss(shape, 4.0) + 10.0 * rng.standard_normal(shape) * taper(shape)).astype(np.float32)
    z = np.zeros(shape)
    # float64: the same CG as the oracle's, iteration for iteration
    info = {}
    got = sw.project_body_force_to_shear(bz.astype(np.float64), *h, dtype="float64", info=info)
    oinfo = {}
    want = O.project_vector_field(z, z, bz.astype(np.float64), *h, 5e-5, 400, dtype=np.float64, info=oinfo)
    assert info["reason"] == "converged" and info["iters"] == oinfo["iters"], (info, oinfo)
    assert info["iters"] < 400
    for a, b in zip(got, want):
        assert rel(a, b) < 1e-10
    # default float32: float32 arrays out, true residual below the default tol
    info = {}
    got = sw.project_body_force_to_shear(bz, *h, info=info)
    assert all(g.dtype == np.float32 and g.shape == shape for g in got)
    assert info["reason"] == "converged" and 0 < info["true_relres"] < 5e-5, info
    f64 = [z, z, bz.astype(np.float64)]
    m = O.interior_mask(shape, np.float64)
    d_in = np.linalg.norm(O.divergence_center(*(f * m for f in f64), *h) * m)
    d_out = np.linalg.norm(O.divergence_center(*(g.astype(np.float64) for g in got), *h) * m)
    assert d_out / d_in < 5e-5 * 1.01, d_out / d_in
    # both approximate the exact projection to the tolerance's order
    exact = O.project_vector_field(z, z, bz.astype(np.float64), *h, 1e-13, 5000, dtype=np.float64)
    for a, b in zip(got, exact):
        assert rel(a, b) < 1e-4


# --------------------------------------------------- float32 reprojection, 4-D


def _div_curl_ratio(u, d):
    """RMS(div u)/RMS(|curl u|) over the inner half-cube (tests/test_divergence_free.py:30-42)."""
    j = [np.gradient(u[..., i].astype(np.float64), d, d, d, edge_order=1) for i in range(3)]
    div = j[0][0] + j[1][1] + j[2][2]
    cx, cy, cz = j[2][1] - j[1][2], j[0][2] - j[2][0], j[1][0] - j[0][1]
    curl = np.sqrt(cx * cx + cy * cy + cz * cz)
    q = u.shape[0] // 4
    sl = (slice(q, u.shape[0] - q),) * 3
    return float(np.sqrt(np.mean(div[sl] ** 2)) / np.sqrt(np.mean(curl[sl] ** 2)))


def test_reproject_every_float32_long_run(sw):
    """The reference's reprojection scenario (tests/test_divergence_free.py:45-56, 96-102): a 32^3 cube, axial
    Gaussian push for the first third, vel_damp 0.02, eta 0.5, reproject_every 20, run for 3x the wall-arrival
    time, default CG tol. Float32 storage within 1e-5 of the float64 oracle on u and v, and the reference's
    div/curl bound (1e-4) at every 20th step."""
    n, d, mu, rho = 32, 1.0e-3, 4580.0, 1000.0
    ax = (np.arange(n) - (n - 1) / 2) * d
    x, y, z = np.meshgrid(ax, ax, ax, indexing="ij")
    sigma = 2.5 * d
    bx = (300.0 * np.exp(-(x**2 + y**2 + z**2) / (2 * sigma**2))).astype(np.float32)
    cs = math.sqrt(mu / rho)
    dt = 0.1 * d / (cs * math.sqrt(3.0))
    n_steps = int(3.0 * (n // 4) * (d / (cs * dt)))
    env = np.zeros(n_steps, np.float32)
    env[: max(1, n_steps // 3)] = 1.0
    opts = {"vel_damp": 0.02, "eta": 0.5, "reproject_every": 20, "env_t": env}
    ratios = []

    def cb(it, u):
        if it % 20 == 19:
            ratios.append(_div_curl_ratio(u, d))

    u32, v32 = sw.shear_fdtd(bx, None, None, rho, mu, d, d, d, dt, n_steps, opts, callback=cb)
    u64, v64 = O.shear_fdtd(bx, None, None, rho, mu, d, d, d, dt, n_steps, opts=opts, dtype=np.float64)
    assert u32.dtype == np.float32
    eu, ev = rel(u32, u64), rel(v32, v64)
    assert eu < 1e-5 and ev < 1e-5, (eu, ev)
    assert ratios and max(ratios) < 1e-4, ratios


def test_4d_forcing_float32(sw):
    """Non-separable (nx,ny,nz,nT) forcing, projected every step (solver.py:118-122, 199): float32 storage
    within 1e-5 of the float64 oracle on u and v."""
    shape, n_steps = (26, 24, 68), 30
    rng = np.random.default_rng(31)
    h = (0.3e-3, 0.3e-3, 0.3e-3)
    rho = 1000.0 * (1 + rng.uniform(-0.01, 0.01, shape))
    mu = 3000.0 + 3000.0 * rng.uniform(0, 1, shape)
    ramp = np.linspace(1.0, 0.0, n_steps)
    g = gauss(shape, 3.0)[..., None]
    bz = 300.0 * g * ramp
    bx = 50.0 * g * np.sin(np.linspace(0, 3, n_steps)) + rng.standard_normal(shape + (1,)) * taper(shape)[..., None]
    opts = {"eta": 1.0, "vel_damp": 0.0, "poisson_tol": 1e-9, "poisson_max_iters": 5000}
    u32, v32 = sw.shear_fdtd(bx, None, bz, rho, mu, *h, 2e-6, n_steps, opts)
    u64, v64 = O.shear_fdtd(bx, None, bz, rho, mu, *h, 2e-6, n_steps, opts=opts, dtype=np.float64)
    eu, ev = rel(u32, u64), rel(v32, v64)
    assert eu < 1e-5 and ev < 1e-5, (eu, ev)
    # the forces resident on the device (default) and sliced on the host each step give the same bits
    u_h, v_h = sw.shear_fdtd(bx, None, bz, rho, mu, *h, 2e-6, n_steps, {**opts, "sw_resident_force": False})
    assert np.array_equal(u_h, u32) and np.array_equal(v_h, v32)


@pytest.mark.parametrize("max_iters", [20, 60])
def test_unconverged_projection_float32_vectors(sw, max_iters):
    """The reference's default projection stops at max_iters before tol (poisson_tol 1e-5, 200 iterations,
    solver.py:95-96, on large grids): the float32-vector CG's iterate at that count within 1e-5 of the float64
    oracle's, on the projected force."""
    rng = np.random.default_rng(12)
    shape = (40, 36, 44)
    f = [(rng.standard_normal(shape) * taper(shape)).astype(np.float32) for _ in range(2)]
    f.append((300.0 * gauss(shape, 4.0) + rng.standard_normal(shape) * taper(shape)).astype(np.float32))
    h = (0.3e-3,) * 3
    info = {}
    got = sw.project_vector_field(*f, *h, 1e-5, max_iters, info=info)
    assert info["reason"] == "max_iters" and info["iters"] == max_iters, info
    want = O.project_vector_field(*(a.astype(np.float64) for a in f), *h, 1e-5, max_iters, dtype=np.float64)
    errs = [rel(a, b) for a, b in zip(got, want)]
    assert max(errs) < 1e-5, errs


def test_4d_forcing_unconverged_projections_float32(sw):
    """Per-step projections that stop at max_iters (20) before tol: float32 storage within 1e-5 of the float64
    oracle on u and v, with the steps' forces projected in batched windows (the default: 16 and 8 steps here),
    each on its own (window 1), and windows of 5 over forces sliced on the host (the same bits as resident)."""
    shape, n_steps = (26, 24, 68), 24
    rng = np.random.default_rng(33)
    h = (0.3e-3, 0.3e-3, 0.3e-3)
    rho = 1000.0 * (1 + rng.uniform(-0.01, 0.01, shape))
    mu = 3000.0 + 3000.0 * rng.uniform(0, 1, shape)
    g = gauss(shape, 3.0)[..., None]
    bz = 300.0 * g * np.linspace(1.0, 0.2, n_steps) + rng.standard_normal(shape + (1,)) * taper(shape)[..., None]
    opts = {"eta": 1.0, "poisson_tol": 1e-5, "poisson_max_iters": 20}
    info = []
    u64, v64 = O.shear_fdtd(None, None, bz, rho, mu, *h, 2e-6, n_steps, opts=opts, dtype=np.float64, proj_info=info)
    assert all(i["iters"] == 20 for i in info), [i["iters"] for i in info]
    runs = {}
    for name, extra in (("default", {}), ("each", {"sw_force_window": 1}), ("w5", {"sw_force_window": 5}),
                        ("w5_host", {"sw_force_window": 5, "sw_resident_force": False})):
        u32, v32 = sw.shear_fdtd(None, None, bz, rho, mu, *h, 2e-6, n_steps, {**opts, **extra})
        eu, ev = rel(u32, u64), rel(v32, v64)
        assert eu < 1e-5 and ev < 1e-5, (name, eu, ev)
        runs[name] = (u32, v32)
    assert np.array_equal(runs["w5"][0], runs["w5_host"][0]) and np.array_equal(runs["w5"][1], runs["w5_host"][1])


def test_4d_forcing_windows_mixed_components_zero_steps_callback_reproject(sw):
    """Per-step forcing through every part of the loop at once: a 4-D bz with all-zero steps (their projections
    exit early), a 3-D bx repeated every step (solver.py:1060-1069), a callback on every step, reprojection every
    7 steps, and windows of 6 steps' projections (the last one short); float32 within 1e-5 and float64 within
    1e-10 of the float64 oracle, and the callback sees every step."""
    shape, n_steps = (24, 20, 66), 20
    rng = np.random.default_rng(35)
    h = (0.3e-3, 0.3e-3, 0.3e-3)
    rho = 1000.0 * (1 + rng.uniform(-0.01, 0.01, shape))
    mu = 3000.0 + 3000.0 * rng.uniform(0, 1, shape)
    g = gauss(shape, 3.0)[..., None]
    ramp = np.linspace(1.0, 0.3, n_steps)
    ramp[[4, 5, 13]] = 0.0
    bz = 300.0 * g * ramp
    bx = 40.0 * gauss(shape, 4.0) + rng.standard_normal(shape) * taper(shape)
    opts = {"eta": 1.0, "vel_damp": 0.01, "poisson_tol": 1e-6, "poisson_max_iters": 40, "reproject_every": 7,
            "sw_force_window": 6}
    u64, v64 = O.shear_fdtd(bx, None, bz, rho, mu, *h, 2e-6, n_steps, opts=opts, dtype=np.float64)
    for dtype, tol in (("float32", 1e-5), ("float64", 1e-10)):
        seen = []
        u, v = sw.shear_fdtd(bx, None, bz, rho, mu, *h, 2e-6, n_steps, {**opts, "sw_dtype": dtype},
                             callback=lambda it, uu: seen.append(it))
        assert seen == list(range(n_steps)), (dtype, seen)
        eu, ev = rel(u, u64), rel(v, v64)
        assert eu < tol and ev < tol, (dtype, eu, ev)


# --------------------------------------------------------- full-size workloads


def test_plain_64_full_size(sw):
    """The reference's CPU-runnable workload at full size: 64^3, h = 0.5 mm, rho 1000, mu 2500, eta 1, Gaussian
    push (sigma 2 mm) for the first 100 us, nT = 500, CG tol 1e-10, snapshots every 10 steps. Float64 within
    1e-10 and float32 within 1e-5 of the float64 oracle, on u, v and every snapshot."""
    from shearwave_amd import workloads as W

    w = W.plain_64()
    h = (w["h"],) * 3
    opts = {"env_t": w["env"], "eta": w["eta"], "poisson_tol": w["tol"], "poisson_max_iters": w["max_iters"]}
    args = (None, None, w["bz"], w["rho"], w["mu"], *h, w["dt"], w["n_steps"])
    u_o, v_o, s_o = O.shear_fdtd_snapshots(*args, opts, w["snapshot_stride"], dtype=np.float64)
    assert s_o.shape == (50, 64, 64, 64, 3)
    for dtype, tol in (("float64", 1e-10), ("float32", 1e-5)):
        u, v, s = sw.shear_fdtd(*args, {**opts, "sw_dtype": dtype}, snapshot_stride=w["snapshot_stride"])
        errs = (rel(u, u_o), rel(v, v_o), rel(s, s_o))
        assert max(errs) < tol, (dtype, errs)


def test_fp64_96_auto_geometry_converged_cg(sw):
    """96^3 float64 with the automatic launch geometry (several x chunks, XCD remap) against the float64 oracle at
    1e-10, the projection converged to tol 1e-13.

    The CG is converged rather than pinned at a mid-way iteration count: mid-way, its iterates amplify rounding
    differences of the dot products (different reduction orders) by ~1e7 (on this force at 150 iterations, relres
    9e-5, a 2e-16 change per dot product moves phi by 5e-9), while converged ones agree to ~1e-15."""
    shape, n_steps = (96, 96, 96), 64
    rng = np.random.default_rng(96)
    h = (0.25e-3, 0.25e-3, 0.25e-3)
    rho = 1000.0 * (1 + rng.uniform(-0.01, 0.01, shape))
    mu = 3000.0 + 9000.0 * (gauss(shape, 9.0, (30, 50, 60)) > 0.5) + 500.0 * rng.uniform(0, 1, shape)
    eta = rng.uniform(0.5, 3.0, shape)
    bz = 300.0 * gauss(shape, 6.0)
    env = np.zeros(n_steps)
    env[:9] = 1.0
    dt = 1.5e-6
    opts = {"env_t": env, "eta": eta, "poisson_tol": 1e-13, "poisson_max_iters": 5000}
    u, v = sw.shear_fdtd(None, None, bz, rho, mu, *h, dt, n_steps, {**opts, "sw_dtype": "float64"})
    u_o, v_o = O.shear_fdtd(None, None, bz, rho, mu, *h, dt, n_steps, opts=opts, dtype=np.float64)
    assert float(np.abs(u_o).max()) > 0
    errs = (rel(u, u_o), rel(v, v_o))
    assert max(errs) < 1e-10, errs


# -------------------------------------------------------------- true residual


@pytest.mark.parametrize("tol", [1e-5, 1e-8, 1e-10])
def test_float32_cg_true_residual_meets_tol(sw, tol):
    """With float32 CG vectors the returned phi's TRUE residual |b - A phi| / |b| (b in float64) is below tol.

    On the interior b - A phi = -div_c f_S, so the projected field itself shows it: with float64 storage (and
    float32 r and p) the oracle's float64 stencils measure |div_c f_S| / |div_c mask f| directly, and it must be
    below tol and agree with the reported true_relres. (A float32 f_S cannot show it below ~4e-7: its own
    rounding has a divergence of that order.) Float32 storage reports the same true residual.
    """
    shape = (48, 40, 72)
    h = (0.25e-3,) * 3
    rng = np.random.default_rng(5)
    f = [(rng.standard_normal(shape) * taper(shape)).astype(np.float32) for _ in range(2)]
    f.append((300.0 * gauss(shape, 5.0)).astype(np.float32))
    m = O.interior_mask(shape, np.float64)
    f64 = [a.astype(np.float64) for a in f]
    d_in = np.linalg.norm(O.divergence_center(*(a * m for a in f64), *h) * m)
    info = {}
    out = sw.project_vector_field(*f64, *h, tol, 20000, dtype="float64", info=info, cg_vectors="float32")
    assert info["reason"] == "converged", info
    ratio = np.linalg.norm(O.divergence_center(*out, *h) * m) / d_in
    assert ratio < tol, (ratio, info)
    assert abs(info["true_relres"] - ratio) < 1e-3 * ratio, (ratio, info)
    info32 = {}
    sw.project_vector_field(*f, *h, tol, 20000, info=info32, cg_vectors="float32")
    assert info32["reason"] == "converged" and 0 < info32["true_relres"] < tol, info32


# ------------------------------------------------------------ failure checks


def test_nonfinite_check_warns_on_unstable_dt(sw):
    """An explicit step above the viscous limit rho h^2 / (6 eta) diverges: the check reports the step."""
    shape, n_steps = (24, 20, 40), 200
    h = 0.5e-3
    eta = 5.0
    dt = 3.0 * 1000.0 * h * h / (6 * eta)  # 3x the viscous limit
    bz = 300.0 * gauss(shape, 3.0)
    opts = {"env_t": np.ones(n_steps), "eta": eta}
    with pytest.warns(RuntimeWarning, match="inf/NaN after step"):
        u, _ = sw.shear_fdtd(None, None, bz, 1000.0, 2500.0, h, h, h, dt, n_steps, opts)
    assert not np.all(np.isfinite(u))
    with warnings.catch_warnings():
        warnings.simplefilter("error", RuntimeWarning)
        u, _ = sw.shear_fdtd(None, None, bz, 1000.0, 2500.0, h, h, h, dt / 10, n_steps, opts)
    assert np.all(np.isfinite(u))


def test_trace_kernels_bounds_checked(sw):
    """sw_gather_traces yields NaN and sw_inject_traces skips an index outside [0, n)."""
    from shearwave_amd import _abi
    from shearwave_amd import engine as E

    lib = _abi.lib()
    n = 1000
    f = torch.arange(n, dtype=torch.float32, device="cuda")
    idx = torch.tensor([0, 999, 1000, -1, 5, 1 << 40], dtype=torch.int64, device="cuda")
    out = torch.zeros(6, dtype=torch.float32, device="cuda")
    _abi.check(lib.sw_gather_traces(_abi.SW_F32, f.data_ptr(), n, idx.data_ptr(), 6, out.data_ptr(), 1,
                                    E.stream_ptr()))
    o = out.cpu().numpy()
    np.testing.assert_array_equal(o[[0, 1, 4]], [0.0, 999.0, 5.0])
    assert np.all(np.isnan(o[[2, 3, 5]]))
    g = torch.ones(6, dtype=torch.float32, device="cuda")
    _abi.check(lib.sw_inject_traces(_abi.SW_F32, f.data_ptr(), n, idx.data_ptr(), 6, g.data_ptr(), 1,
                                    E.stream_ptr()))
    want = np.arange(n, dtype=np.float32)
    want[[0, 999, 5]] += 1.0
    np.testing.assert_array_equal(f.cpu().numpy(), want)


# ----------------------------------------------------------- in-loop strain


@pytest.mark.parametrize("dtype,tol", [("float32", 2e-6), ("float64", 1e-12)])
def test_in_loop_strain_matches_numpy_strain_of_the_snapshots(sw, dtype, tol):
    """The invariants computed inside the loop (x-marching kernel, partials times 1/(2h)) against
    compute_strain_invariants(compute_strain_tensor(u)) (strain.py:16-135, NumPy in the same precision) of
    the snapshot taken at the same step: within a few roundings (rel L2 per invariant)."""
    from oracle import strain as S

    p = _problem(shape=(34, 40, 70), n_steps=40, seed=23)
    args = (None, None, p["bz"], p["rho"], p["mu"], *p["h"], p["dt"], p["n_steps"])
    opts = {**p["opts"], "sw_dtype": dtype, "sw_strain_stride": 10}
    u, v, snaps, strain = sw.shear_fdtd(*args, opts, snapshot_stride=10, return_strain=True)
    assert snaps.shape[0] == 4 and strain["volumetric"].shape == (4,) + p["shape"]
    for k in range(4):
        s = snaps[k]
        want = S.compute_strain_invariants(S.compute_strain_tensor(s[..., 0], s[..., 1], s[..., 2], *p["h"]))
        for key in ("volumetric", "deviatoric", "max_shear", "von_mises"):
            assert float(np.abs(want[key]).max()) > 0
            assert rel(strain[key][k], want[key]) < tol, (k, key, rel(strain[key][k], want[key]))


# ------------------------------------------------- spectral preconditioner


@pytest.mark.parametrize("shape", [(40, 21, 70), (33, 34, 35), (64, 64, 64), (9, 70, 12)])
def test_preconditioned_projection(sw, shape, monkeypatch):
    """The float32-vector CG with the spectral preconditioner (forced on: SW_CG_PRECOND=1) converges in a few
    iterations to a TRUE residual below tol, and the projection agrees with the float64 oracle's converged one
    (odd and even axis lengths: every pairing of flux and value chain ends; unequal spacings)."""
    monkeypatch.setenv("SW_CG_PRECOND", "1")
    h = (0.3e-3, 0.25e-3, 0.35e-3)
    rng = np.random.default_rng(sum(shape))
    f = [(rng.standard_normal(shape) * taper(shape)) for _ in range(2)]
    f.append(300.0 * gauss(shape, 4.0) + rng.standard_normal(shape) * taper(shape))
    m = O.interior_mask(shape, np.float64)
    d_in = np.linalg.norm(O.divergence_center(*(a * m for a in f), *h) * m)
    info = {}
    out = sw.project_vector_field(*f, *h, 1e-10, 5000, dtype="float64", info=info, cg_vectors="float32")
    assert info["reason"] == "converged" and info["iters"] <= 12, info
    ratio = np.linalg.norm(O.divergence_center(*out, *h) * m) / d_in
    assert ratio < 1e-10 and abs(info["true_relres"] - ratio) < 1e-3 * ratio, (ratio, info)
    want = O.project_vector_field(*f, *h, 1e-13, 20000, dtype=np.float64)
    for a, b in zip(out, want):
        assert rel(a, b) < 1e-9
    info32 = {}
    out32 = sw.project_vector_field(*(a.astype(np.float32) for a in f), *h, 1e-9, 5000, info=info32)
    assert info32["reason"] == "converged" and info32["iters"] <= 12 and info32["true_relres"] < 1e-9, info32
    for a, b in zip(out32, want):
        assert rel(a, b) < 1e-6


def test_preconditioner_only_where_the_plain_cg_converges(sw):
    """By default the preconditioned CG stands in for the plain one only at tol <= 1e-7 with max_iters at least
    the plain CG's iteration estimate: the reference's defaults (tol 1e-5, 200 iterations) keep the plain CG and
    its iterates, tight tolerances take a few preconditioned iterations."""
    shape = (48, 40, 72)
    h = (0.25e-3,) * 3
    rng = np.random.default_rng(6)
    f = [(rng.standard_normal(shape) * taper(shape)).astype(np.float32) for _ in range(3)]
    runs = {}
    for tol, max_iters in ((1e-5, 200), (1e-8, 20000), (1e-8, 40)):
        info = {}
        sw.project_vector_field(*f, *h, tol, max_iters, info=info)
        runs[(tol, max_iters)] = info
    assert runs[(1e-5, 200)]["iters"] > 30  # plain CG
    assert runs[(1e-8, 20000)]["iters"] <= 12 and runs[(1e-8, 20000)]["true_relres"] < 1e-8
    assert runs[(1e-8, 40)]["reason"] == "max_iters" and runs[(1e-8, 40)]["iters"] == 40  # plain, stopped early


@pytest.mark.parametrize("dtype,tile_y", [("float32", 0), ("float64", 0), ("float32", 12)])
@pytest.mark.parametrize("stride", [4, 5])
def test_in_loop_strain_fused_into_the_next_pass(sw, dtype, tile_y, stride):
    """The in-loop invariants that the next force-free two-step pass computes from its input state equal
    those of their own kernel bit for bit. Stride 5 puts outputs before single steps, stride 4 before
    two-step passes; the force of the first steps puts some before forced passes (own kernel), and the
    last step's invariants come after the last pass (own kernel). Float32 also with the 12-row tile."""
    import torch

    from shearwave_amd import engine as E

    p = _problem(shape=(40, 36, 70), n_steps=60, seed=29)
    shape = (40, 36, 70)
    got = []
    for separate in (True, False):
        s = E.Solve(shape, p["h"], p["dt"], p["rho"], p["mu"], p["opts"]["eta"], dtype=dtype,
                    vel_damp=p["opts"]["vel_damp"], kernel="march2", tile_y=tile_y)
        s.set_force(None, None, p["bz"], 1e-10, 5000)
        inv = torch.full((p["n_steps"] // stride, 4) + shape, float("nan"), dtype=E.torch_dtype(dtype),
                         device=E.device())
        s.run(p["n_steps"], p["opts"]["env_t"], inv_stride=stride, inv_out=inv, inv_separate=separate)
        got.append((inv.cpu().numpy(), [t.cpu().numpy() for t in s.u_components()]))
    (inv_a, u_a), (inv_b, u_b) = got
    assert np.isfinite(inv_a).all() and float(np.abs(inv_a).max()) > 0
    np.testing.assert_array_equal(inv_b, inv_a)
    for a, b in zip(u_a, u_b):
        np.testing.assert_array_equal(b, a)
