"""Seeded inputs of the golden cases (shared by the writer and the tests).

Inputs are regenerated from seeds, so ``ref_fp32.npz`` only holds outputs.
Each case exercises one part of the reference's float32 behaviour.
"""

import numpy as np

CASES = ("separable", "separable_np64_dt", "nonseparable", "reproject", "project")


def _gauss(shape, sigma, centre=None):
    c = [(n - 1) / 2 for n in shape] if centre is None else centre
    g = np.meshgrid(*(np.arange(n) - cc for n, cc in zip(shape, c)), indexing="ij")
    return np.exp(-(g[0] ** 2 + g[1] ** 2 + g[2] ** 2) / (2 * sigma**2))


def _taper(shape):
    s = [np.sin(np.pi * np.arange(n) / (n - 1)) for n in shape]
    return s[0][:, None, None] * s[1][None, :, None] * s[2][None, None, :]


def build_case(name):
    if name in ("separable", "separable_np64_dt"):
        rng = np.random.default_rng(101)
        shape = (12, 10, 14)
        h = (2.0e-4, 2.5e-4, 1.8e-4)
        rho = (1000.0 * (1 + rng.uniform(-0.01, 0.01, shape))).astype(np.float32)
        mu = (4000.0 * (1 + rng.uniform(0, 1.5, shape))).astype(np.float32)
        eta = rng.uniform(0.5, 3.0, shape).astype(np.float32)
        bz = (0.03 * _gauss(shape, 2.5)).astype(np.float32)
        bx = (0.01 * rng.standard_normal(shape) * _taper(shape)).astype(np.float32)
        n_steps = 30 if name == "separable" else 12
        t = np.arange(n_steps, dtype=np.float32)
        env = (np.sin(2 * np.pi * 0.07 * t) * (1 + 0.5 * np.cos(2 * np.pi * 0.19 * t))).astype(np.float32)
        cs = float(np.sqrt(mu.max() / rho.min()))
        dT = 0.25 * min(h) / (cs * 3**0.5)
        dT = min(dT, 0.4 * float(rho.min()) * min(h) ** 2 / (6 * float(eta.max())))
        if name == "separable_np64_dt":
            dT = np.float64(dT)  # float64 scalar: NumPy 2 promotes the float32 arithmetic it touches
        opts = {"env_t": env, "eta": eta, "vel_damp": 0.02, "poisson_tol": 1e-6, "poisson_max_iters": 300}
        return dict(kind="fdtd", bx=bx, by=None, bz=bz, rho=rho, mu=mu, h=h, dT=dT, n_steps=n_steps,
                    opts=opts, stride=5 if name == "separable" else 0)
    if name == "nonseparable":
        rng = np.random.default_rng(202)
        shape = (10, 9, 11)
        h = (3.0e-4, 3.0e-4, 3.0e-4)
        n_steps = 6
        b3 = (0.02 * _gauss(shape, 2.0)).astype(np.float32)
        env = rng.uniform(-1, 1, n_steps).astype(np.float32)
        bz4 = (b3[..., None] * env[None, None, None, :]).astype(np.float32)
        by3 = (0.01 * rng.standard_normal(shape) * _taper(shape)).astype(np.float32)
        u0 = (1e-7 * rng.standard_normal(shape + (3,))).astype(np.float32)
        v0 = (1e-4 * rng.standard_normal(shape + (3,))).astype(np.float32)
        opts = {"eta": 0.5, "initial_u": u0, "initial_v": v0, "poisson_tol": 1e-5, "poisson_max_iters": 200}
        return dict(kind="fdtd", bx=None, by=by3, bz=bz4, rho=1000.0, mu=3500.0, h=h, dT=4e-6, n_steps=n_steps,
                    opts=opts)
    if name == "reproject":
        shape = (10, 10, 10)
        h = (1e-3, 1e-3, 1e-3)
        n_steps = 12
        bx = (300.0 * _gauss(shape, 2.5)).astype(np.float32)
        env = np.zeros(n_steps, np.float32)
        env[:4] = 1.0
        dT = 0.1 * 1e-3 / ((4580.0 / 1000.0) ** 0.5 * 3**0.5)
        opts = {"env_t": env, "vel_damp": 0.02, "eta": 0.5, "reproject_every": 5}
        return dict(kind="fdtd", bx=bx, by=None, bz=None, rho=1000.0, mu=4580.0, h=h, dT=dT, n_steps=n_steps,
                    opts=opts)
    if name == "project":
        rng = np.random.default_rng(303)
        shape = (16, 16, 16)
        tp = _taper(shape).astype(np.float32)
        f = [(rng.standard_normal(shape).astype(np.float32) * tp) for _ in range(3)]
        return dict(kind="project", fx=f[0], fy=f[1], fz=f[2], h=(1 / 16, 1 / 16, 1 / 16), tol=1e-8, max_iters=5000)
    raise KeyError(name)
