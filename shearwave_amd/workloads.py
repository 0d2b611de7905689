"""Seeded synthetic inputs of the benchmark workloads (NumPy, host side).

* ``plain_64``: the reference's CPU-runnable case: 64^3, h = 0.5 mm,
  homogeneous rho = 1000, mu = 2.5 kPa, eta = 1 Pa s, a Gaussian axial push
  (sigma = 2 mm) at the centre for the first 100 us, nT = 500, CG tol 1e-10,
  snapshot stride 10.
* ``dropin_256``: the single drop-in solve: 256^3, h = 0.25 mm,
  rho = 1000 (1 + U(-0.01, 0.01)), mu = 3 kPa with 8 seeded spherical
  inclusions (radius U(2, 6) mm, mu U(6, 20) kPa, later ones overwrite),
  eta U(0.5, 3) Pa s per voxel, Gaussian push (sigma = 1.5 mm) along z at the
  centre for t < 100 us, nT = 2000, CG tol 1e-8, invariants every 20 steps.

The time step is min(0.3 h / (c_max sqrt 3), 0.5 rho_min h^2 / (6 eta_max)):
the CFL-0.3 step alone violates the explicit viscous limit rho h^2 / (6 eta)
and diverges; the factor-0.5 viscous cap is the rule of the reference's
example driver (examples/sparse_transcranial_shear_3d.py:1258-1265).
"""

from __future__ import annotations

import math

import numpy as np


def stable_dt(h, rho_min, mu_max, eta_max, cfl=0.3):
    dt_cfl = cfl * h / (math.sqrt(mu_max / rho_min) * math.sqrt(3.0))
    dt_visc = 0.5 * rho_min * h * h / (6.0 * eta_max) if eta_max > 0 else math.inf
    return min(dt_cfl, dt_visc), dt_cfl


def gaussian_push(shape, h, sigma, amplitude=300.0, centre=None):
    c = [(n - 1) / 2 * h for n in shape] if centre is None else centre
    ax = [np.arange(n) * h - cc for n, cc in zip(shape, c)]
    r2 = ax[0][:, None, None] ** 2 + ax[1][None, :, None] ** 2 + ax[2][None, None, :] ** 2
    return amplitude * np.exp(-r2 / (2 * sigma * sigma))


def inclusion_map(shape, h, rng, n_incl, background, r_range=(2e-3, 6e-3), mu_range=(6e3, 20e3)):
    mu = np.full(shape, float(background))
    ext = [n * h for n in shape]
    ax = [np.arange(n) * h for n in shape]
    for _ in range(n_incl):
        r = rng.uniform(*r_range)
        c = [rng.uniform(r, e - r) for e in ext]
        val = rng.uniform(*mu_range)
        # only the bounding box of the sphere is touched
        lo = [max(0, int((cc - r) / h)) for cc in c]
        hi = [min(n, int((cc + r) / h) + 2) for cc, n in zip(c, shape)]
        sub = (
            (ax[0][lo[0]:hi[0], None, None] - c[0]) ** 2
            + (ax[1][None, lo[1]:hi[1], None] - c[1]) ** 2
            + (ax[2][None, None, lo[2]:hi[2]] - c[2]) ** 2
        ) < r * r
        mu[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]][sub] = val
    return mu


def plain_64(n=64):
    h = 0.5e-3
    rho, mu, eta = 1000.0, 2500.0, 1.0
    dt, dt_cfl = stable_dt(h, rho, mu, eta)
    n_steps = 500
    bz = gaussian_push((n, n, n), h, 2e-3)
    env = (np.arange(n_steps) * dt < 100e-6).astype(np.float64)
    return dict(shape=(n, n, n), h=h, rho=rho, mu=mu, eta=eta, dt=dt, dt_cfl=dt_cfl, n_steps=n_steps, bx=None,
                by=None, bz=bz, env=env, tol=1e-10, max_iters=5000, snapshot_stride=10)


def dropin_256(n=256, seed=2, n_steps=2000):
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    h = 0.25e-3
    rho = 1000.0 * (1.0 + rng.uniform(-0.01, 0.01, shape))
    mu = inclusion_map(shape, h, rng, 8, 3000.0)
    eta = rng.uniform(0.5, 3.0, shape)
    dt, dt_cfl = stable_dt(h, float(rho.min()), float(mu.max()), float(eta.max()))
    bz = gaussian_push(shape, h, 1.5e-3)
    env = (np.arange(n_steps) * dt < 100e-6).astype(np.float64)
    return dict(shape=shape, h=h, rho=rho, mu=mu, eta=eta, dt=dt, dt_cfl=dt_cfl, n_steps=n_steps, bx=None, by=None,
                bz=bz, env=env, tol=1e-8, max_iters=20000, strain_stride=20)


def force4d_128(n=128, n_steps=100, seed=6):
    """Non-separable forcing (solver.py:118-122, 199): a (n, n, n, nT) float32 bz, a Gaussian push
    (sigma 2 mm) whose centre moves 16 voxels along z over the run while its amplitude ramps from 300 to 0,
    projected every step at the reference's defaults (tol 1e-5, 200 iterations); h = 0.5 mm, rho 1000,
    mu 3 kPa with 3 seeded inclusions, eta 1."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    h = 0.5e-3
    rho, eta = 1000.0, 1.0
    mu = inclusion_map(shape, h, rng, 3, 3000.0)
    dt, dt_cfl = stable_dt(h, rho, float(mu.max()), eta)
    sigma = 2e-3
    ax = (np.arange(n) - (n - 1) / 2) * h
    gxy = np.exp(-(ax[:, None] ** 2 + ax[None, :] ** 2) / (2 * sigma**2)).astype(np.float32)
    c = (np.linspace(-8, 8, n_steps) * h)[None, :]
    amp = np.linspace(300.0, 0.0, n_steps)[None, :]
    gz = (amp * np.exp(-((ax[:, None] - c) ** 2) / (2 * sigma**2))).astype(np.float32)  # (nz, nT)
    bz = np.empty(shape + (n_steps,), np.float32)
    np.multiply(gxy[:, :, None, None], gz[None, None, :, :], out=bz)
    return dict(shape=shape, h=h, rho=rho, mu=mu, eta=eta, dt=dt, dt_cfl=dt_cfl, n_steps=n_steps, bz=bz,
                tol=1e-5, max_iters=200)


# algorithmic HBM bytes per voxel of each kernel (read + write, once each)
def alg_step_bytes(itemsize, viscous=True, forced=False):
    """Algorithmic bytes of one time step per voxel, in the reference's (u, u_prev) form.

    Read u, u_prev (3 components each), mu/rho, eta/rho and, on a force-active
    step, f/rho; write u_next: 44 B in float32, 88 B in float64, +12/24 B when
    forced, -4/8 B when inviscid. This is the unit of the roofline figure
    (bytes per voxel-step), whatever form a kernel stores.
    """
    vec = 3 * itemsize
    b = 2 * vec + vec + itemsize * (2 if viscous else 1)
    if forced:
        b += vec
    return b


def step_bytes(itemsize, form="increment", viscous=True, forced=False):
    """Time step: u, b read (stencil, counted once), u' (and d') written, mu/rho (and eta/rho) read."""
    vec = 3 * itemsize
    b = 2 * vec + vec  # u, b read; u' written
    if form == "increment":
        b += vec  # d' written
    b += itemsize * (2 if viscous else 1)
    if forced:
        b += vec
    return b


CG_BYTES_PER_ITER = 8 * 8  # dir: r, p_old read, p written; step: phi, r, p read, phi, r written (float64)
CG_TEXTBOOK_BYTES_PER_ITER = 11 * 8  # the three-kernel form: A p, p.Ap, phi and r updates, r.r (float64)


def batch_48(seed=3, n_solves=48, h=0.5e-3):
    """The batching workload: 48 independent solves of uneven sizes.

    (nx, ny) ~ U{64..192} and nz ~ U{64..256}, rounded to multiples of 8;
    h = 0.5 mm; rho = 1000 (1 + U(-0.01, 0.01)); mu = 2 kPa with 3-8 seeded
    inclusions U(2, 20) kPa; eta U(0.5, 3) Pa s per voxel; nT ~ U{500..2000};
    a Gaussian push (sigma = 2 mm) along z at a random interior position,
    active for t < 100 us; float32 storage, CG tol 1e-8; per-solve dt by
    the stability rule above.
    """
    from .batch import SolveSpec

    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n_solves):
        nx, ny = (int(8 * round(rng.integers(64, 193) / 8)) for _ in range(2))
        nz = int(8 * round(rng.integers(64, 257) / 8))
        shape = (nx, ny, nz)
        rho = 1000.0 * (1.0 + rng.uniform(-0.01, 0.01, shape))
        mu = inclusion_map(shape, h, rng, int(rng.integers(3, 9)), 2000.0, mu_range=(2e3, 20e3))
        eta = rng.uniform(0.5, 3.0, shape)
        n_steps = int(rng.integers(500, 2001))
        dt, dt_cfl = stable_dt(h, float(rho.min()), float(mu.max()), float(eta.max()))
        centre = [rng.uniform(0.25, 0.75) * (n - 1) * h for n in shape]
        bz = gaussian_push(shape, h, 2e-3, centre=centre)
        env = (np.arange(n_steps) * dt < 100e-6).astype(np.float64)
        specs.append(SolveSpec(shape=shape, h=(h, h, h), dt=dt, rho=rho, mu=mu, eta=eta, force=(None, None, bz),
                               env=env, n_steps=n_steps, tol=1e-8, max_iters=20000,
                               meta={"dt_cfl": dt_cfl, "index": k}))
    return specs
