"""CPU oracle for the shear_fdtd hot path: a NumPy restatement of the reference.

THIS MODULE IS TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker. The product (``shearwave_amd``) never imports, calls or falls back to
it.

Every function restates the algorithm of the reference module
``src/shearwave/solver.py`` (cited per function) with the arithmetic type as a
parameter:

* ``dtype=np.float64`` is the oracle proper: every array and every
  intermediate is float64.
* ``dtype=np.float32`` mirrors the reference's own storage (float32 fields,
  float64 CG vectors cast back to float32), in the reference's order of
  operations, so that it reproduces the reference's float32 rounding.

Pinning (what checks each function; see ``tests/test_oracle_*.py``):

* ``laplacian_center`` / ``divergence_center`` / ``gradient_center``: exact on
  polynomials of degree 2 / 1 / 1, zero boundary layer, brute-force loops on
  tiny grids.
* ``poisson_operator`` / ``poisson_cg``: the operator against an explicit
  matrix assembled element by element from its closed form (including the
  rows next to the wall), CG against a dense direct solve, and the discrete
  identity ``|div_c(f - grad_c phi)| = |r_k|``.
* ``project_vector_field``: the reference's own invariant tests (pure
  gradient removed, solenoidal field kept, output divergence-free) and golden
  vectors written by the reference itself in float32.
* ``shear_fdtd``: Dirichlet sine eigenmodes against the closed-form roots of
  the scalar recurrence (elastic, Kelvin-Voigt, damped; with and without
  ``initial_v``), a brute-force loop implementation on a 5x6x7 grid with
  heterogeneous coefficients, the reference's divergence-free invariant, and
  float32 golden vectors written by the reference.
"""

from __future__ import annotations

import numpy as np

# ---------------------------------------------------------------------------
# stencil primitives
# ---------------------------------------------------------------------------


def _shift(f, axis, k):
    """Interior-sized view of ``f`` displaced by ``k`` voxels along ``axis``."""
    idx = [slice(1, -1), slice(1, -1), slice(1, -1)]
    n = f.shape[axis]
    idx[axis] = slice(1 + k, n - 1 + k)
    return f[tuple(idx)]


def interior_mask(shape, dtype=np.float32):
    """1 on the interior, 0 on the one-voxel boundary layer (solver.py:152-153)."""
    m = np.zeros(shape[:3], dtype=dtype)
    m[1:-1, 1:-1, 1:-1] = 1.0
    return m


def laplacian_center(field, dX, dY, dZ):
    """Compact 7-point Laplacian, zero on the boundary layer.

    Restates solver.py:686-695: per axis ``(f[i+1] - 2 f[i] + f[i-1]) / h**2``,
    axes summed x, y, z in that order, result in the input dtype.
    """
    out = np.zeros_like(field)
    c = _shift(field, 0, 0)
    acc = None
    for axis, h in ((0, dX), (1, dY), (2, dZ)):
        term = (_shift(field, axis, 1) - 2.0 * c + _shift(field, axis, -1)) / (h**2)
        acc = term if acc is None else acc + term
    out[1:-1, 1:-1, 1:-1] = acc
    return out


def divergence_center(fx, fy, fz, dX, dY, dZ):
    """2h central divergence, zero on the boundary layer (solver.py:708-717)."""
    dtype = np.result_type(fx.dtype, fy.dtype, fz.dtype)
    out = np.zeros_like(fx, dtype=dtype)
    acc = None
    for axis, f, h in ((0, fx, dX), (1, fy, dY), (2, fz, dZ)):
        term = (_shift(f, axis, 1) - _shift(f, axis, -1)) / (2.0 * h)
        acc = term if acc is None else acc + term
    out[1:-1, 1:-1, 1:-1] = acc
    return out


def gradient_center(phi, dX, dY, dZ):
    """2h central gradient, zero on the boundary layer (solver.py:720-729)."""
    out = []
    for axis, h in ((0, dX), (1, dY), (2, dZ)):
        g = np.zeros_like(phi)
        g[1:-1, 1:-1, 1:-1] = (_shift(phi, axis, 1) - _shift(phi, axis, -1)) / (2.0 * h)
        out.append(g)
    return tuple(out)


# ---------------------------------------------------------------------------
# Helmholtz-Hodge projection
# ---------------------------------------------------------------------------


def poisson_operator(x, dX, dY, dZ):
    """``A x = -div_c(grad_c(x))`` (solver.py:1024-1026)."""
    gx, gy, gz = gradient_center(x, dX, dY, dZ)
    return -divergence_center(gx, gy, gz, dX, dY, dZ)


def poisson_cg(rhs, dX, dY, dZ, mask_center, tol=1e-4, max_iters=200, info=None):
    """Conjugate gradients on ``A phi = rhs`` in float64 (solver.py:1010-1050).

    Stops (after updating phi and r) when ``sqrt(rs_new / rs0) < tol``; breaks
    before the update on a non-finite or non-positive ``p.Ap`` and after it on
    a non-finite ``r.r``. The result is cast back to ``rhs.dtype``.

    ``info`` (a dict, optional) receives ``iters`` (number of phi updates),
    ``relres`` (last ``sqrt(rs/rs0)``) and ``reason``.
    """
    eps = np.finfo(np.float64).eps
    rhs64 = rhs.astype(np.float64, copy=False)
    m64 = mask_center.astype(np.float64, copy=False)
    phi = np.zeros_like(rhs64)
    r = (rhs64 - poisson_operator(phi, dX, dY, dZ)) * m64
    p = r.copy()
    rs_old = np.sum(r * r)
    rs0 = max(rs_old, eps)
    iters = 0
    rel = float(np.sqrt(rs_old / rs0))
    reason = "max_iters"
    for _ in range(max_iters):
        ap = poisson_operator(p, dX, dY, dZ) * m64
        denom = np.sum(p * ap)
        if not np.isfinite(denom) or denom <= 0.0:
            reason = "denom"
            break
        alpha = rs_old / denom
        phi = (phi + alpha * p) * m64
        r = (r - alpha * ap) * m64
        iters += 1
        rs_new = np.sum(r * r)
        if not np.isfinite(rs_new):
            reason = "nonfinite"
            break
        rel = float(np.sqrt(rs_new / rs0))
        if rel < tol:
            reason = "converged"
            break
        beta = rs_new / max(rs_old, eps)
        p = (r + beta * p) * m64
        rs_old = rs_new
    if info is not None:
        info.update(iters=iters, relres=rel, reason=reason, rs0=float(rs0))
    return phi.astype(rhs.dtype, copy=False)


def project_vector_field(fx, fy, fz, dX, dY, dZ, tol, max_iters, dtype=np.float32, info=None):
    """Divergence-free projection of a force field (solver.py:644-683).

    Masks the field first, takes its divergence, returns the masked field
    unchanged when ``max|div| < 1e-12`` (solver.py:672-676), otherwise solves
    for phi by CG and returns ``mask * (f - grad_c phi)``.
    """
    shape = fx.shape
    m = interior_mask(shape, dtype)
    fx = fx.astype(dtype, copy=False) * m
    fy = fy.astype(dtype, copy=False) * m
    fz = fz.astype(dtype, copy=False) * m
    div_f = divergence_center(fx, fy, fz, dX, dY, dZ)
    div_f *= m
    if info is not None:
        info["max_div"] = float(np.max(np.abs(div_f)))
    if np.max(np.abs(div_f)) < 1e-12:
        if info is not None:
            info.update(iters=0, relres=0.0, reason="early_exit")
        return fx * m, fy * m, fz * m
    phi = poisson_cg(-div_f, dX, dY, dZ, m, tol=tol, max_iters=max_iters, info=info)
    gx, gy, gz = gradient_center(phi, dX, dY, dZ)
    return (fx - gx) * m, (fy - gy) * m, (fz - gz) * m


def project_body_force_to_shear(bz, dX, dY, dZ, tol=5e-5, max_iters=400, dtype=np.float32):
    """``project_vector_field(0, 0, bz)`` (solver.py:628-641)."""
    bz = bz.astype(dtype)
    z = np.zeros_like(bz)
    return project_vector_field(z, z.copy(), bz.copy(), dX, dY, dZ, tol, max_iters, dtype=dtype)


# ---------------------------------------------------------------------------
# input preparation
# ---------------------------------------------------------------------------


def prepare_body_component(comp, nx, ny, nz, n_steps, dtype=np.float32):
    """solver.py:1060-1069: None -> zeros, 3-D -> repeated over time, shape check."""
    if comp is None:
        return np.zeros((nx, ny, nz, n_steps), dtype=dtype)
    comp = np.asarray(comp, dtype=dtype)
    if comp.ndim == 3:
        comp = np.repeat(comp[..., np.newaxis], n_steps, axis=3)
    if comp.ndim != 4 or comp.shape != (nx, ny, nz, n_steps):
        raise ValueError("Body-force component has incompatible shape.")
    return comp


def prepare_material_field(field, nx, ny, nz, dtype=np.float32):
    """solver.py:1072-1079: scalar -> filled array, array -> cast + shape check."""
    if np.isscalar(field):
        return np.full((nx, ny, nz), float(field), dtype=dtype)
    field = np.asarray(field, dtype=dtype)
    if field.shape != (nx, ny, nz):
        raise ValueError("Material field has incompatible shape.")
    return field


# ---------------------------------------------------------------------------
# time loop
# ---------------------------------------------------------------------------


def shear_fdtd(
    bx,
    by,
    bz,
    rho,
    mu,
    dX,
    dY,
    dZ,
    dT,
    n_steps,
    opts=None,
    callback=None,
    dtype=np.float32,
    proj_info=None,
):
    """Explicit reduced Kelvin-Voigt time loop (solver.py:27-243).

    ``u_next = 2 u - u_prev + dT^2 (mu/rho lap u + eta/rho lap v + f/rho)``
    with ``v = mask (u - u_prev)/dT``, optional velocity damping
    ``u_next = u + (1 - vel_damp)(u_next - u)``, then the Dirichlet mask.
    Returns ``(u, (u - u_prev)/dT)``. ``dtype`` selects the oracle precision.
    ``proj_info`` (a list, optional) collects one CG info dict per projection.
    """
    ref = bx if bx is not None else (by if by is not None else bz)
    if ref is None:
        raise ValueError("At least one body-force component must be provided.")

    opts = opts or {}
    vel_damp = opts.get("vel_damp", 0.0)
    poisson_tol = opts.get("poisson_tol", 1e-5)
    poisson_max_iters = opts.get("poisson_max_iters", 200)
    reproject_every = int(opts.get("reproject_every", 0) or 0)
    if reproject_every < 0:
        raise ValueError("reproject_every must be a non-negative integer.")
    verbose = opts.get("verbose", False)
    progress_every = opts.get("progress_every", max(1, n_steps // 10))

    nx, ny, nz = ref.shape[:3]
    env_t = opts.get("env_t")
    if env_t is not None:
        env_t = np.asarray(env_t, dtype=dtype).ravel()
        if env_t.shape[0] != n_steps:
            raise ValueError(f"env_t length ({env_t.shape[0]}) must equal n_steps ({n_steps}).")
        zero = np.zeros((nx, ny, nz), dtype=dtype)
        b3 = [zero if b is None else np.asarray(b, dtype=dtype) for b in (bx, by, bz)]
        separable = True
    else:
        b4 = [prepare_body_component(b, nx, ny, nz, n_steps, dtype) for b in (bx, by, bz)]
        separable = False

    rho_f = prepare_material_field(rho, nx, ny, nz, dtype)
    mu_f = prepare_material_field(mu, nx, ny, nz, dtype)
    mu_over_rho = mu_f / rho_f
    inv_rho = 1.0 / rho_f

    eta = opts.get("eta", 0.0)
    if np.isscalar(eta):
        eta = None if float(eta) == 0.0 else prepare_material_field(float(eta), nx, ny, nz, dtype)
    elif eta is not None:
        eta = prepare_material_field(eta, nx, ny, nz, dtype)
    eta_over_rho = None if eta is None else eta / rho_f

    u_prev = np.zeros((nx, ny, nz, 3), dtype=dtype)
    u_curr = np.zeros_like(u_prev)
    mask = interior_mask((nx, ny, nz), dtype)
    mask4 = mask[..., None]
    dt2 = dT * dT

    initial_u = opts.get("initial_u")
    initial_v = opts.get("initial_v")
    if initial_u is not None:
        initial_u = np.asarray(initial_u, dtype=dtype)
        if initial_u.shape != u_curr.shape:
            raise ValueError("initial_u must match shape (nx,ny,nz,3)")
        u_curr = initial_u * mask4
    if initial_v is not None:
        initial_v = np.asarray(initial_v, dtype=dtype)
        if initial_v.shape != u_curr.shape:
            raise ValueError("initial_v must match shape (nx,ny,nz,3)")
        u_prev = u_curr - dT * initial_v
        u_prev *= mask4

    def project(fx, fy, fz):
        info = {} if proj_info is not None else None
        out = project_vector_field(
            fx.astype(dtype, copy=True),
            fy.astype(dtype, copy=True),
            fz.astype(dtype, copy=True),
            dX,
            dY,
            dZ,
            poisson_tol,
            poisson_max_iters,
            dtype=dtype,
            info=info,
        )
        if proj_info is not None:
            proj_info.append(info)
        return out

    if separable:
        f0 = project(*b3)

    for it in range(n_steps):
        if separable:
            s = env_t[it]
            f = (f0[0] * s, f0[1] * s, f0[2] * s)
        else:
            f = project(b4[0][..., it], b4[1][..., it], b4[2][..., it])

        acc = [mu_over_rho * laplacian_center(u_curr[..., c], dX, dY, dZ) + f[c] * inv_rho for c in range(3)]
        if eta_over_rho is not None:
            vel = (u_curr - u_prev) / dT
            vel *= mask4
            for c in range(3):
                acc[c] += eta_over_rho * laplacian_center(vel[..., c], dX, dY, dZ)

        u_next = np.empty_like(u_curr)
        for c in range(3):
            u_next[..., c] = 2.0 * u_curr[..., c] - u_prev[..., c] + dt2 * acc[c]
        if vel_damp > 0.0:
            u_next = u_curr + (1.0 - vel_damp) * (u_next - u_curr)
        u_next *= mask4

        if callback is not None:
            callback(it, u_next)
        if verbose and (it + 1) % progress_every == 0:
            print(f"FDTD it={it + 1}/{n_steps}")

        u_prev, u_curr = u_curr, u_next.copy()

        if reproject_every and (it + 1) % reproject_every == 0:
            for arr in (u_curr, u_prev):
                px, py, pz = project(arr[..., 0], arr[..., 1], arr[..., 2])
                arr[..., 0], arr[..., 1], arr[..., 2] = px, py, pz

    return u_curr.copy(), (u_curr - u_prev) / dT


def shear_fdtd_snapshots(bx, by, bz, rho, mu, dX, dY, dZ, dT, n_steps, opts, stride, dtype=np.float32):
    """``snapshot_stride`` semantics of solver.py:469-507 on the NumPy loop.

    Snapshot k is ``u`` after step ``(k+1)*stride - 1``, i.e. a callback firing
    when ``(it+1) % stride == 0`` (tests/test_snapshot_stride.py:204-219).
    Returns ``(u, v, snaps)`` with ``snaps`` shaped ``(n_steps//stride, nx, ny, nz, 3)``.
    """
    snaps = []

    def cb(it, u):
        if (it + 1) % stride == 0:
            snaps.append(np.array(u, copy=True))

    u, v = shear_fdtd(bx, by, bz, rho, mu, dX, dY, dZ, dT, n_steps, opts=opts, callback=cb, dtype=dtype)
    return u, v, np.stack(snaps, axis=0)
