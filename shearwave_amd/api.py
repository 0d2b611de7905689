"""Drop-in API of shearwave's hot path on MI355X.

Same names, arguments, option keys, return values and error messages as the
reference's ``src/shearwave/solver.py`` and ``src/shearwave/strain.py``
(cited per function). The arithmetic runs in the gfx950 HIP library; inputs
are NumPy arrays (or torch device tensors) and outputs are NumPy arrays in
the reference's dtypes, or device tensors when asked for.

Build-only options (namespaced ``sw_``):
  * ``sw_dtype``: storage type, ``"float32"`` (default, as the reference) or ``"float64"``;
  * ``sw_form``: ``"increment"`` (default) or ``"leapfrog"`` state storage;
  * ``sw_kernel``: ``"auto"``, ``"march2"`` (two steps per pass wherever no
    output falls between them; the increments equal single steps' bit for
    bit, and u is stored with one rounding per pass, u_n + (d_{n+1} +
    d_{n+2})), ``"march3"`` (three force-free float32 steps per pass where no
    output falls between them, two-step passes elsewhere; slower than
    ``"march2"`` on gfx950), ``"march"`` (one step per pass) or ``"direct"``
    time-step kernel; ``"auto"`` is ``"march2"`` on grids the march tiles fit;
  * ``sw_tile_y``, ``sw_tile_z``, ``sw_x_chunk``: march-kernel tile shape (0 = automatic);
  * ``sw_autotune``: with ``sw_kernel`` ``"auto"`` and no tile shape given,
    the two-step pass's tile rows and x chunk come from the autotuner
    (``shearwave_amd.autotune``), which times the candidates once per grid
    class on the solve's own buffers and caches the fastest (default True
    for loops of at least ``AUTOTUNE_MIN_VOXEL_STEPS`` voxel-steps); the
    results do not depend on the tile shape;
  * ``sw_cg_vectors``: storage of the projection's CG vectors r and p,
    ``"float32"`` or ``"float64"`` (default: ``sw_dtype``); phi, the dot
    products and the residual norm are float64 in both, as poisson_cg keeps
    them (solver.py:1020-1041);
  * ``sw_strain_stride``: with ``shear_fdtd(..., return_strain=True)``, the
    four strain invariants every that many steps, computed in-loop;
  * ``sw_resident_force``: with 4-D (per-step) body forces, upload each
    component once and take every step's slice on the device (default True;
    used when they take at most half of the free device memory), instead of
    slicing and uploading each step's force from the host;
  * ``sw_force_window``: with per-step forces and float32 storage, the
    forces of up to this many steps (default 16) are projected together in
    one batched CG before those steps run, where the plain CG serves them
    (1 = each step's on its own, as the float64 form always does);
  * ``sw_check_finite``: check u for inf/NaN every that many steps (default
    ``max(1, n_steps // 20)``, 0 = never) and warn (``RuntimeWarning``) with
    the first step found; the returned arrays are unchanged, as the
    reference returns them (an unstable dT fills them with inf/NaN).
"""

from __future__ import annotations

import warnings

import numpy as np
import torch

from . import engine as E


# Loops shorter than this many voxel-steps keep the default launch shape: the
# autotuner's first measurement of a grid class (about 30 passes per
# candidate) would cost more than it saves.
AUTOTUNE_MIN_VOXEL_STEPS = 1 << 30


def _is_dev(x):
    return isinstance(x, torch.Tensor)


def _np_out(t, dtype=None):
    a = E.to_host(t)
    return a if dtype is None else a.astype(dtype, copy=False)


def _prepare_material(field, nx, ny, nz):
    """Scalar or (nx, ny, nz) array check (solver.py:1072-1079); returns a float64 source."""
    if np.isscalar(field):
        return float(field)
    if _is_dev(field):
        if tuple(field.shape) != (nx, ny, nz):
            raise ValueError("Material field has incompatible shape.")
        return field
    arr = np.asarray(field)
    if arr.shape != (nx, ny, nz):
        raise ValueError("Material field has incompatible shape.")
    return arr


def _check_body(comp, nx, ny, nz, n_steps):
    """Shape check of _prepare_body_component (solver.py:1060-1069), without the repeat."""
    if comp is None:
        return None
    shape = tuple(comp.shape)
    if len(shape) == 3 and shape == (nx, ny, nz):
        return comp
    if len(shape) == 4 and shape == (nx, ny, nz, n_steps):
        return comp
    raise ValueError("Body-force component has incompatible shape.")


def _resident_body(comps, dtype):
    """The body-force components of a per-step solve on the device, each uploaded once in its
    source type (float32 or float64; others as `dtype`) and sliced per step there: the host's
    slice of step `it`, ``b[..., it]``, reads one value per nT (a cache line per value). Returns
    None when they would take more than half of the free device memory."""
    arrs = []
    need = 0
    for c in comps:
        if c is None or _is_dev(c):
            arrs.append(c)
            continue
        a = np.asarray(c)
        if a.dtype not in (np.float32, np.float64):
            a = a.astype(np.float32 if dtype == "float32" else np.float64)
        arrs.append(a)
        need += a.nbytes
    free, _total = torch.cuda.mem_get_info(E.device())
    if 2 * need > free:
        return None
    # a pageable copy (about 30 GB/s): a pinned staging block this large would first cost its allocation
    # (about 3.4 GB/s on first use)
    dev = E.device()
    return [a if a is None or _is_dev(a) else torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in arrs]


def _force_window(solve, opts, n_steps, tol, max_iters):
    """Steps whose per-step forces are projected together in one batched CG (1: each on its own).

    ``opts['sw_force_window']`` (default 16) caps it; the window is used where
    the batched CG serves the solve (float32 storage and CG vectors, the
    plain CG: ``Solve.batched_projection_fits``) and its buffers (about 44 B
    per voxel and step: the force, its projection and the CG vectors) take
    at most a quarter of the free device memory. Each projection is
    poisson_cg's as in the single-solve path, in another summation order.
    """
    w = min(int(opts.get("sw_force_window", 16) or 1), n_steps)
    if w <= 1 or not solve.batched_projection_fits(tol, max_iters):
        return 1
    free, _total = torch.cuda.mem_get_info(E.device())
    return max(1, min(w, free // (4 * 44 * solve.n)))


def _trace_flat(trace_indices, shape):
    if not isinstance(trace_indices, (tuple, list)) or len(trace_indices) != 3:
        raise ValueError("trace_indices must be a 3-tuple: (ix, iy, iz).")
    ix, iy, iz = (np.asarray(a, dtype=np.int64) for a in trace_indices)
    if not (ix.shape == iy.shape == iz.shape):
        raise ValueError("trace index arrays must share the same shape.")
    if ix.ndim != 1:
        raise ValueError("trace index arrays must be 1D.")
    idx = []
    for a, n in zip((ix, iy, iz), shape):
        a = np.where(a < 0, a + n, a)
        if np.any((a < 0) | (a >= n)):
            raise IndexError("trace index out of bounds")
        idx.append(a)
    return (idx[0] * shape[1] + idx[1]) * shape[2] + idx[2]


def shear_fdtd(  # noqa: PLR0912, PLR0915
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
    *,
    trace_indices=None,
    return_traces=False,
    return_device=False,
    snapshot_stride=None,
    return_strain=False,
):
    """3-D collocated shear FDTD with Kelvin-Voigt viscosity (solver.py:27-243) on the GPU.

    Also takes the keyword arguments of shear_fdtd_jax (solver.py:246-566):
    ``trace_indices`` / ``return_traces`` (u_z after every step at the given
    voxels, shape (n_traces, n_steps)), ``snapshot_stride`` (u after every
    stride steps, shape (n_steps//stride, nx, ny, nz, 3)) and
    ``return_device`` (torch tensors left in HBM instead of NumPy arrays).
    Returns (u, v), (u, v, snaps) or (u, v, traces); with ``return_strain`` a
    dict of in-loop strain invariants (every ``opts['sw_strain_stride']``
    steps, each (n_k, nx, ny, nz)) is appended.
    """
    ref = bx if bx is not None else (by if by is not None else bz)
    if ref is None:
        raise ValueError("At least one body-force component must be provided.")
    if return_traces and trace_indices is None:
        raise ValueError("trace_indices must be provided when return_traces=True.")
    if snapshot_stride is not None:
        if not isinstance(snapshot_stride, (int, np.integer)) or int(snapshot_stride) <= 0:
            raise ValueError("snapshot_stride must be a positive integer.")
        snapshot_stride = int(snapshot_stride)
        if n_steps % snapshot_stride != 0:
            raise ValueError(f"n_steps ({n_steps}) must be a multiple of snapshot_stride ({snapshot_stride}).")
        if return_traces:
            raise ValueError("return_traces and snapshot_stride are not both supported in v1.")

    opts = opts or {}
    vel_damp = opts.get("vel_damp", 0.0)
    poisson_tol = opts.get("poisson_tol", 1e-5)
    poisson_max_iters = opts.get("poisson_max_iters", 200)
    reproject_every = int(opts.get("reproject_every", 0) or 0)
    if reproject_every < 0:
        raise ValueError("reproject_every must be a non-negative integer.")
    verbose = opts.get("verbose", False)
    progress_every = opts.get("progress_every", max(1, n_steps // 10))
    dtype = E.dtype_name(opts.get("sw_dtype", "float32"))
    np_dtype = np.float32 if dtype == "float32" else np.float64
    form = opts.get("sw_form", "increment")
    strain_stride = int(opts.get("sw_strain_stride", 0) or 0)
    finite_stride = int(opts.get("sw_check_finite", max(1, n_steps // 20)) or 0)
    if return_strain and strain_stride <= 0:
        raise ValueError("return_strain needs opts['sw_strain_stride'] > 0.")

    nx, ny, nz = (int(s) for s in ref.shape[:3])
    shape = (nx, ny, nz)
    env_t = opts.get("env_t")
    if env_t is not None:
        env_t = np.asarray(env_t, dtype=np_dtype).ravel()
        if env_t.shape[0] != n_steps:
            raise ValueError(f"env_t length ({env_t.shape[0]}) must equal n_steps ({n_steps}).")
        comps = [None if b is None else b for b in (bx, by, bz)]
        per_step = False
    else:
        comps = [_check_body(b, nx, ny, nz, n_steps) for b in (bx, by, bz)]
        per_step = any(c is not None and len(c.shape) == 4 for c in comps)

    rho_f = _prepare_material(rho, nx, ny, nz)
    mu_f = _prepare_material(mu, nx, ny, nz)
    eta = opts.get("eta", 0.0)
    if np.isscalar(eta):
        eta = None if float(eta) == 0.0 else float(eta)
    elif eta is not None:
        eta = _prepare_material(eta, nx, ny, nz)

    initial_u = opts.get("initial_u")
    initial_v = opts.get("initial_v")
    for name, a in (("initial_u", initial_u), ("initial_v", initial_v)):
        if a is not None and tuple(a.shape) != shape + (3,):
            raise ValueError(f"{name} must match shape (nx,ny,nz,3)")

    trace_flat = None
    if trace_indices is not None:
        trace_flat = _trace_flat(trace_indices, shape)

    solve = E.Solve(shape, (dX, dY, dZ), float(dT), rho_f, mu_f, eta, dtype=dtype, form=form,
                    vel_damp=float(vel_damp), kernel=opts.get("sw_kernel", "auto"),
                    tile_y=int(opts.get("sw_tile_y", 0)), x_chunk=int(opts.get("sw_x_chunk", 0)),
                    tile_z=int(opts.get("sw_tile_z", 0)), cg_vectors=opts.get("sw_cg_vectors"))
    if (opts.get("sw_autotune", True) and opts.get("sw_kernel", "auto") == "auto" and not opts.get("sw_tile_y")
            and not opts.get("sw_x_chunk") and solve.two_step_fits()
            and n_steps * solve.n >= AUTOTUNE_MIN_VOXEL_STEPS):
        solve.autotune()
    if initial_u is not None or initial_v is not None:
        solve.set_initial(initial_u, initial_v)

    dev = E.device()
    tdt = E.torch_dtype(dtype)
    snaps = None
    if snapshot_stride:
        snaps = torch.empty((n_steps // snapshot_stride,) + shape + (3,), dtype=tdt, device=dev)
    inv = None
    if return_strain:
        inv = torch.empty((n_steps // strain_stride, 4) + shape, dtype=tdt, device=dev)
    tr_idx = tr_out = None
    if return_traces:
        tr_idx = torch.from_numpy(trace_flat).to(dev)
        tr_out = torch.zeros((trace_flat.shape[0], n_steps), dtype=tdt, device=dev)
    outs = dict(snap_stride=snapshot_stride or 0, snap_out=snaps, inv_stride=strain_stride if return_strain else 0,
                inv_out=inv, trace_idx=tr_idx, trace_out=tr_out, finite_stride=finite_stride)

    if per_step:
        env = np.ones(n_steps)
        if opts.get("sw_resident_force", True):
            comps = _resident_body(comps, dtype) or comps
    elif env_t is not None:
        env = env_t.astype(np.float64)
    else:
        env = np.ones(n_steps)  # 3-D forces without env_t: the same projected force every step

    if not per_step:
        if verbose and env_t is not None:
            print("Pre-computing Helmholtz projection (separable force)...")
        solve.set_force(*comps, poisson_tol, poisson_max_iters)

    window = 0
    if per_step:
        window = _force_window(solve, opts, n_steps, poisson_tol, poisson_max_iters)

    def step_force(t):
        return [None if c is None else (c[..., t] if len(c.shape) == 4 else c) for c in comps]

    if callback is None and not reproject_every and not per_step and not verbose:
        solve.run(n_steps, env, **outs)
    else:
        it = 0
        w_outs = None
        while it < n_steps:
            if per_step and window > 1:
                k = 1
                if it % window == 0:  # the next window's forces, projected in one batched CG
                    w_outs = None
                    w_outs, _ = solve.project_forces([step_force(t) for t in range(it, min(it + window, n_steps))],
                                                     poisson_tol, poisson_max_iters)
                solve.use_force(w_outs[it % window])
            elif per_step:
                k = 1
                solve.set_force(*step_force(it), poisson_tol, poisson_max_iters)
            elif callback is not None or verbose:
                k = 1
            else:  # reprojection only: run up to the next reprojection
                k = reproject_every - (it % reproject_every)
                k = min(k, n_steps - it)
            solve.run(k, env[it:it + k], **outs)
            it += k
            if callback is not None:
                callback(it - 1, _np_out(solve.u_aos()))
            if verbose and it % progress_every == 0:
                print(f"FDTD it={it}/{n_steps}")
            if reproject_every and it % reproject_every == 0:
                solve.reproject(poisson_tol, poisson_max_iters)

    u_aos, v_aos = solve.export()
    if finite_stride:
        bad = solve.first_nonfinite_step()
        if bad is not None:
            warnings.warn(f"shear_fdtd: u holds inf/NaN after step {bad} (unstable dT or inputs?)", RuntimeWarning,
                          stacklevel=2)
    v_dtype = np.float64 if (dtype == "float32" and isinstance(dT, np.float64)) else None
    if return_device:
        res = [u_aos, v_aos if v_dtype is None else v_aos.to(torch.float64)]
        if snaps is not None:
            res.append(snaps)
        if return_traces:
            res.append(tr_out)
    else:
        res = [_np_out(u_aos), _np_out(v_aos, v_dtype)]
        if snaps is not None:
            res.append(_np_out(snaps))
        if return_traces:
            res.append(_np_out(tr_out))
    if return_strain:
        keys = ("volumetric", "deviatoric", "max_shear", "von_mises")
        res.append({k: (inv[:, j] if return_device else _np_out(inv[:, j])) for j, k in enumerate(keys)})
    return tuple(res)


def shear_fdtd_jax(bx, by, bz, rho, mu, dX, dY, dZ, dT, n_steps, opts=None, callback=None, trace_indices=None,
                   return_traces=False, return_device=False, snapshot_stride=None):
    """shear_fdtd_jax's contract (solver.py:246-566), served by the same gfx950 engine.

    Keeps its restrictions: no callback, no reproject_every; ``opts['jit']`` is accepted and ignored.
    """
    if callback is not None:
        raise ValueError("callback is not supported in shear_fdtd_jax.")
    if (opts or {}).get("reproject_every"):
        raise NotImplementedError("reproject_every is only implemented in the NumPy solver (shear_fdtd).")
    return shear_fdtd(bx, by, bz, rho, mu, dX, dY, dZ, dT, n_steps, opts, None, trace_indices=trace_indices,
                      return_traces=return_traces, return_device=return_device, snapshot_stride=snapshot_stride)


# ------------------------------------------------------------------ projection


def project_vector_field(fx, fy, fz, dX, dY, dZ, tol, max_iters, *, dtype="float32", return_device=False, info=None,
                         cg_vectors=None):
    """Divergence-free projection (solver.py:644-683) on the GPU.

    Returns float32 arrays like the reference (``dtype="float64"`` keeps
    float64). phi, the CG dot products and the residual norm are float64 on
    the device; the CG vectors r and p are stored in ``cg_vectors`` (default:
    ``dtype``). ``info`` (a dict) receives iters, relres, max_div and the stop
    reason.
    """
    shape = tuple(int(s) for s in fx.shape)
    dev_in = _is_dev(fx)
    f = torch.empty((3,) + shape, dtype=E.torch_dtype(dtype), device=E.device())
    for c, comp in enumerate((fx, fy, fz)):
        f[c].copy_(E.to_device(comp, dtype, shape))
    rep = E.project_inplace(f, (dX, dY, dZ), tol, max_iters, vectors=cg_vectors)
    if info is not None:
        info.update(rep)
    if return_device or dev_in:
        return f[0], f[1], f[2]
    return tuple(_np_out(f[c]) for c in range(3))


def project_body_force_to_shear(bz, dX, dY, dZ, tol=5e-5, max_iters=400, **kw):
    """``project_vector_field(0, 0, bz)`` (solver.py:628-641)."""
    z = np.zeros(tuple(bz.shape), dtype=np.float32)
    return project_vector_field(z, z, bz, dX, dY, dZ, tol, max_iters, **kw)


# ---------------------------------------------------------- static elasticity


def _tol_is_f32(tol):
    """Whether ``rel < tol``, rel a float32 NumPy scalar, compares in float32.

    Under NumPy 2's promotion a Python float (or int) adapts to float32; a
    NumPy scalar keeps its own type, so a float64 one compares in float64.
    """
    if isinstance(tol, np.generic):
        return np.result_type(np.float32, tol.dtype) == np.float32
    return True


def elastic_solve_static(bz, dX, dY, dZ, mu, tol=1e-6, max_iters=1500, *, verbose=False, return_device=False,
                         info=None):
    """Solve mu * laplace(u) = -bz with zero Dirichlet boundary via CG (solver.py:569-625), on the GPU.

    As the reference: CG on the SPD form (-mu lap) u = bz, float32 vectors and
    scalars, the masked updates in the same order, and every dot product
    summed in NumPy's float32 order, so the result (and the iteration count)
    equals the reference's. Returns float32 u (a device tensor for device
    input or ``return_device``); ``info`` (a dict) receives iters, relres and
    the stop reason. ``verbose`` prints the reference's progress lines once
    the solve has finished.
    """
    import ctypes as C

    from ._abi import CG_REASONS, StaticInfo, check, lib

    nx, ny, nz = (int(s) for s in bz.shape)
    shape = (nx, ny, nz)
    mu_src = _prepare_material(mu, nx, ny, nz)
    rhs = E.to_device(bz, "float32", shape)
    mu_t = None if isinstance(mu_src, float) else E.to_device(mu_src, "float32", shape)
    n_it = max(int(max_iters), 0)  # range(max_iters)
    u = torch.empty(shape, dtype=torch.float32, device=E.device())
    ws = E.static_workspace(shape)
    hist = torch.empty(max(2 * n_it, 2), dtype=torch.float32, device=E.device()) if verbose else None
    rep = StaticInfo()
    check(
        lib().sw_elastic_solve_static(
            rhs.data_ptr(), None if mu_t is None else mu_t.data_ptr(), mu_src if mu_t is None else 0.0, nx, ny, nz,
            float(dX), float(dY), float(dZ), float(tol), int(_tol_is_f32(tol)), n_it, u.data_ptr(), ws.data_ptr(),
            ws.numel(), None if hist is None else hist.data_ptr(), C.byref(rep), E.stream_ptr(),
        )
    )
    iters = int(rep.iters)
    if verbose and iters:
        for it, (rs_new, rel) in enumerate(hist[: 2 * iters].view(-1, 2).cpu().numpy()):
            if it == 0 or (it + 1) % 25 == 0 or rel < tol:
                print(f"CG (GPU): it={it + 1}/{max_iters}, |r|={np.sqrt(rs_new):.3e}, rel={rel:.3e}")
    if info is not None:
        info.update(iters=iters, relres=float(rep.relres) if iters else None,
                    reason=CG_REASONS.get(int(rep.reason), str(rep.reason)))
    return u if (return_device or _is_dev(bz)) else _np_out(u)


# -------------------------------------------------------------------- stencils


_TORCH_TO_NP = {torch.float32: np.float32, torch.float64: np.float64, torch.float16: np.float16}


def _np_dtype(a):
    """NumPy dtype of an array or tensor (a device tensor is not copied)."""
    if _is_dev(a) or isinstance(a, torch.Tensor):
        return np.dtype(_TORCH_TO_NP.get(a.dtype, np.float64))
    return np.asarray(a).dtype


def _field_dtype(*arrays):
    dt = np.result_type(*[_np_dtype(a) for a in arrays])
    return "float64" if dt == np.float64 else ("float32" if dt == np.float32 else "float64")


def laplacian_center(field, dX, dY, dZ):
    """Compact 7-point Laplacian, 0 on the boundary layer (solver.py:686-695)."""
    from ._abi import check, lib

    dt = _field_dtype(field)
    shape = tuple(int(s) for s in field.shape)
    f = E.to_device(field, dt, shape)
    out = torch.empty_like(f)
    check(lib().sw_laplacian_center(E.sw_code(dt), f.data_ptr(), out.data_ptr(), *shape, dX, dY, dZ, E.stream_ptr()))
    return out if _is_dev(field) else _np_out(out)


def divergence_center(fx, fy, fz, dX, dY, dZ):
    """2h central divergence, 0 on the boundary layer (solver.py:708-717)."""
    from ._abi import check, lib

    dt = _field_dtype(fx, fy, fz)
    shape = tuple(int(s) for s in fx.shape)
    f = [E.to_device(a, dt, shape) for a in (fx, fy, fz)]
    out = torch.empty_like(f[0])
    check(lib().sw_divergence_center(E.sw_code(dt), *(a.data_ptr() for a in f), out.data_ptr(), *shape, dX, dY, dZ,
                                     E.stream_ptr()))
    return out if _is_dev(fx) else _np_out(out)


def gradient_center(phi, dX, dY, dZ):
    """2h central gradient, 0 on the boundary layer (solver.py:720-729)."""
    from ._abi import check, lib

    dt = _field_dtype(phi)
    shape = tuple(int(s) for s in phi.shape)
    p = E.to_device(phi, dt, shape)
    g = torch.empty((3,) + shape, dtype=p.dtype, device=p.device)
    check(lib().sw_gradient_center(E.sw_code(dt), p.data_ptr(), g[0].data_ptr(), g[1].data_ptr(), g[2].data_ptr(),
                                   *shape, dX, dY, dZ, E.stream_ptr()))
    if _is_dev(phi):
        return g[0], g[1], g[2]
    return tuple(_np_out(g[c]) for c in range(3))


def hann1d(n):
    """Hann window of length n as float32 (solver.py:1053-1057); a host utility."""
    if n <= 1:
        return np.ones((n,), dtype=np.float32)
    return np.hanning(n).astype(np.float32)


# ---------------------------------------------------------------------- strain


def _validate_strain_inputs(ux, uy, uz, dX, dY, dZ):
    """strain.py:138-159."""
    if ux.ndim != 3:
        raise ValueError(f"Displacement arrays must be 3-D, got ndim={ux.ndim}")
    if tuple(ux.shape) != tuple(uy.shape) or tuple(ux.shape) != tuple(uz.shape):
        raise ValueError(
            "All displacement components must share the same shape: "
            f"ux={tuple(ux.shape)}, uy={tuple(uy.shape)}, uz={tuple(uz.shape)}"
        )
    for name, val in (("dX", dX), ("dY", dY), ("dZ", dZ)):
        if val <= 0:
            raise ValueError(f"Grid spacing {name} must be positive, got {val}")


def compute_strain_tensor(ux, uy, uz, dX, dY, dZ):
    """Voigt strain [exx, eyy, ezz, gyz, gxz, gxy], engineering shear (strain.py:16-70)."""
    from ._abi import check, lib

    _validate_strain_inputs(ux, uy, uz, dX, dY, dZ)
    dt = _field_dtype(ux, uy, uz)
    shape = tuple(int(s) for s in ux.shape)
    u = [E.to_device(a, dt, shape) for a in (ux, uy, uz)]
    eps = torch.empty(shape + (6,), dtype=u[0].dtype, device=u[0].device)
    check(lib().sw_strain_tensor(E.sw_code(dt), *(a.data_ptr() for a in u), eps.data_ptr(), *shape, dX, dY, dZ,
                                 E.stream_ptr()))
    return eps if _is_dev(ux) else _np_out(eps)


def compute_strain_invariants(epsilon):
    """Volumetric, deviatoric, max-shear and von Mises strain (strain.py:73-135).

    von_mises is float64 for float32 input, as in the reference (NumPy 2
    promotes through the float64 scalar np.sqrt(2/3)).
    """
    from ._abi import check, lib

    if epsilon.shape[-1] != 6:
        raise ValueError(f"Expected last dimension == 6 (Voigt notation), got {epsilon.shape[-1]}")
    dt = _field_dtype(epsilon)
    lead = tuple(int(s) for s in epsilon.shape[:-1])
    e = E.to_device(epsilon, dt, lead + (6,))
    n = int(np.prod(lead)) if lead else 1
    outs = [torch.empty(lead, dtype=e.dtype, device=e.device) for _ in range(3)]
    vm = torch.empty(lead, dtype=torch.float64, device=e.device)
    check(lib().sw_strain_invariants(E.sw_code(dt), e.data_ptr(), n, *(o.data_ptr() for o in outs), vm.data_ptr(),
                                     E.stream_ptr()))
    res = {"volumetric": outs[0], "deviatoric": outs[1], "max_shear": outs[2], "von_mises": vm}
    if _is_dev(epsilon):
        return res
    return {k: _np_out(v) for k, v in res.items()}


# ---------------------------------------------------------- radiation force


def _gate_indices(gate, nt):
    """The time samples a gate selects, as NumPy fancy indexing on the time axis would (repeats kept)."""
    idx = np.arange(nt)[gate]
    return np.ascontiguousarray(np.atleast_1d(idx).astype(np.int32))


def _gate_args(idx):
    """(device indices or None, start, count): a contiguous ascending gate goes as a range, without indices."""
    n = int(idx.shape[0])
    if n == 0 or (n > 0 and np.all(np.diff(idx) == 1)):
        return None, (int(idx[0]) if n else 0), n
    return torch.from_numpy(np.ascontiguousarray(idx.astype(np.int32))).to(E.device()), 0, n


def _auto_gate(p_dev, n_vox, nt, dt):
    """Samples whose spatial mean of p^2 reaches 10% of its peak (radiation_force.py:66-72).

    The sums run on the device in float64; the mean, the threshold and the
    comparison run on the host in the pressure's dtype, as the reference does.
    """
    from ._abi import check, lib

    need = int(lib().sw_pressure_energy_workspace(n_vox, nt))
    work = torch.empty(max(need, 8) // 8, dtype=torch.float64, device=E.device())
    env = torch.empty(nt, dtype=torch.float64, device=E.device())
    check(lib().sw_pressure_energy(E.sw_code(dt), p_dev.data_ptr(), n_vox, nt, env.data_ptr(), work.data_ptr(),
                                   work.numel() * 8, E.stream_ptr()))
    npdt = np.float32 if dt == "float32" else np.float64
    mean = (env.cpu().numpy() / n_vox).astype(npdt)
    threshold = npdt(0.1) * mean.max()
    mask = mean >= threshold
    return np.arange(nt) if not mask.any() else np.nonzero(mask)[0]


def _radforce_inputs(pressure_4d, maps):
    if pressure_4d.ndim != 4:
        raise ValueError(f"pressure_4d must be 4-D (nx, ny, nz, nt), got ndim={pressure_4d.ndim}")
    shape3 = tuple(int(s) for s in pressure_4d.shape[:3])
    for m in maps:
        if tuple(m.shape) != shape3:
            raise ValueError(f"maps must have shape {shape3}, got {tuple(m.shape)}")
    return shape3, int(pressure_4d.shape[3])


def compute_radiation_force(pressure_4d, c_map, rho_map, alpha_coeff_map, f0, gate=None):
    """Acoustic radiation force b = 2 alpha_Np Isppa / c (radiation_force.py:27-83) on the GPU.

    Same arguments and gate semantics as the reference (a slice, an index
    array or None for the 10%-of-peak auto-gate); float32 inputs give float32,
    anything else float64. Time reductions run in float64 on the device.
    """
    from ._abi import check, lib

    shape3, nt = _radforce_inputs(pressure_4d, (c_map, rho_map, alpha_coeff_map))
    dt = _field_dtype(pressure_4d, c_map, rho_map, alpha_coeff_map)
    n_vox = int(np.prod(shape3))
    p = E.to_device(pressure_4d, dt, shape3 + (nt,))
    maps = [E.to_device(m, dt, shape3) for m in (c_map, rho_map, alpha_coeff_map)]
    idx = _auto_gate(p, n_vox, nt, dt) if gate is None else _gate_indices(gate, nt)
    g, g0, n_gate = _gate_args(idx)
    out = torch.empty(shape3, dtype=p.dtype, device=p.device)
    check(lib().sw_radiation_force(E.sw_code(dt), p.data_ptr(), n_vox, nt, None if g is None else g.data_ptr(), g0,
                                   n_gate, maps[0].data_ptr(), maps[1].data_ptr(), maps[2].data_ptr(), float(f0),
                                   out.data_ptr(), E.stream_ptr()))
    return out if _is_dev(pressure_4d) else _np_out(out)


def compute_radiation_force_velocity(pressure_4d, velocity_u_4d, velocity_w_4d, velocity_v_4d, c_map, rho_map,
                                     alpha_coeff_map, f0, gate=None):
    """Vector radiation force f_i = 2 alpha_Np <p v_i> / c (radiation_force.py:86-157) on the GPU.

    ``rho_map`` is accepted for interface parity and not used, as in the reference.
    """
    from ._abi import check, lib

    shape3, nt = _radforce_inputs(pressure_4d, (c_map, rho_map, alpha_coeff_map))
    for v in (velocity_u_4d, velocity_w_4d, velocity_v_4d):
        if tuple(v.shape) != tuple(pressure_4d.shape):
            raise ValueError(f"velocity fields must have shape {tuple(pressure_4d.shape)}, got {tuple(v.shape)}")
    dt = _field_dtype(pressure_4d, velocity_u_4d, velocity_w_4d, velocity_v_4d, c_map, rho_map, alpha_coeff_map)
    n_vox = int(np.prod(shape3))
    full = shape3 + (nt,)
    p = E.to_device(pressure_4d, dt, full)
    vel = [E.to_device(v, dt, full) for v in (velocity_u_4d, velocity_w_4d, velocity_v_4d)]
    c = E.to_device(c_map, dt, shape3)
    alpha = E.to_device(alpha_coeff_map, dt, shape3)
    idx = _auto_gate(p, n_vox, nt, dt) if gate is None else _gate_indices(gate, nt)
    g, g0, n_gate = _gate_args(idx)
    out = torch.empty((3,) + shape3, dtype=p.dtype, device=p.device)
    check(lib().sw_radiation_force_velocity(E.sw_code(dt), p.data_ptr(), *(v.data_ptr() for v in vel), n_vox, nt,
                                            None if g is None else g.data_ptr(), g0, n_gate, c.data_ptr(),
                                            alpha.data_ptr(), float(f0), out[0].data_ptr(), out[1].data_ptr(),
                                            out[2].data_ptr(), E.stream_ptr()))
    if _is_dev(pressure_4d):
        return out[0], out[1], out[2]
    return tuple(_np_out(out[k]) for k in range(3))



# ---------------------------------------------------------- acoustic strain


def _acoustic_inputs(pressure_4d, c_map, rho_map, gate):
    """Device record and maps in the result dtype, and the gate as (indices or None, start, count)."""
    shape3, nt = _radforce_inputs(pressure_4d, (c_map, rho_map))
    dt = _field_dtype(pressure_4d, c_map, rho_map)
    n_vox = int(np.prod(shape3))
    p = E.to_device(pressure_4d, dt, shape3 + (nt,))
    c = E.to_device(c_map, dt, shape3)
    rho = E.to_device(rho_map, dt, shape3)
    idx = _auto_gate(p, n_vox, nt, dt) if gate is None else _gate_indices(gate, nt)
    g, g0, n_gate = _gate_args(idx)
    return shape3, nt, dt, n_vox, p, c, rho, (None if g is None else g.data_ptr(), g0, n_gate, g)


def _ret(t, like):
    return t if _is_dev(like) else _np_out(t)


def compute_acoustic_strain_volumetric(pressure_4d, c_map, rho_map, gate=None):
    """epsilon = -p / K over the gated samples, (nx, ny, nz, nt_gated) (acoustic_strain.py:75-104), on the GPU."""
    from ._abi import check, lib

    shape3, nt, dt, n_vox, p, c, rho, (gp, g0, n_gate, _keep) = _acoustic_inputs(pressure_4d, c_map, rho_map, gate)
    out = torch.empty(shape3 + (n_gate,), dtype=p.dtype, device=p.device)
    check(lib().sw_acoustic_strain_volumetric(E.sw_code(dt), p.data_ptr(), n_vox, nt, gp, g0, n_gate, c.data_ptr(),
                                              rho.data_ptr(), out.data_ptr(), E.stream_ptr()))
    return _ret(out, pressure_4d)


def compute_acoustic_strain_peak(pressure_4d, c_map, rho_map, gate=None):
    """{"peak", "rms", "mean"} of the acoustic strain over the gated samples (acoustic_strain.py:107-151)."""
    from ._abi import check, lib

    shape3, nt, dt, n_vox, p, c, rho, (gp, g0, n_gate, _keep) = _acoustic_inputs(pressure_4d, c_map, rho_map, gate)
    if n_gate == 0:  # np.max over the empty sample axis (acoustic_strain.py:143)
        raise ValueError("zero-size array to reduction operation maximum which has no identity")
    out = torch.empty((3,) + shape3, dtype=p.dtype, device=p.device)
    check(lib().sw_acoustic_strain_peak(E.sw_code(dt), p.data_ptr(), n_vox, nt, gp, g0, n_gate, c.data_ptr(),
                                        rho.data_ptr(), out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
                                        E.stream_ptr()))
    return {k: _ret(out[j], pressure_4d) for j, k in enumerate(("peak", "rms", "mean"))}


def compute_acoustic_strain_gradient(pressure_4d, c_map, rho_map, dX, dY, dZ, gate=None):
    """Spatial gradient of -p_t / K for every gated sample (acoustic_strain.py:154-203), on the GPU.

    Each component is (nx, ny, nz, nt_gated) in the pressure's dtype, zero on the boundary layer.
    """
    from ._abi import check, lib

    shape3, nt, dt, n_vox, p, c, rho, (gp, g0, n_gate, _keep) = _acoustic_inputs(pressure_4d, c_map, rho_map, gate)
    out = torch.empty((3,) + shape3 + (n_gate,), dtype=p.dtype, device=p.device)
    check(lib().sw_acoustic_strain_gradient(E.sw_code(dt), p.data_ptr(), *shape3, nt, gp, g0, n_gate, c.data_ptr(),
                                            rho.data_ptr(), float(dX), float(dY), float(dZ), out[0].data_ptr(),
                                            out[1].data_ptr(), out[2].data_ptr(), E.stream_ptr()))
    p_dtype = _np_dtype(pressure_4d)
    res = [out[j] for j in range(3)]
    if _is_dev(pressure_4d):
        return tuple(res)
    return tuple(_np_out(r).astype(p_dtype, copy=False) for r in res)


def compute_acoustic_strain_gradient_peak(pressure_4d, c_map, rho_map, dX, dY, dZ, gate=None):
    """Float64 running statistics of the strain gradient over the gated samples (acoustic_strain.py:206-277)."""
    from ._abi import check, lib

    shape3, nt, dt, n_vox, p, c, rho, (gp, g0, n_gate, _keep) = _acoustic_inputs(pressure_4d, c_map, rho_map, gate)
    out = torch.empty((5,) + shape3, dtype=torch.float64, device=p.device)
    check(lib().sw_acoustic_strain_gradient_peak(E.sw_code(dt), p.data_ptr(), *shape3, nt, gp, g0, n_gate,
                                                 c.data_ptr(), rho.data_ptr(), float(dX), float(dY), float(dZ),
                                                 *(out[j].data_ptr() for j in range(5)), E.stream_ptr()))
    keys = ("peak_magnitude", "peak_grad_x", "peak_grad_y", "peak_grad_z", "rms_magnitude")
    return {k: _ret(out[j], pressure_4d) for j, k in enumerate(keys)}
