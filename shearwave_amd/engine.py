"""Device-resident solve state and the host calls around the C ABI.

:class:`Solve` owns one solve's fields in HBM (SoA, C order, z fastest) as
torch tensors and drives the gfx950 kernels through ``_abi``. PyTorch only
allocates device memory and provides the stream; all arithmetic on the fields
happens in the HIP library. Without a visible GPU or without the built
library, everything here raises.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _abi
from ._abi import check, lib, ptrs

_DTYPES = {
    "float32": (torch.float32, _abi.SW_F32, np.float32),
    "float64": (torch.float64, _abi.SW_F64, np.float64),
}
_KERNELS = {"auto": _abi.SW_KERNEL_AUTO, "direct": _abi.SW_KERNEL_DIRECT, "march": _abi.SW_KERNEL_MARCH,
            "march2": _abi.SW_KERNEL_MARCH2, "march3": _abi.SW_KERNEL_MARCH3}
_FORMS = {"increment": _abi.SW_FORM_INCREMENT, "leapfrog": _abi.SW_FORM_LEAPFROG}


def dtype_name(dtype) -> str:
    if isinstance(dtype, str):
        name = dtype
    elif isinstance(dtype, torch.dtype):
        name = str(dtype).replace("torch.", "")
    else:
        name = np.dtype(dtype).name
    if name not in _DTYPES:
        raise ValueError(f"storage dtype must be float32 or float64, got {dtype!r}")
    return name


def torch_dtype(name):
    return _DTYPES[dtype_name(name)][0]


def sw_code(name):
    return _DTYPES[dtype_name(name)][1]


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("shearwave_amd runs on a ROCm GPU; torch.cuda.is_available() is False")
    lib()  # fail loudly if the HIP library is missing
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr():
    return torch.cuda.current_stream().cuda_stream


def to_host(t: torch.Tensor) -> np.ndarray:
    """A device tensor as a NumPy array, through pinned host memory.

    Device-to-host copies into pageable memory run at about 4 GB/s on this
    platform, into pinned memory at about 57 GB/s. The array is a view of the
    pinned tensor, which stays alive with it; torch's host allocator caches
    the pinned blocks for reuse.
    """
    t = t.detach()
    if t.device.type != "cuda":
        return t.numpy()
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t)
    return h.numpy()


def to_device(x, dtype, shape=None) -> torch.Tensor:
    """Contiguous device tensor of `dtype` from a NumPy array, scalar or tensor."""
    dev = device()
    tdt = torch_dtype(dtype)
    if isinstance(x, torch.Tensor):
        t = x.to(device=dev, dtype=tdt).contiguous()
    elif np.isscalar(x):
        if shape is None:
            raise ValueError("scalar needs a shape")
        t = torch.full(tuple(shape), float(x), dtype=tdt, device=dev)
    else:
        a = np.asarray(x)
        if a.dtype in (np.float32, np.float64):
            # upload in the source type, convert on the device (the same IEEE rounding as a host cast)
            t = torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(tdt)
        else:
            t = torch.from_numpy(np.ascontiguousarray(a, dtype=_DTYPES[dtype_name(dtype)][2])).to(dev)
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
    return t


def upload_async(a: np.ndarray) -> torch.Tensor:
    """A host array on the device through pinned memory, queued on the current stream without
    waiting for it (a pageable copy would wait for the stream's queued work first)."""
    a = np.ascontiguousarray(a)
    if not a.flags.writeable:
        a = a.copy()
    h = torch.from_numpy(a).pin_memory()
    return h.to(device(), non_blocking=True)


# Byte shift between consecutive per-voxel fields of one solve. Fields of one
# size allocated apart start at addresses that agree modulo the allocator's
# granularity, so a kernel reading many of them at one voxel (the two-step
# pass reads eight and writes six) sends every stream to the same L2 channel
# at the same time; shifting each field by a further 9 x 128 B spreads them
# over the channels (the 256^3 float32 two-step pass: 230 -> 221 us per
# launch, measured against fields exactly 2^26 B apart).
FIELD_STAGGER = 1152


class FieldArena:
    """Per-voxel fields of one grid carved from one zeroed device allocation.

    Field k starts k * step elements after the first (256-byte aligned) one,
    step = the field's size rounded up to 128 B plus FIELD_STAGGER, so every
    field is 128-byte aligned and consecutive fields sit on different L2
    channels at the same voxel. take(lead) returns the next prod(lead) fields
    as one tensor of shape lead + grid (each field contiguous).
    """

    def __init__(self, shape, dtype, n_fields):
        tdt = dtype if isinstance(dtype, torch.dtype) else torch_dtype(dtype)
        item = torch.empty((), dtype=tdt).element_size()
        self.shape = tuple(int(v) for v in shape)
        n = int(np.prod(self.shape))
        self.step = (-(-n * item // 128) * 128 + FIELD_STAGGER) // item
        self.buf = torch.zeros(n_fields * self.step + 256 // item, dtype=tdt, device=device())
        self.off = ((-self.buf.data_ptr()) % 256) // item
        self.left = int(n_fields)

    def take(self, lead=()):
        lead = tuple(int(v) for v in lead)
        cnt = int(np.prod(lead)) if lead else 1
        if cnt > self.left:
            raise ValueError("field arena exhausted")
        strides, st = [], self.step
        for d in reversed(lead):
            strides.insert(0, st)
            st *= d
        _, ny, nz = self.shape
        t = torch.as_strided(self.buf, lead + self.shape, tuple(strides) + (ny * nz, nz, 1), self.off)
        self.off += cnt * self.step
        self.left -= cnt
        return t


_WORKSPACES: dict = {}


def _workspace(kind, shape, nbytes_fn) -> torch.Tensor:
    """Cached device workspace of one kind for a grid shape.

    One per device, shape and stream: calls on one stream are ordered, so a
    workspace (which also holds the CG state the host polls) is never used by
    two calls at once; calls on different streams get different workspaces.
    """
    key = (kind, torch.cuda.current_device(), tuple(shape), stream_ptr())
    ws = _WORKSPACES.get(key)
    if ws is None:
        nbytes = int(nbytes_fn(*shape))
        if nbytes <= 0:
            raise ValueError(f"bad grid {shape}")
        ws = torch.empty(nbytes, dtype=torch.uint8, device=device())
        _WORKSPACES[key] = ws
    return ws


def project_workspace(shape) -> torch.Tensor:
    """Cached CG workspace of the projection for a grid shape."""
    return _workspace("project", shape, lib().sw_project_workspace_bytes)


def static_workspace(shape) -> torch.Tensor:
    """Cached workspace of the static elasticity solve for a grid shape."""
    return _workspace("static", shape, lib().sw_static_workspace_bytes)


def project_inplace(f, spacing, tol, max_iters, *, scale_out=None, inv_rho=None, scale=1.0, vectors=None) -> dict:
    """Helmholtz-Hodge projection of f (3, nx, ny, nz) in place on the device.

    Optionally writes scale * f_S * inv_rho into scale_out (3, nx, ny, nz).
    ``vectors`` is the storage of the CG vectors r and p (``"float32"`` or
    ``"float64"``; default: f's dtype); phi, the dot products and the residual
    norm are float64 either way. Returns the CG report: iters, relres,
    max_div, reason and true_relres (|b - A phi| / |b| of the returned phi,
    b formed in float64; with float32 vectors the CG replaces its recursive
    residual by it until it meets tol).
    """
    shape = tuple(f.shape[1:])
    ws = project_workspace(shape)
    info = _abi.CgInfo()
    so = ptrs(*scale_out) if scale_out is not None else None
    vec = sw_code(f.dtype if vectors is None else vectors)
    check(
        lib().sw_project_vector_field(
            sw_code(f.dtype), vec, f[0].data_ptr(), f[1].data_ptr(), f[2].data_ptr(), *shape, *spacing, float(tol),
            int(max_iters), ws.data_ptr(), ws.numel(), so, None if inv_rho is None else inv_rho.data_ptr(),
            float(scale), C.byref(info), stream_ptr(),
        )
    )
    return {
        "iters": int(info.iters),
        "relres": float(info.relres),
        "max_div": float(info.max_div),
        "reason": _abi.CG_REASONS.get(int(info.reason), str(info.reason)),
        "true_relres": float(info.true_relres),
    }


class Solve:
    """One shear_fdtd solve resident in HBM.

    Fields: mu/rho, eta/rho (None when inviscid), 1/rho, the state (u, b) and
    its next-step buffers, and the loop's force dT^2 f_S / rho. The state form
    is ``increment`` (u, d = u - u_prev) by default: in float32 storage it
    keeps the velocity to float32 precision (the reference's (u, u_prev) form
    loses it to cancellation); ``leapfrog`` stores (u, u_prev).
    """

    def __init__(self, shape, spacing, dt, rho, mu, eta=None, *, dtype="float32", form="increment",
                 vel_damp=0.0, kernel="auto", tile_y=0, x_chunk=0, tile_z=0, cg_vectors=None):
        self.shape = tuple(int(s) for s in shape)
        self.spacing = tuple(float(h) for h in spacing)
        self.dt = float(dt)
        self.dtype = dtype_name(dtype)
        self.code = sw_code(self.dtype)
        # storage of the projection's CG vectors r and p (phi and the reductions stay float64)
        self.cg_vectors = self.dtype if cg_vectors is None else dtype_name(cg_vectors)
        self.form = form
        self.n = int(np.prod(self.shape))
        tdt = torch_dtype(self.dtype)
        dev = device()
        st = stream_ptr()
        rho64 = to_device(rho, "float64", self.shape)
        mu64 = to_device(mu, "float64", self.shape)
        eta64 = None if eta is None else to_device(eta, "float64", self.shape)
        # every per-voxel field in one allocation, staggered across the L2 channels (FieldArena):
        # mu/rho, eta/rho, 1/rho, the state and next-state buffers, the loop force
        n_state = 4 if form == "increment" else 3
        arena = FieldArena(self.shape, tdt, 3 + 3 * n_state + 3)
        self.m = arena.take()
        self.e = None if eta is None else arena.take()
        self.inv_rho = arena.take()
        check(lib().sw_prepare_coefficients(
            self.code, rho64.data_ptr(), mu64.data_ptr(), None if eta64 is None else eta64.data_ptr(),
            self.m.data_ptr(), None if self.e is None else self.e.data_ptr(), self.inv_rho.data_ptr(), self.n, st))
        self._bufs = [arena.take((3,)) for _ in range(n_state)]
        self._force_slot = arena.take((3,))
        self._by_ptr = {}
        for t in self._bufs:
            for c in range(3):
                self._by_ptr[t[c].data_ptr()] = t[c]
        self.force = None
        self.step_count = 0
        # first global step after which a non-finite check found inf/NaN in u (-1: none)
        self._nonfinite = torch.full((1,), -1, dtype=torch.int64, device=dev)
        s = _abi.FdtdState()
        s.dtype = self.code
        s.form = _FORMS[form]
        s.nx, s.ny, s.nz = self.shape
        s.dx, s.dy, s.dz = self.spacing
        s.dt = self.dt
        s.vel_damp = float(vel_damp)
        s.mu_over_rho = self.m.data_ptr()
        s.eta_over_rho = None if self.e is None else self.e.data_ptr()
        for c in range(3):
            s.force[c] = None
            s.u[c] = self._bufs[0][c].data_ptr()
            s.b[c] = self._bufs[1][c].data_ptr()
            s.u_next[c] = self._bufs[2][c].data_ptr()
            s.b_next[c] = self._bufs[3][c].data_ptr() if form == "increment" else None
        s.kernel = _KERNELS[kernel]
        s.tile_y = int(tile_y)
        s.x_chunk = int(x_chunk)
        s.tile_z = int(tile_z)
        self.state = s

    def two_step_fits(self):
        """Whether this solve's loop runs two-step passes whose launch shape the autotuner can pick."""
        nx, ny, nz = self.shape
        return (self.form == "increment" and self.state.kernel in (_KERNELS["auto"], _KERNELS["march2"])
                and nz >= 32 and ny >= 8 and nx >= 8 and ny * nz < (1 << 28))

    def autotune(self, tuner=None):
        """Tile rows and x chunk of the two-step pass from the autotuner, which times the candidates on this
        solve's own buffers once per grid class (reading the state and writing only the next-step buffers)
        and caches the fastest. The state is unchanged, and every candidate computes each voxel with the
        same arithmetic, so results do not depend on the choice. Returns (tile rows, x chunk)."""
        from .autotune import default_tuner

        if not self.two_step_fits():
            raise ValueError("the autotuner picks the launch shape of the two-step pass, which this solve does not run")
        ty, xc = (tuner or default_tuner()).pass_shape(self)
        self.state.tile_y, self.state.x_chunk = int(ty), int(xc)
        return int(ty), int(xc)

    # -------------------------------------------------------------- inputs
    def set_force(self, fx, fy, fz, tol, max_iters):
        """Project the force pattern (solver.py:189-192) and keep dT^2 f_S / rho for the loop."""
        f = torch.empty((3,) + self.shape, dtype=torch_dtype(self.dtype), device=device())
        for c, comp in enumerate((fx, fy, fz)):
            if comp is None:
                f[c].zero_()
            else:
                f[c].copy_(to_device(comp, self.dtype, self.shape))
        self.use_force(self._force_slot)
        info = project_inplace(f, self.spacing, tol, max_iters, scale_out=self.force, inv_rho=self.inv_rho,
                               scale=self.dt * self.dt, vectors=self.cg_vectors)
        return info, f

    def batched_projection_fits(self, tol, max_iters):
        """Whether :meth:`project_forces` serves this solve's projections: float32 storage and CG vectors
        (the batched CG's), where the single-solve projection would run the plain CG too (not the
        spectral-preconditioned one, which converges in a few iterations)."""
        return (self.dtype == "float32" and self.cg_vectors == "float32"
                and not lib().sw_project_preconditioned(_abi.SW_F32, *self.shape, float(tol), int(max_iters)))

    def project_forces(self, forces, tol, max_iters):
        """Project several force patterns in one batched CG (``batch.batched_projection``), each as
        :meth:`set_force` would (solver.py:189-192: mask, divergence, poisson_cg, gradient subtraction; the
        batched CG's float32 vectors and float64 reductions). ``forces``: a list of (fx, fy, fz) as set_force
        takes them. Returns (outs, infos): outs[k] (3, nx, ny, nz) is pattern k's dT^2 f_S / rho, for
        :meth:`use_force`."""
        from .batch import ProjectionJob, batched_projection

        if self.dtype != "float32":
            raise ValueError("the batched projection runs on float32 storage")
        k = len(forces)
        f = torch.empty((k, 3) + self.shape, dtype=torch.float32, device=device())
        outs = torch.empty((k, 3) + self.shape, dtype=torch.float32, device=device())
        for j, comps in enumerate(forces):
            for c, comp in enumerate(comps):
                if comp is None:
                    f[j, c].zero_()
                else:
                    f[j, c].copy_(to_device(comp, self.dtype, self.shape))
        jobs = [ProjectionJob(shape=self.shape, h=self.spacing, f=f[j], out=outs[j], inv_rho=self.inv_rho,
                              scale=self.dt * self.dt, tol=float(tol), max_iters=int(max_iters)) for j in range(k)]
        infos = batched_projection(jobs)
        return outs, infos

    def use_force(self, f):
        """The loop's force: f (3, nx, ny, nz) of this solve's storage type, dT^2 f_S / rho (kept referenced)."""
        if self.force is not f:
            self.force = f
            for c in range(3):
                self.state.force[c] = f[c].data_ptr()

    def set_initial(self, u0=None, v0=None):
        """Initial state from AoS (nx, ny, nz, 3) arrays (solver.py:157-171)."""
        aos = self.shape + (3,)
        u0d = None if u0 is None else to_device(u0, self.dtype, aos)
        v0d = None if v0 is None else to_device(v0, self.dtype, aos)
        u = [self.state.u[c] for c in range(3)]
        b = [self.state.b[c] for c in range(3)]
        check(lib().sw_import_state(
            self.code, self.state.form, None if u0d is None else u0d.data_ptr(),
            None if v0d is None else v0d.data_ptr(), self.dt, (C.c_void_p * 3)(*u), (C.c_void_p * 3)(*b),
            *self.shape, stream_ptr()))

    # ----------------------------------------------------------------- loop
    def run(self, n_steps, env=None, *, snap_stride=0, snap_out=None, inv_stride=0, inv_out=None,
            trace_idx=None, trace_out=None, finite_stride=0, inv_separate=False):
        """Advance n_steps steps; env (host, length n_steps) scales the force per step.

        With ``finite_stride`` > 0, u is checked for inf/NaN every that many
        global steps (after the pass that reaches the step), without
        synchronising; :meth:`first_nonfinite_step` reads the result. The
        in-loop strain invariants (``inv_stride``) are computed by the pass
        after their step where that is a force-free two-step pass, else by
        their own kernel; ``inv_separate`` always uses their own kernel (the
        same values).
        """
        n_steps = int(n_steps)
        if n_steps <= 0:
            return
        env_arr = None
        if env is not None:
            env_np = np.ascontiguousarray(np.asarray(env, dtype=np.float64).ravel())
            if env_np.shape[0] != n_steps:
                raise ValueError("env length must equal n_steps")
            env_arr = env_np.ctypes.data_as(C.POINTER(C.c_double))
        out = _abi.FdtdOutputs()
        out.step0 = self.step_count
        out.snap_stride = int(snap_stride)
        out.snap_out = None if snap_out is None else snap_out.data_ptr()
        out.inv_stride = int(inv_stride)
        out.inv_out = None if inv_out is None else inv_out.data_ptr()
        if trace_idx is not None and trace_idx.numel() > 0:
            out.trace_idx = trace_idx.data_ptr()
            out.n_traces = trace_idx.numel()
            out.trace_out = trace_out.data_ptr()
            out.trace_ld = trace_out.shape[1]
        if finite_stride:
            out.finite_stride = int(finite_stride)
            out.nonfinite_step = self._nonfinite.data_ptr()
        out.inv_mode = 1 if inv_separate else 0
        check(lib().sw_fdtd_run(C.byref(self.state), n_steps, env_arr, C.byref(out), stream_ptr()))
        self.step_count += n_steps

    def first_nonfinite_step(self):
        """The global step after which a check of :meth:`run` first found inf/NaN
        in u, or None (synchronises with the stream)."""
        k = int(self._nonfinite.item())
        return None if k < 0 else k

    def step_into(self, u_src, d_src, u_dst, d_dst, env_value=None):
        """One step from caller buffers (u_src, d_src) into (u_dst, d_dst), each
        (3, nx, ny, nz) in the storage type, increment form, without touching
        the solve's own state (the adjoint's recomputation writes each state of
        a segment straight into its slot)."""
        assert self.form == "increment"
        st = self.state
        keep = [list(getattr(st, f)) for f in ("u", "b", "u_next", "b_next")]
        try:
            for c in range(3):
                st.u[c] = u_src[c].data_ptr()
                st.b[c] = d_src[c].data_ptr()
                st.u_next[c] = u_dst[c].data_ptr()
                st.b_next[c] = d_dst[c].data_ptr()
            env_arr = None
            if env_value is not None:
                env_np = np.asarray([float(env_value)], dtype=np.float64)
                env_arr = env_np.ctypes.data_as(C.POINTER(C.c_double))
            out = _abi.FdtdOutputs()
            check(lib().sw_fdtd_run(C.byref(st), 1, env_arr, C.byref(out), stream_ptr()))
        finally:
            for f, vals in zip(("u", "b", "u_next", "b_next"), keep):
                arr = getattr(st, f)
                for c in range(3):
                    arr[c] = vals[c]

    def step2_into(self, u_src, d_src, u_mid, d_mid, u_dst, d_dst, env_pair=None):
        """Two steps from (u_src, d_src) into (u_dst, d_dst) in one two-step pass that also writes the state
        after the first step into (u_mid, d_mid), bit for bit what step_into writes (the adjoint's
        recomputation). Returns False, doing nothing, on grids the two-step pass does not take."""
        assert self.form == "increment"
        st = self.state
        keep = [list(getattr(st, f)) for f in ("u", "b", "u_next", "b_next")]
        try:
            for c in range(3):
                st.u[c] = u_src[c].data_ptr()
                st.b[c] = d_src[c].data_ptr()
                st.u_next[c] = u_dst[c].data_ptr()
                st.b_next[c] = d_dst[c].data_ptr()
            env_arr = None
            if env_pair is not None:
                env_np = np.asarray(env_pair, dtype=np.float64).reshape(2)
                env_arr = env_np.ctypes.data_as(C.POINTER(C.c_double))
            rc = lib().sw_fdtd_step2_mid(C.byref(st), env_arr, _abi.ptrs(*u_mid), _abi.ptrs(*d_mid), stream_ptr())
            if rc == _abi.SW_ERR_UNSUPPORTED:
                return False
            check(rc)
            return True
        finally:
            for f, vals in zip(("u", "b", "u_next", "b_next"), keep):
                arr = getattr(st, f)
                for c in range(3):
                    arr[c] = vals[c]

    # --------------------------------------------------------------- access
    def _vec(self, ptr_array):
        return [self._by_ptr[ptr_array[c]] for c in range(3)]

    def u_components(self):
        return self._vec(self.state.u)

    def b_components(self):
        return self._vec(self.state.b)

    def export(self, u_aos=None, v_aos=None):
        """u and v = (u - u_prev)/dT as AoS device tensors (solver.py:241-243)."""
        aos = self.shape + (3,)
        tdt = torch_dtype(self.dtype)
        if u_aos is None:
            u_aos = torch.empty(aos, dtype=tdt, device=device())
        if v_aos is None:
            v_aos = torch.empty(aos, dtype=tdt, device=device())
        check(lib().sw_export_state(
            self.code, self.state.form, (C.c_void_p * 3)(*self.state.u), (C.c_void_p * 3)(*self.state.b),
            1.0 / self.dt, u_aos.data_ptr(), v_aos.data_ptr(), self.n, stream_ptr()))
        return u_aos, v_aos

    def u_aos(self, out=None):
        aos = self.shape + (3,)
        if out is None:
            out = torch.empty(aos, dtype=torch_dtype(self.dtype), device=device())
        u = self.state.u
        check(lib().sw_soa_to_aos3(self.code, u[0], u[1], u[2], out.data_ptr(), self.n, stream_ptr()))
        return out

    def reproject(self, tol, max_iters):
        """Project u and u_prev to divergence-free (solver.py:236-239).

        As the reference, u and u_prev are projected separately (each CG stops
        on its own residual). In the increment form of float32 storage the
        state is (u, d) with |d| << |u|, and d' = P(u) - P(u - d) cancels: so
        u_prev = u - d (exact in float64), both projections (float64 fields
        and CG vectors) and the difference all run in float64, and d' is
        rounded to float32 once.
        """
        u = self.u_components()
        b = self.b_components()
        if self.form == "increment":
            f64 = self.dtype == "float64"
            u64 = torch.stack(u).to(torch.float64)  # copies: the projection works in place
            b64 = torch.stack(b).to(torch.float64)
            prev = torch.empty_like(u64)
            check(lib().sw_combine3(_abi.SW_F64, ptrs(*u64), ptrs(*b64), 1.0, -1.0, ptrs(*prev), self.n,
                                    stream_ptr()))
            infos = [project_inplace(v, self.spacing, tol, max_iters, vectors=None if f64 else "float64")
                     for v in (u64, prev)]
            d64 = torch.empty_like(u64)
            check(lib().sw_combine3(_abi.SW_F64, ptrs(*u64), ptrs(*prev), 1.0, -1.0, ptrs(*d64), self.n,
                                    stream_ptr()))
            for c in range(3):
                u[c].copy_(u64[c])
                b[c].copy_(d64[c])
            return infos
        uvec = torch.stack(u)
        prev = torch.stack(b)
        infos = [project_inplace(v, self.spacing, tol, max_iters, vectors=self.cg_vectors) for v in (uvec, prev)]
        for c in range(3):
            u[c].copy_(uvec[c])
            b[c].copy_(prev[c])
        return infos
