"""Gradients of sensor-plane displacement traces with respect to the mu and eta fields.

The forward model is the drop-in's time loop (increment form) with a fixed
projected force; its outputs are u_z after every step at a set of sensor
voxels (a plane for elastography). The adjoint is a hand-written kernel
(``sw_adjoint_step``); memory is bounded by sqrt(nT) checkpointing: the
forward pass keeps (u, d) every K ~ sqrt(nT) steps, and the backward pass
recomputes each segment from its checkpoint before sweeping the adjoint back
through it. The recomputation pairs steps into two-step passes as the forward
run does (``sw_fdtd_step2_mid`` also writes each pass's intermediate state),
so every recomputed state equals the forward run's bit for bit. The gradient
sums accumulate in float64.

:class:`SensorTraces` is a ``torch.autograd.Function``: ``traces =
SensorTraces.apply(mu, eta, problem)``, then any loss of ``traces`` can be
back-propagated to mu and eta.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _abi
from . import engine as E
from ._abi import check, lib, ptrs


class Problem:
    """Everything of a solve except mu and eta: grid, rho, force, envelope, sensors."""

    def __init__(self, shape, h, dt, rho, force, env, sensors, *, dtype="float32", vel_damp=0.0, tol=1e-10,
                 max_iters=20000, checkpoint_every=None, viscous=True):
        self.shape = tuple(int(s) for s in shape)
        self.h = tuple(float(v) for v in h)
        self.dt = float(dt)
        self.dtype = E.dtype_name(dtype)
        self.code = E.sw_code(self.dtype)
        self.env = np.asarray(env, dtype=np.float64)
        self.n_steps = int(self.env.shape[0])
        self.vel_damp = float(vel_damp)
        self.viscous = bool(viscous)
        self.K = int(checkpoint_every or max(1, math.isqrt(self.n_steps)))
        dev = E.device()
        self.rho = E.to_device(rho, "float64", self.shape)
        sens = np.asarray(sensors, dtype=np.int64).ravel()
        if np.any(sens < 0) or np.any(sens >= int(np.prod(self.shape))) or len(np.unique(sens)) != len(sens):
            raise ValueError("sensor indices must be unique flat voxel indices")
        self.sensors = torch.from_numpy(sens).to(dev)
        # the projected force is fixed (it depends on rho only): project it once
        eta0 = torch.ones(self.shape, dtype=torch.float64, device=dev) if self.viscous else None
        self.solve = E.Solve(self.shape, self.h, self.dt, self.rho, torch.ones_like(self.rho), eta0,
                             dtype=self.dtype, form="increment", vel_damp=self.vel_damp)
        self.cg = self.solve.set_force(*force, tol, max_iters)[0]

    # ----------------------------------------------------------------------
    def set_materials(self, mu, eta):
        s = self.solve
        mu64 = E.to_device(mu, "float64", self.shape)
        eta64 = E.to_device(eta, "float64", self.shape) if self.viscous else None
        check(lib().sw_prepare_coefficients(
            self.code, self.rho.data_ptr(), mu64.data_ptr(), None if eta64 is None else eta64.data_ptr(),
            s.m.data_ptr(), None if eta64 is None else s.e.data_ptr(), s.inv_rho.data_ptr(), s.n, E.stream_ptr()))

    def _state(self):
        return torch.stack(self.solve.u_components()), torch.stack(self.solve.b_components())

    def _restore(self, u, d):
        for c, t in enumerate(self.solve.u_components()):
            t.copy_(u[c])
        for c, t in enumerate(self.solve.b_components()):
            t.copy_(d[c])

    def forward(self, mu, eta, keep_checkpoints=True):
        """Traces (n_sensors, n_steps) of u_z after every step; checkpoints every K steps."""
        self.set_materials(mu, eta)
        s = self.solve
        s.step_count = 0
        s.set_initial(None, None)
        traces = torch.zeros((self.sensors.numel(), self.n_steps), dtype=E.torch_dtype(self.dtype),
                             device=self.rho.device)
        ckpts = []
        for n0 in range(0, self.n_steps, self.K):
            if keep_checkpoints:
                ckpts.append(self._state())
            k = min(self.K, self.n_steps - n0)
            s.run(k, self.env[n0:n0 + k], trace_idx=self.sensors, trace_out=traces)
        return traces, ckpts

    def gradient(self, ckpts, gtraces):
        """dJ/dmu and dJ/deta (float64) given dJ/dtraces, by the checkpointed adjoint sweep."""
        s = self.solve
        dev = self.rho.device
        tdt = E.torch_dtype(self.dtype)
        g = torch.as_tensor(gtraces, device=dev).to(tdt).contiguous()
        # the adjoint states, the recomputed forward states of a segment and the gradient sums, each set
        # carved from one allocation with its fields staggered across the L2 channels (E.FieldArena)
        adj = E.FieldArena(self.shape, tdt, 12)
        ub = [adj.take((3,)) for _ in range(2)]
        db = [adj.take((3,)) for _ in range(2)]
        sums = E.FieldArena(self.shape, torch.float64, 2)
        gm = sums.take()
        ge = sums.take() if self.viscous else None
        st = E.stream_ptr()
        n_sens = self.sensors.numel()
        N = self.n_steps

        def inject(buf, n):
            # dJ/du_n for n >= 1: traces column n-1 holds u_z after step n
            check(lib().sw_inject_traces(self.code, buf[2].data_ptr(), buf[2].numel(), self.sensors.data_ptr(), n_sens,
                                         g[:, n - 1].data_ptr() if g.stride(1) == 1 else g[:, n - 1].contiguous().data_ptr(),
                                         g.stride(0), st))

        cur = 0
        inject(ub[cur], N)
        # the forward states of one segment, each step written into its own slot
        segs = E.FieldArena(self.shape, tdt, 6 * self.K)
        seg_u = segs.take((self.K, 3))
        seg_d = segs.take((self.K, 3))
        for seg in range(len(ckpts) - 1, -1, -1):
            n0 = seg * self.K
            n1 = min(n0 + self.K, N)
            # recompute the forward states u_n, d_n for n = n0 .. n1-1, in two-step passes paired as the
            # forward run paired them (from the checkpoint); each pass also writes its intermediate state
            seg_u[0].copy_(ckpts[seg][0])
            seg_d[0].copy_(ckpts[seg][1])
            n = n0
            while n < n1 - 1:
                k = n - n0
                if n + 2 <= n1 - 1 and s.step2_into(seg_u[k], seg_d[k], seg_u[k + 1], seg_d[k + 1], seg_u[k + 2],
                                                    seg_d[k + 2], self.env[n:n + 2]):
                    n += 2
                else:
                    s.step_into(seg_u[k], seg_d[k], seg_u[k + 1], seg_d[k + 1], self.env[n])
                    n += 1
            for n in range(n1 - 1, n0 - 1, -1):
                u_n, d_n = seg_u[n - n0], seg_d[n - n0]
                nxt = 1 - cur
                check(lib().sw_adjoint_step(
                    self.code, ptrs(*u_n), ptrs(*d_n), ptrs(*ub[cur]), ptrs(*db[cur]), ptrs(*ub[nxt]),
                    ptrs(*db[nxt]), s.m.data_ptr(), None if ge is None else s.e.data_ptr(), gm.data_ptr(),
                    None if ge is None else ge.data_ptr(), *self.shape, *self.h, self.dt, self.vel_damp, st))
                cur = nxt
                if n >= 1:
                    inject(ub[cur], n)
        inv_rho = 1.0 / self.rho
        return gm * inv_rho, (ge * inv_rho if ge is not None else None)


class SensorTraces(torch.autograd.Function):
    """traces = SensorTraces.apply(mu, eta, problem): differentiable in mu and eta."""

    @staticmethod
    def forward(ctx, mu, eta, problem):
        traces, ckpts = problem.forward(mu, eta)
        ctx.problem = problem
        ctx.ckpts = ckpts
        ctx.dtypes = (mu.dtype, eta.dtype if isinstance(eta, torch.Tensor) else None)
        return traces

    @staticmethod
    def backward(ctx, gtraces):
        gm, ge = ctx.problem.gradient(ctx.ckpts, gtraces)
        ctx.ckpts = None
        gmu = gm.to(ctx.dtypes[0])
        geta = None if ge is None or ctx.dtypes[1] is None else ge.to(ctx.dtypes[1])
        return gmu, geta, None


def plane_sensors(shape, x):
    """Flat indices of the whole y-z plane at depth index x."""
    nx, ny, nz = shape
    yz = np.arange(ny * nz, dtype=np.int64)
    return x * ny * nz + yz
