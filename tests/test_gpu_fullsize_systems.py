"""Batching, serving and differentiation on their full-size workloads.

The small-grid tests hold each system against the float64 oracle; these run
the measured workloads themselves (the 48-solve batch, full-size serving
requests, the 128^3 nT = 800 gradient) and check them against the float64
form of the plain drop-in, which the parity tests hold within 1e-10 of the
oracle, and, for the gradient, against central differences.
"""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def plain(sw, sp, dtype):
    """The plain drop-in on one solve spec (host arrays out)."""
    opts = {"env_t": sp.env, "eta": sp.eta, "vel_damp": sp.vel_damp, "poisson_tol": sp.tol,
            "poisson_max_iters": sp.max_iters, "sw_dtype": dtype}
    return sw.shear_fdtd(None, None, sp.force[2], sp.rho, sp.mu, *sp.h, sp.dt, sp.n_steps, opts)


def test_batch_48_solves_full_size(sw):
    """The 48-solve batching workload (grids U{64..192}^2 x U{64..256}, nT U{500..2000}): every result finite,
    the smallest, largest, longest and two more solves within 1e-5 of the float64 drop-in on u and v, and two
    solves within 1e-6 of the float32 drop-in."""
    from shearwave_amd import workloads as W
    from shearwave_amd.batch import Batch

    specs = W.batch_48()
    b = Batch(specs)
    info = b.project()
    assert all(i["reason"] == "converged" for i in info), info
    b.run(collect=True)
    for k in range(len(specs)):
        u, v = b.result(k)
        assert u.shape == specs[k].shape + (3,) and np.isfinite(u).all() and np.isfinite(v).all()
    sizes = [sp.voxels for sp in specs]
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), int(np.argmax([sp.n_steps for sp in specs])), 7,
                    31})
    for k in picks:
        u64, v64 = plain(sw, specs[k], "float64")
        u, v = b.result(k)
        assert rel(u, u64) < 1e-5 and rel(v, v64) < 1e-5, (k, specs[k].shape, rel(u, u64), rel(v, v64))
    for k in picks[:2]:
        u32, v32 = plain(sw, specs[k], "float32")
        u, v = b.result(k)
        assert rel(u, u32) < 1e-6 and rel(v, v32) < 1e-6, (k, rel(u, u32), rel(v, v32))


def test_serving_full_size_requests(sw):
    """Sixteen requests of the serving workload (grids U{64..256}^3, nT U{300..1500}) arriving at 100 req/s:
    all served, overlapping on the device, three of them within 1e-5 of the float64 drop-in."""
    from shearwave_amd.serve import Server, request_specs, with_arrivals

    specs = request_specs(16, seed=4)
    srv = Server(tick_passes=10)
    reqs, summary = srv.run(with_arrivals(specs, 100.0, seed=9))
    assert summary["requests"] == 16 and all(r.u is not None and r.v is not None for r in reqs)
    assert summary["active_solves_max"] > 1  # requests shared the device
    assert 0 < summary["p50_latency_s"] <= summary["p99_latency_s"]
    by_id = {r.rid: r for r in reqs}
    sizes = [int(np.prod(sp.shape)) for sp in specs]
    for k in sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 5}):
        u64, v64 = plain(sw, specs[k], "float64")
        u, v = by_id[k].u, by_id[k].v
        assert rel(u, u64) < 1e-5 and rel(v, v64) < 1e-5, (k, specs[k].shape, rel(u, u64), rel(v, v64))


def test_gradient_128_nt800_full_size(sw):
    """The differentiation workload (128^3, nT = 800, a 128 x 128 sensor plane, float64): the adjoint gradient
    of the misfit w.r.t. mu against central differences at four of its largest entries (1e-5), and the float32
    gradient within 1e-3 of the float64 one."""
    from shearwave_amd import workloads as W
    from shearwave_amd.adjoint import Problem, SensorTraces, plane_sensors

    n, n_steps, h = 128, 800, 0.5e-3
    rng = np.random.default_rng(5)
    shape = (n, n, n)
    rho = np.full(shape, 1000.0)
    mu_true = W.inclusion_map(shape, h, rng, 3, 3000.0, mu_range=(6e3, 15e3))
    dt, _ = W.stable_dt(h, 1000.0, float(mu_true.max()), 1.0)
    bz = W.gaussian_push(shape, h, 1.5e-3)
    env = (np.arange(n_steps) * dt < 100e-6).astype(np.float64)
    dev = torch.device("cuda")
    mu_t = torch.full(shape, 3000.0, dtype=torch.float64, device=dev)
    eta_t = torch.ones(shape, dtype=torch.float64, device=dev)
    grads = {}
    for dtype in ("float64", "float32"):
        prob = Problem(shape, (h,) * 3, dt, rho, (None, None, bz), env, plane_sensors(shape, n // 4), dtype=dtype,
                       tol=1e-10)
        target, _ = prob.forward(torch.tensor(mu_true, device=dev), eta_t, keep_checkpoints=False)
        target = target.to(torch.float64)
        mu_r = mu_t.clone().requires_grad_(True)
        tr = SensorTraces.apply(mu_r, eta_t.clone().requires_grad_(True), prob)
        (0.5 * ((tr.to(torch.float64) - target) ** 2).sum()).backward()
        grads[dtype] = (prob, target, mu_r.grad)
    prob, target, gmu = grads["float64"]
    assert math.isfinite(float(gmu.abs().max())) and float(gmu.abs().max()) > 0

    def J(mu):
        tr, _ = prob.forward(mu, eta_t, keep_checkpoints=False)
        return float(0.5 * ((tr.to(torch.float64) - target) ** 2).sum())

    for idx in torch.argsort(gmu.abs().flatten(), descending=True)[:4].cpu().numpy():
        e = torch.zeros(gmu.numel(), dtype=torch.float64, device=dev)
        eps = 1e-4 * float(mu_t.flatten()[idx])
        e[idx] = eps
        e = e.reshape(shape)
        fd = (J(mu_t + e) - J(mu_t - e)) / (2 * eps)
        ad = float(gmu.flatten()[idx])
        assert abs(fd - ad) < 1e-5 * abs(ad), (int(idx), fd, ad)
    g32 = grads["float32"][2]
    assert float(torch.linalg.norm(g32 - gmu) / torch.linalg.norm(gmu)) < 1e-3
