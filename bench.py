"""Benchmark: the single drop-in shear_fdtd solve on one MI355X.

One bench step is one complete solve of the 256^3 workload with every input
already resident in HBM: the Helmholtz-Hodge projection of the push (CG to
1e-8, float64 reductions), 2000 fused time steps (float32 storage, increment
form), and the four strain invariants every 20 steps into a device-resident
ring. ``value`` is whole-solve throughput in Gvoxel-steps/s (256^3 * 2000 /
seconds per solve, projection and strain included).

The JSON line also carries the roofline of the dominant kernel (in float32
the two-step pass, which advances two force-free viscous steps per launch;
timed with HIP events on the stream it runs on; algorithmic bytes per launch
= 44 B per voxel-step, the reference's (u, u_prev) traffic of one step, x 2
steps x voxels) and a CPU baseline (the NumPy oracle's float32 mode, the
reference's own arithmetic, on a bounded sample of the same loop, one core,
with the float64 oracle beside it). Further records: the projection by the
preconditioned and the plain CG, the float64 form of the solve, the solve
through the drop-in API from host arrays, the 64^3 CPU-runnable case, the
48-solve batch, serving at six arrival rates, the gradient workload, the
radiation force and acoustic strain, the static solve, and non-separable
(per-step projected) forcing.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--nt 2000] [--no-cpu] [--no-fp64]
                    [--no-batch] [--no-serve] [--no-adjoint] [--no-radforce] [--no-static] [--no-force4d]
                    [--only radforce|static|batch|serve|adjoint|force4d] [--config1]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Gvoxel-steps/s & %HBM roofline; batched solves/s + p50/p99 at 1e-5 rel err"
HBM_PEAK_GBS = 8000.0


_T0 = time.perf_counter()


def progress(msg):
    """One line on stderr as a section of the run ends (the JSON record goes to stdout at the end)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nt", type=int, default=2000)
    ap.add_argument("--dtype", default="float32", choices=["float32", "float64"])
    ap.add_argument("--no-fp64", action="store_true", help="skip the float64 solve")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="float32 oracle steps timed for cpu_baseline, after 2 warm-up steps")
    ap.add_argument("--kernel-steps", type=int, default=200)
    ap.add_argument("--kernel", default="auto")
    ap.add_argument("--tile-y", type=int, default=0)
    ap.add_argument("--x-chunk", type=int, default=0)
    ap.add_argument("--no-batch", action="store_true", help="skip the batching comparison")
    ap.add_argument("--batch-solves", type=int, default=48)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving run")
    ap.add_argument("--serve-rates", default="5,10,20,50,100,200")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiation run")
    ap.add_argument("--no-radforce", action="store_true", help="skip the radiation-force run")
    ap.add_argument("--no-static", action="store_true", help="skip the static elasticity solve run")
    ap.add_argument("--no-force4d", action="store_true", help="skip the non-separable forcing run")
    ap.add_argument("--config1", action="store_true", help="only run the 64^3 CPU-vs-GPU comparison and exit")
    ap.add_argument("--only", choices=["radforce", "static", "batch", "serve", "adjoint", "force4d"],
                    help="only run that section (radiation force + acoustic strain, the static solve, the batching "
                         "comparison or the differentiation run), print its record and exit")
    return ap.parse_args()


def traffic_from_profile():
    """HBM bytes per step-kernel launch from the committed PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", "fdtd_step_256_fp32_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


class DropIn:
    """Resident inputs + state of one drop-in solve (the bench step)."""

    def __init__(self, w, dtype, kernel, tile_y, x_chunk):
        import torch

        from shearwave_amd import engine as E

        self.w = w
        self.E = E
        self.solve = E.Solve(w["shape"], (w["h"],) * 3, w["dt"], w["rho"], w["mu"], w["eta"], dtype=dtype,
                             form="increment", kernel=kernel, tile_y=tile_y, x_chunk=x_chunk)
        # as the drop-in API does: with no tile shape given, the autotuner's (measured here, before timing)
        self.tuned = None
        if kernel == "auto" and not tile_y and not x_chunk and self.solve.two_step_fits():
            self.tuned = self.solve.autotune()
        self.bz = E.to_device(w["bz"], dtype, w["shape"])
        self.n_inv = w["n_steps"] // w["strain_stride"]
        self.inv = torch.empty((self.n_inv, 4) + tuple(w["shape"]), dtype=E.torch_dtype(dtype), device=E.device())
        self.env = w["env"]
        self.last_cg = None

    def solve_tuner_entry(self):
        from shearwave_amd.autotune import default_tuner

        t = default_tuner()
        s = self.solve
        return t.entry(t.pass_key(s.dtype, s.e is not None, s.shape))

    def run(self):
        s = self.solve
        s.step_count = 0
        s.set_initial(None, None)
        info, _ = s.set_force(None, None, self.bz, self.w["tol"], self.w["max_iters"])
        self.last_cg = info
        s.run(self.w["n_steps"], self.env, inv_stride=self.w["strain_stride"], inv_out=self.inv)


def pretune_dropin(specs, dtype="float32"):
    """The drop-in autotuner's launch shape for every grid class of `specs` that shear_fdtd would tune,
    measured before a timed run (as the batch tuner is), so that no timed plain solve pays a first
    measurement. Materials do not change the timing; the tuner keys on storage type, viscosity and shape."""
    from shearwave_amd import api
    from shearwave_amd import engine as E
    from shearwave_amd.autotune import default_tuner

    tuner = default_tuner()
    for k, sp in enumerate(specs):
        if sp.n_steps * sp.voxels >= api.AUTOTUNE_MIN_VOXEL_STEPS:
            inviscid = sp.eta is None or (np.isscalar(sp.eta) and float(sp.eta) == 0.0)
            s = E.Solve(sp.shape, sp.h, sp.dt, 1000.0, 3000.0, None if inviscid else 1.0, dtype=dtype)
            if s.two_step_fits():
                s.autotune(tuner)
            del s
        if (k + 1) % 25 == 0:
            progress(f"drop-in launch shapes: {k + 1} of {len(specs)} solves")


def time_step_kernel(solve, n_launch, steps_per_launch):
    """Average duration of one launch of the force-free step kernel (HIP events on the launch stream).

    `steps_per_launch` is 2 for the two-step pass (n_launch launches advance
    2 n_launch steps), 1 for the single-step kernels.
    """
    import torch

    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    # warm with a run as long as the timed one: the first run after a short warm-up
    # measured up to 15% slow (float64, 50 steps at 256^3)
    solve.run(max(10, n_launch * steps_per_launch))
    start.record()
    solve.run(n_launch * steps_per_launch)
    stop.record()
    stop.synchronize()
    return start.elapsed_time(stop) / 1e3 / n_launch


def _oracle_steps(shape, h, m, e, dt, n_steps, seed=0):
    """Seconds for `n_steps` force-free viscous steps of a float64 NumPy loop (after one warm-up step).

    The per-process work of the CPU throughput estimate below (materials as scalars).
    """
    from oracle import solver as O

    rng = np.random.default_rng(seed)
    mask = O.interior_mask(shape, np.float64)
    u_prev = 1e-9 * rng.standard_normal(tuple(shape) + (3,)) * mask[..., None]
    u = 1e-9 * rng.standard_normal(tuple(shape) + (3,)) * mask[..., None]

    def one():
        nonlocal u_prev, u
        vel = (u - u_prev) / dt * mask[..., None]
        nxt = np.empty_like(u)
        for c in range(3):
            acc = m * O.laplacian_center(u[..., c], h, h, h) + e * O.laplacian_center(vel[..., c], h, h, h)
            nxt[..., c] = 2.0 * u[..., c] - u_prev[..., c] + dt * dt * acc
        nxt *= mask[..., None]
        u_prev, u = u, nxt

    one()  # warm-up
    t0 = time.perf_counter()
    for _ in range(n_steps):
        one()
    return time.perf_counter() - t0


def oracle_loop_seconds(w, dtype, n_steps, warm):
    """Seconds per step of the oracle's own time loop (oracle.solver.shear_fdtd) on the workload's grid and fields.

    ``dtype`` float32 is the reference's arithmetic: the oracle's float32 mode
    is bit-identical to solver.py (tests/golden/ref_fp32.npz). Force-free
    steps: bz = 0 with env_t = 0 (the projection exits at once), a random
    initial state so that every step does its full work. Timed from the
    callback after step ``warm - 1`` to the one after step ``warm + n_steps -
    1``: exactly ``n_steps`` whole loop iterations.
    """
    from oracle import solver as O

    shape = tuple(w["shape"])
    rng = np.random.default_rng(0)
    u0 = (1e-9 * rng.standard_normal(shape + (3,))).astype(dtype)
    v0 = (1e-6 * rng.standard_normal(shape + (3,))).astype(dtype)
    total = warm + n_steps
    opts = {"env_t": np.zeros(total), "eta": w["eta"], "initial_u": u0, "initial_v": v0}
    stamps = []
    O.shear_fdtd(None, None, np.zeros(shape, dtype=dtype), w["rho"], w["mu"], w["h"], w["h"], w["h"], w["dt"], total,
                 opts=opts, callback=lambda it, u: stamps.append(time.perf_counter()), dtype=dtype)
    return (stamps[-1] - stamps[warm - 1]) / n_steps


def cpu_projection_sample(w, iters, full_iters):
    """The float64 oracle projection of the same force with the CG capped at `iters` iterations (one core).

    Returns seconds for that call, its CG iterations, and the projection time
    extrapolated to `full_iters` iterations (divergence and gradient-subtract
    once, each CG iteration at the sampled per-iteration cost).
    """
    from oracle import solver as O

    z = np.zeros(w["shape"])
    bz = np.asarray(w["bz"], dtype=np.float64)
    h = w["h"]
    t0 = time.perf_counter()
    O.project_vector_field(z, z, bz, h, h, h, 0.0, 1, dtype=np.float64)
    t1 = time.perf_counter() - t0
    info = {}
    t0 = time.perf_counter()
    O.project_vector_field(z, z, bz, h, h, h, 0.0, iters, dtype=np.float64, info=info)
    tn = time.perf_counter() - t0
    per_iter = (tn - t1) / max(iters - 1, 1)
    return {"seconds": round(tn, 3), "cg_iters": info.get("iters", iters), "per_iter_s": round(per_iter, 4),
            "extrapolated_s": round(t1 + per_iter * (full_iters - 1), 1), "full_iters": full_iters}


def cpu_model():
    """The host CPU's model name (what lscpu prints), or the platform string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(w, n_steps=20, warm=2, n_steps64=3, proj_iters=3, full_iters=0):
    """The reference's CPU path on the same 256^3 workload, one core.

    `value` is the oracle's float32 mode (the reference's own arithmetic) on a
    slice of `n_steps` force-free steps after `warm` warm-up steps; the
    float64 oracle is timed beside it on `n_steps64` steps; the projection
    (float64 CG in both, as poisson_cg) is sampled at `proj_iters` CG
    iterations and extrapolated to `full_iters`. NumPy ufuncs are
    single-threaded here.
    """
    vox = int(np.prod(w["shape"]))
    t32 = oracle_loop_seconds(w, np.float32, n_steps, warm)
    t64 = oracle_loop_seconds(w, np.float64, n_steps64, 1)
    proj = cpu_projection_sample(w, proj_iters, full_iters) if proj_iters > 0 else None
    n_total = w["n_steps"]
    rec = {
        "value": round(1.0 / t32 * vox / 1e9, 5),
        "unit": "Gvoxel-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{n_steps} force-free viscous steps of the {w['shape'][0]}^3 loop after {warm} warm-up steps, "
                  f"oracle.solver.shear_fdtd in float32 mode (bit-identical to solver.py's float32 arithmetic): "
                  f"{t32:.3f} s per step; loop only, the projection is sampled apart",
        "seconds_per_step": round(t32, 4),
        "float64": {"gvox_steps_per_s": round(1.0 / t64 * vox / 1e9, 5), "seconds_per_step": round(t64, 4),
                    "sample": f"{n_steps64} steps after 1 warm-up step, float64 oracle"},
        "projection_sample": proj,
        "numpy": np.__version__,
        "host": platform.machine(),
        "cpu_model": cpu_model(),
        "os_cpu_count": os.cpu_count(),
    }
    if proj is not None:
        solve_s = proj["extrapolated_s"] + n_total * t32
        rec["solve_s_extrapolated"] = round(solve_s, 1)
        rec["solve_note"] = (f"projection (extrapolated to {full_iters} CG iterations) + {n_total} float32 steps at "
                             f"the sampled rate; in-loop strain not included")
    return rec


def _cpu_worker(args):
    shape, n_steps, seed = args
    t = _oracle_steps(shape, 5e-4, 2.5, 1.0e-3, 2e-5, n_steps, seed)
    return int(np.prod(shape)) * n_steps / t


def cpu_parallel_throughput(shape=(128, 128, 160), n_steps=2, max_procs=16):
    """Independent float64 oracle loops, one process per core (at most max_procs): aggregate voxel-steps/s.

    The throughput a CPU node gets on the batching and serving workloads by
    running one solve per core. Loop only (the projection is not timed),
    which flatters the CPU. Run before the GPU is initialised: the workers are
    spawned processes.
    """
    import multiprocessing as mp

    n = max(1, min(os.cpu_count() or 1, max_procs))
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(n) as pool:
        rates = pool.map(_cpu_worker, [(shape, n_steps, s) for s in range(n)])
    return {
        "processes": n,
        "gvox_steps_per_s": round(sum(rates) / 1e9, 4),
        "per_process_gvox_steps_per_s": round(min(rates) / 1e9, 5),
        "sample": f"{n} processes, each {n_steps} force-free steps of a {shape[0]}x{shape[1]}x{shape[2]} float64 "
                  f"oracle loop after a warm-up step; loop only, projection excluded; {time.perf_counter() - t0:.1f} s",
    }


def bench_batch(n_solves=48, check=8):
    """Config 3: 48 uneven solves, batched vs the plain per-solve drop-in (host arrays in and out)."""
    import torch

    import shearwave_amd as sw
    from shearwave_amd import workloads as W
    from shearwave_amd.autotune import default_tuner
    from shearwave_amd.batch import Batch

    specs = W.batch_48(n_solves=n_solves)
    vox_steps = sum(sp.voxels * sp.n_steps for sp in specs)
    tuner = default_tuner()
    tile_y, cap = tuner.batch_shape()  # tuned (or read from the cache) before timing
    pretune_dropin(specs)
    progress("batch: launch shapes tuned")

    def plain(sp, dtype="float32"):
        opts = {"env_t": sp.env, "eta": sp.eta, "vel_damp": sp.vel_damp, "poisson_tol": sp.tol,
                "poisson_max_iters": sp.max_iters, "sw_dtype": dtype}
        return sw.shear_fdtd(None, None, sp.force[2], sp.rho, sp.mu, *sp.h, sp.dt, sp.n_steps, opts)

    plain(specs[0])  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_out = [plain(sp) for sp in specs]
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    progress("batch: plain drop-in solves done")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b = Batch(specs, tuner=tuner)
    t_setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    info = b.project()
    t_proj = time.perf_counter() - t1
    t1 = time.perf_counter()
    b.run(collect=True)  # results exported to host memory while the remaining solves run
    t_run = time.perf_counter() - t1
    t_loop = b.last_loop_s
    run_timing = {k: round(v, 3) for k, v in b.last_run_timing.items()}
    batch_out = [b.result(k) for k in range(len(specs))]
    b_perm, graph_steps = [s.perm for s in b.solves], b.graph_steps
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0

    errs = []
    for k in range(min(check, len(specs))):
        u64, v64 = plain(specs[k], "float64")
        ub, vb = batch_out[k]
        errs.append(max(float(np.linalg.norm(ub - u64) / np.linalg.norm(u64)),
                        float(np.linalg.norm(vb - v64) / np.linalg.norm(v64))))
    agree = max(float(np.linalg.norm(pb[0] - pp[0]) / np.linalg.norm(pp[0])) for pb, pp in zip(batch_out, plain_out))
    del b
    torch.cuda.empty_cache()
    return {
        "workload": f"{len(specs)} solves, grids U{{64..192}}^2 x U{{64..256}}, nT U{{500..2000}}, float32, seed 3",
        "voxel_steps": vox_steps,
        "batched_solves_per_s": round(len(specs) / t_batch, 3),
        "plain_dropin_solves_per_s": round(len(specs) / t_plain, 3),
        "speedup": round(t_plain / t_batch, 3),
        "batched_s": round(t_batch, 3),
        "plain_s": round(t_plain, 3),
        "batched_phases_s": {"upload_setup": round(t_setup, 3), "projection": round(t_proj, 3),
                             "loop_and_results": round(t_run, 3), "loop_on_device": round(t_loop, 3)},
        "batched_loop_gvox_steps_per_s": round(vox_steps / t_loop / 1e9, 2),
        "solves_stored_permuted": sum(1 for s in b_perm if s != (0, 1, 2)),
        "graph_steps": graph_steps,
        "tile_y": tile_y,
        "x_chunk_cap": cap,
        "tuner": tuner.entry(tuner.key()),
        "run_timing_s": run_timing,
        "cg_iters_max": max(i["iters"] for i in info),
        "max_rel_err_vs_fp64": max(errs) if errs else None,
        "checked_solves": len(errs),
        "max_rel_diff_batched_vs_plain": agree,
    }


def bench_serve(rates=(5.0, 10.0, 20.0, 50.0, 100.0, 200.0), n_requests=200, tick_passes=10):
    """Config 4: Poisson arrivals, continuous batching; solves/s and p50/p99 latency per rate.

    The same seeded pool of requests (grids U{64..256}^3, nT U{300..1500}) is
    served at every rate, each rate with its own seeded Poisson arrivals.
    """
    import torch

    import shearwave_amd as sw
    from shearwave_amd.autotune import default_tuner
    from shearwave_amd.serve import Server, request_specs, with_arrivals

    tuner = default_tuner()
    t0 = time.perf_counter()
    pool = request_specs(n_requests, seed=4)
    t_gen = time.perf_counter() - t0
    tuner.batch_shape()  # tuned (or read from the cache) before timing
    pretune_dropin(pool)
    progress("serve: launch shapes tuned")
    out = {"workload": f"{n_requests} requests per rate from one seeded pool (seed 4): grids U{{64..256}}^3 "
                       f"(multiples of 8), nT U{{300..1500}}, float32, CG tol 1e-8; Poisson arrivals seeded per rate; "
                       f"ticks of {tick_passes} passes (one or two steps each), 3 ticks in flight, hipGraph replay",
           "pool_generation_s": round(t_gen, 1), "rates": []}
    for k, rate in enumerate(rates):
        reqs = with_arrivals(pool, rate, seed=100 + k)
        torch.cuda.synchronize()
        srv = Server(tick_passes=tick_passes, tuner=tuner)
        done, summ = srv.run(reqs)
        summ = {kk: (round(v, 4) if isinstance(v, float) else v) for kk, v in summ.items()}
        summ["rate_req_per_s"] = rate
        out["rates"].append(summ)
        progress(f"serve: {rate:g} req/s done ({summ['solves_per_s']} solves/s)")
        for r in done:
            r.u = r.v = None
        del srv, done
        torch.cuda.empty_cache()
    # the plain drop-in on the same requests, one after another (arrivals ignored: its best case)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sp in pool:
        opts = {"env_t": sp.env, "eta": sp.eta, "poisson_tol": sp.tol, "poisson_max_iters": sp.max_iters}
        sw.shear_fdtd(None, None, sp.force[2], sp.rho, sp.mu, *sp.h, sp.dt, sp.n_steps, opts)
    torch.cuda.synchronize()
    plain_rate = len(pool) / (time.perf_counter() - t0)
    out["plain_dropin_solves_per_s"] = round(plain_rate, 3)
    best = max(r["solves_per_s"] for r in out["rates"])
    out["best_served_over_plain"] = round(best / plain_rate, 3)
    return out


def bench_radforce(shape=(160, 160, 160), nt=256, on=(64, 192), reps=5, cpu_shape=(48, 48, 48)):
    """Rows f3 and f4: radiation force and acoustic strain over a resident float32 pressure record (nx, ny, nz, nt).

    Times the two kernels with HIP events on their stream (the auto-gate's
    energy pass over the whole record, the force pass over the gated
    samples) and the whole drop-in call on device tensors. Algorithmic bytes:
    energy = the record once; force = the gated samples once, c, rho, alpha
    read and b written once. The CPU baseline is the float32 oracle on a
    `cpu_shape` block of the same record, one core.
    """
    import torch

    import shearwave_amd as sw
    from shearwave_amd import api
    from shearwave_amd import engine as E
    from shearwave_amd._abi import check, lib

    n_vox = int(np.prod(shape))
    dev = E.device()
    g = torch.Generator(device=dev).manual_seed(7)
    p = torch.randn((n_vox, nt), generator=g, device=dev, dtype=torch.float32)
    t = torch.arange(nt, device=dev)
    p *= torch.where((t >= on[0]) & (t < on[1]), 1.0e5, 2.0e3).to(torch.float32)
    c = 1450.0 + 150.0 * torch.rand(shape, generator=g, device=dev, dtype=torch.float32)
    rho = 950.0 + 150.0 * torch.rand(shape, generator=g, device=dev, dtype=torch.float32)
    alpha = 0.3 + 0.9 * torch.rand(shape, generator=g, device=dev, dtype=torch.float32)
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    env = torch.empty(nt, dtype=torch.float64, device=dev)
    need = int(lib().sw_pressure_energy_workspace(n_vox, nt))
    work = torch.empty(need // 8 + 1, dtype=torch.float64, device=dev)
    st = E.stream_ptr()

    def energy():
        check(lib().sw_pressure_energy(0, p.data_ptr(), n_vox, nt, env.data_ptr(), work.data_ptr(), need, st))

    gate = api._auto_gate(p, n_vox, nt, "float32")
    gidx, g0, n_gate = api._gate_args(gate)

    def force():
        check(lib().sw_radiation_force(0, p.data_ptr(), n_vox, nt, None if gidx is None else gidx.data_ptr(), g0,
                                       n_gate, c.data_ptr(), rho.data_ptr(), alpha.data_ptr(), 1.0e6, out.data_ptr(),
                                       st))

    def timed(fn):
        fn()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / 1e3 / reps

    t_e = timed(energy)
    t_f = timed(force)
    # acoustic strain kernels on the same record and gate (acoustic_strain.py)
    n_gate = len(gate)
    g_ptr = None if gidx is None else gidx.data_ptr()
    stats = torch.empty((3, n_vox), dtype=torch.float32, device=dev)
    gstats = torch.empty((5, n_vox), dtype=torch.float64, device=dev)
    vol = torch.empty((n_vox, n_gate), dtype=torch.float32, device=dev)

    def strain_peak():
        check(lib().sw_acoustic_strain_peak(0, p.data_ptr(), n_vox, nt, g_ptr, g0, n_gate, c.data_ptr(),
                                            rho.data_ptr(), stats[0].data_ptr(), stats[1].data_ptr(),
                                            stats[2].data_ptr(), st))

    def gradient_peak():
        check(lib().sw_acoustic_strain_gradient_peak(0, p.data_ptr(), *shape, nt, g_ptr, g0, n_gate, c.data_ptr(),
                                                     rho.data_ptr(), 0.3e-3, 0.3e-3, 0.3e-3,
                                                     *(gstats[j].data_ptr() for j in range(5)), st))

    def volumetric():
        check(lib().sw_acoustic_strain_volumetric(0, p.data_ptr(), n_vox, nt, g_ptr, g0, n_gate, c.data_ptr(),
                                                  rho.data_ptr(), vol.data_ptr(), st))

    t_sp = timed(strain_peak)
    t_gp = timed(gradient_peak)
    t_vol = timed(volumetric)
    del vol
    grad = torch.empty((3, n_vox, n_gate), dtype=torch.float32, device=dev)

    def gradient():
        check(lib().sw_acoustic_strain_gradient(0, p.data_ptr(), *shape, nt, g_ptr, g0, n_gate, c.data_ptr(),
                                                rho.data_ptr(), 0.3e-3, 0.3e-3, 0.3e-3, grad[0].data_ptr(),
                                                grad[1].data_ptr(), grad[2].data_ptr(), st))

    t_grad = timed(gradient)
    del grad, stats, gstats
    torch.cuda.empty_cache()
    gated = n_vox * n_gate * 4  # the gated samples, read once
    p4 = p.view(shape + (nt,))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sw.compute_radiation_force(p4, c, rho, alpha, 1.0e6)
    torch.cuda.synchronize()
    t_call = (time.perf_counter() - t0) / reps
    e_bytes = n_vox * nt * 4
    f_bytes = n_vox * (len(gate) * 4 + 3 * 4 + 4)
    # CPU baseline: the float32 oracle on a block of the same record
    from oracle import radiation_force as R

    cn = int(np.prod(cpu_shape))
    pc = p[:cn].cpu().numpy().reshape(cpu_shape + (nt,))
    cc, rc_, ac = (m.flatten()[:cn].cpu().numpy().reshape(cpu_shape) for m in (c, rho, alpha))
    t0 = time.perf_counter()
    R.compute_radiation_force(pc, cc, rc_, ac, 1.0e6)
    t_cpu = time.perf_counter() - t0
    del p, p4, work
    torch.cuda.empty_cache()
    return {
        "workload": f"pressure record {shape[0]}x{shape[1]}x{shape[2]} x nt={nt} float32 ({e_bytes / 1e9:.2f} GB), "
                    f"burst in samples {on[0]}..{on[1] - 1}, auto-gate, f0 = 1 MHz, resident in HBM",
        "gate_samples": int(len(gate)),
        "energy_pass": {"us": round(t_e * 1e6, 1), "GBs": round(e_bytes / t_e / 1e9, 1),
                        "frac": round(e_bytes / t_e / 1e9 / HBM_PEAK_GBS, 4)},
        "force_pass": {"us": round(t_f * 1e6, 1), "GBs": round(f_bytes / t_f / 1e9, 1),
                       "frac": round(f_bytes / t_f / 1e9 / HBM_PEAK_GBS, 4)},
        "acoustic_strain": {
            "peak_rms_mean": {"us": round(t_sp * 1e6, 1), "GBs": round((gated + n_vox * 20) / t_sp / 1e9, 1),
                              "frac": round((gated + n_vox * 20) / t_sp / 1e9 / HBM_PEAK_GBS, 4)},
            "gradient_statistics": {"us": round(t_gp * 1e6, 1),
                                    "GBs": round((gated + n_vox * (8 + 40)) / t_gp / 1e9, 1),
                                    "frac": round((gated + n_vox * 48) / t_gp / 1e9 / HBM_PEAK_GBS, 4)},
            "volumetric": {"us": round(t_vol * 1e6, 1), "GBs": round((2 * gated + n_vox * 8) / t_vol / 1e9, 1),
                           "frac": round((2 * gated + n_vox * 8) / t_vol / 1e9 / HBM_PEAK_GBS, 4)},
            "gradient": {"us": round(t_grad * 1e6, 1), "GBs": round((4 * gated + n_vox * 8) / t_grad / 1e9, 1),
                         "frac": round((4 * gated + n_vox * 8) / t_grad / 1e9 / HBM_PEAK_GBS, 4)},
            "bytes": "gated samples read once; maps (c, rho) read and outputs written once",
        },
        "dropin_call_ms": round(t_call * 1e3, 3),
        "gvoxel_samples_per_s": round(n_vox * nt / t_call / 1e9, 2),
        "cpu_baseline": {"gvoxel_samples_per_s": round(cn * nt / t_cpu / 1e9, 4), "cores": 1, "kind": "port",
                         "sample": f"float32 oracle on {cpu_shape[0]}x{cpu_shape[1]}x{cpu_shape[2]} x nt={nt}, "
                                   f"auto-gate; {t_cpu:.2f} s"},
    }


def bench_static(n=256, k1=10, k2=60, cpu_iters=40):
    """Row f4: the static elasticity solve (solver.py:569-625) on a resident n^3 float32 force.

    Per-iteration time from two calls that differ only in their iteration cap
    (tol = 0 never stops early): (T(k2) - T(k1)) / (k2 - k1), timed with HIP
    events on the call's stream, so launches and the host's stop-flag polls
    are included and the set-up is not. Algorithmic bytes per voxel and
    iteration: the dir kernel reads r, p_old, mu and writes p (16 B), the step
    kernel reads p, u, r, mu and writes u, r (24 B): 40 B. Also the
    reference's quasi-static validation solve (72x48x72, tol 1e-6,
    validate_quasistatic.py:24-33, 83-85) end to end on device tensors, and
    the float32 oracle's rate over `cpu_iters` iterations of that case on one
    core.
    """
    import torch

    import shearwave_amd as sw
    from shearwave_amd import engine as E

    dev = E.device()
    g = torch.Generator(device=dev).manual_seed(11)
    shape = (n, n, n)
    w = torch.sin(torch.pi * (torch.arange(n, device=dev, dtype=torch.float32) + 0.5) / n)
    bz = 0.02 * torch.randn(shape, generator=g, device=dev, dtype=torch.float32) * w[:, None, None] * w[None, :, None]
    bz = (bz * w[None, None, :]).contiguous()
    mu = 2000.0 + 4000.0 * torch.rand(shape, generator=g, device=dev, dtype=torch.float32)
    d = 2.0e-4

    def call(k):
        return sw.elastic_solve_static(bz, d, d, d, mu, tol=0.0, max_iters=k)

    def timed(k, reps=2):
        call(k)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            call(k)
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / 1e3 / reps

    t1, t2 = timed(k1), timed(k2)
    per_it = (t2 - t1) / (k2 - k1)
    vox = n**3
    alg = 40 * vox
    # the reference's validation solve
    vshape, vd = (72, 48, 72), 2.0e-4
    xg, yg, zg = np.meshgrid(*(np.arange(m) - m // 2 for m in vshape), indexing="ij")
    G = np.exp(-0.5 * ((xg / 8.0) ** 2 + (yg / 8.0) ** 2 + (zg / 6.0) ** 2)).astype(np.float32)
    G /= np.max(G)
    taper = np.ones(vshape, np.float32)
    for ax, m in enumerate(vshape):
        taper = taper * np.hanning(m).astype(np.float32).reshape([-1 if q == ax else 1 for q in range(3)])
    # the force of the validation run: its tapered Gaussian, projected to the shear part
    _, _, vbz = sw.project_body_force_to_shear((0.02 * G * taper).astype(np.float32), vd, vd, vd, tol=5e-5,
                                               max_iters=400)
    vbz_t = torch.from_numpy(vbz).to(dev)
    mu_v = 1000.0 * 2.0**2
    info = {}
    sw.elastic_solve_static(vbz_t, vd, vd, vd, mu_v, tol=1e-6, max_iters=1500, info=info)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sw.elastic_solve_static(vbz_t, vd, vd, vd, mu_v, tol=1e-6, max_iters=1500)
    torch.cuda.synchronize()
    t_val = time.perf_counter() - t0
    from oracle import static_solve as S

    t0 = time.perf_counter()
    S.elastic_solve_static(vbz, vd, vd, vd, mu_v, tol=0.0, max_iters=cpu_iters)
    t_cpu = (time.perf_counter() - t0) / cpu_iters
    vvox = int(np.prod(vshape))
    del bz, mu
    torch.cuda.empty_cache()
    return {
        "workload": f"{n}^3 float32 force and mu field resident in HBM, h = {d}, CG iterations timed",
        "us_per_iteration": round(per_it * 1e6, 1),
        "gvoxel_iters_per_s": round(vox / per_it / 1e9, 2),
        "alg_bytes_per_voxel_iter": 40,
        "GBs": round(alg / per_it / 1e9, 1),
        "frac": round(alg / per_it / 1e9 / HBM_PEAK_GBS, 4),
        "validation_solve": {"shape": list(vshape), "iters": info.get("iters"), "reason": info.get("reason"),
                             "ms": round(t_val * 1e3, 2)},
        "cpu_baseline": {"gvoxel_iters_per_s": round(vvox / t_cpu / 1e9, 4), "cores": 1, "kind": "port",
                         "sample": f"float32 oracle, {cpu_iters} iterations of the 72x48x72 validation solve, "
                                   f"{t_cpu * 1e3:.1f} ms per iteration"},
    }


def bench_adjoint(n=128, n_steps=800, n_fd=32):
    """Config 5: gradient of a sensor-plane misfit w.r.t. mu and eta (128^3, nT=800)."""
    import torch

    from shearwave_amd import workloads as W
    from shearwave_amd.adjoint import Problem, SensorTraces, plane_sensors

    rng = np.random.default_rng(5)
    shape = (n, n, n)
    h = 0.5e-3
    rho = np.full(shape, 1000.0)
    mu_bg = np.full(shape, 3000.0)
    mu_true = W.inclusion_map(shape, h, rng, 3, 3000.0, mu_range=(6e3, 15e3))
    eta = np.full(shape, 1.0)
    dt, dt_cfl = W.stable_dt(h, 1000.0, float(mu_true.max()), 1.0)
    bz = W.gaussian_push(shape, h, 1.5e-3)
    env = (np.arange(n_steps) * dt < 100e-6).astype(np.float64)
    sensors = plane_sensors(shape, n // 4)
    dev = torch.device("cuda")
    res = {"workload": f"{n}^3, h=0.5 mm, nT={n_steps}, loss = 1/2 sum (u_z - u_z_target)^2 over the y-z plane "
                       f"x={n // 4} ({n}x{n} sensors) and every step; target mu: 3 seeded inclusions; "
                       "gradient at the background mu, eta = 1 Pa s"}
    out = {}
    for dtype in ("float32", "float64"):
        prob = Problem(shape, (h,) * 3, dt, rho, (None, None, bz), env, sensors, dtype=dtype, tol=1e-10)
        mu_t = torch.tensor(mu_bg, device=dev)
        eta_t = torch.tensor(eta, device=dev)
        target, _ = prob.forward(torch.tensor(mu_true, device=dev), eta_t, keep_checkpoints=False)
        target = target.to(torch.float64)
        # two rounds: the first also grows the allocator's cache by the checkpoints (about 1.5 GB
        # at 128^3 float32), which the sections before this one may have left fragmented
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            mu_r = mu_t.clone().requires_grad_(True)
            eta_r = eta_t.clone().requires_grad_(True)
            tr = SensorTraces.apply(mu_r, eta_r, prob)
            loss = 0.5 * ((tr.to(torch.float64) - target) ** 2).sum()
            torch.cuda.synchronize()
            t_fwd = time.perf_counter() - t0
            loss.backward()
            torch.cuda.synchronize()
            t_all = time.perf_counter() - t0
        loss = loss.detach()
        out[dtype] = (prob, mu_t, eta_t, target, mu_r.grad, eta_r.grad)
        res[dtype] = {"forward_s": round(t_fwd, 3), "forward_plus_adjoint_s": round(t_all, 3),
                      "checkpoints": math.ceil(n_steps / prob.K), "checkpoint_every": prob.K, "loss": float(loss)}
    prob, mu_t, eta_t, target, gmu, geta = out["float64"]

    def J(mu, et):
        tr, _ = prob.forward(mu, et, keep_checkpoints=False)
        return float(0.5 * ((tr.to(torch.float64) - target) ** 2).sum())

    cand = torch.argsort(gmu.abs().flatten(), descending=True)[: 20 * n_fd].cpu().numpy()
    picks = rng.choice(cand, n_fd, replace=False)
    errs = []
    for idx in picks:
        e = torch.zeros(shape, dtype=torch.float64, device=dev).flatten()
        eps = 1e-4 * float(mu_t.flatten()[idx])
        e[idx] = eps
        e = e.reshape(shape)
        fd = (J(mu_t + e, eta_t) - J(mu_t - e, eta_t)) / (2 * eps)
        ad = float(gmu.flatten()[idx])
        errs.append(abs(fd - ad) / abs(ad))
    g32 = out["float32"][4]
    res["fd_check"] = {"voxels": n_fd, "field": "mu", "dtype": "float64", "max_rel_err": max(errs),
                       "median_rel_err": float(np.median(errs)), "eps": "1e-4 relative, central"}
    res["fp32_vs_fp64_grad_rel_l2"] = float(torch.linalg.norm(g32 - gmu) / torch.linalg.norm(gmu))
    res["jax_gradient_check"] = "not run: jax is not installed and there is no network"
    del out
    torch.cuda.empty_cache()
    return res


def dropin_host_to_host(w, dtype="float32"):
    """The same 256^3 solve through the drop-in API with host arrays in and out (PCIe included).

    shearwave_amd.shear_fdtd uploads rho, mu, eta (float64) and bz, projects,
    runs the loop and returns u and v as NumPy arrays (pinned downloads). No
    in-loop strain (its 100 evaluations would be 27 GB of host output). One
    warm-up call, then one timed call.
    """
    import torch

    import shearwave_amd as sw

    h = w["h"]
    opts = {"env_t": w["env"], "eta": w["eta"], "poisson_tol": w["tol"], "poisson_max_iters": w["max_iters"],
            "sw_dtype": dtype}
    args = (None, None, w["bz"], w["rho"], w["mu"], h, h, h, w["dt"], w["n_steps"])
    sw.shear_fdtd(*args, opts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u, v = sw.shear_fdtd(*args, opts)
    t = time.perf_counter() - t0
    vox = int(np.prod(w["shape"]))
    item = 4 if dtype == "float32" else 8
    return {"seconds": round(t, 4), "gvox_steps_per_s": round(vox * w["n_steps"] / t / 1e9, 3),
            "upload_bytes": 3 * 8 * vox + item * vox, "download_bytes": 2 * 3 * item * vox,
            "note": "host NumPy arrays in, host arrays out (u, v); projection and the 2000-step loop on the GPU; "
                    "no in-loop strain"}


def bench_force4d(cpu_steps=1, window=16, proj_steps=(0, 33, 66)):
    """Row f1: non-separable forcing (solver.py:118-122, 199), a 128^3 x nT=100 float32 bz projected every step
    at the reference's defaults (tol 1e-5, 200 iterations), through the drop-in API from host arrays.

    Three calls, each timed whole: the default (the force uploaded once and
    sliced on the device, the forces of each window of `window` steps
    projected in one batched CG before the window's steps run), the same
    with each step's force projected on its own (``sw_force_window`` 1), and
    each step's slice cut and uploaded from the host and projected on its own
    (``sw_resident_force`` False as well). The projections alone, with HIP
    events: a window's batched CG, and single steps' CG; GB/s at 34 B per
    voxel-iteration, the float32-vector CG's traffic. One reprojection
    (``reproject_every``) of a state of the same grid at the same defaults.
    CPU baseline: the float32 oracle (the reference's arithmetic) over
    `cpu_steps` steps of the same slices, and one projection of the
    reprojection's, one core.
    """
    import torch

    import shearwave_amd as sw
    from oracle import solver as O  # the checker
    from shearwave_amd import engine as E
    from shearwave_amd import workloads as W

    w = W.force4d_128()
    n_steps, h, bz = w["n_steps"], w["h"], w["bz"]
    vox = int(np.prod(w["shape"]))
    opts = {"eta": w["eta"], "poisson_tol": w["tol"], "poisson_max_iters": w["max_iters"], "sw_force_window": window}
    mats = (w["rho"], w["mu"], h, h, h, w["dt"])
    sw.shear_fdtd(None, None, np.ascontiguousarray(bz[..., :2]), *mats, 2, opts)  # warm-up
    runs = {}
    for name, extra in (("batched_windows", {}), ("per_step", {"sw_force_window": 1}),
                        ("per_step_host_sliced", {"sw_force_window": 1, "sw_resident_force": False})):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sw.shear_fdtd(None, None, bz, *mats, n_steps, {**opts, **extra})
        t = time.perf_counter() - t0
        runs[name] = {"seconds": round(t, 4), "ms_per_step": round(t / n_steps * 1e3, 3)}
        progress(f"4-D forcing: {name} call done ({t:.2f} s)")

    # the projections alone: one window's batched CG, and single steps' CG, on a solve of the workload
    s = E.Solve(w["shape"], (h, h, h), w["dt"], w["rho"], w["mu"], w["eta"])
    bz_d = torch.from_numpy(bz).to(E.device())

    def events():
        return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    forces = [(None, None, bz_d[..., t]) for t in range(window)]
    s.project_forces(forces, w["tol"], w["max_iters"])  # warm
    a, b = events()
    a.record()
    _, infos = s.project_forces(forces, w["tol"], w["max_iters"])
    b.record()
    b.synchronize()
    ms_w = a.elapsed_time(b)
    it_w = sum(max(1, i["iters"]) for i in infos)
    batched = {"steps": window, "ms": round(ms_w, 3), "ms_per_step": round(ms_w / window, 3),
               "iters": sorted({i["iters"] for i in infos}),
               "us_per_step_iter": round(ms_w * 1e3 / it_w, 2),
               "GBs": round(34 * vox * it_w / (ms_w / 1e3) / 1e9, 1)}
    singles = []
    for k in proj_steps:
        a, b = events()
        a.record()
        info, _ = s.set_force(None, None, bz_d[..., k], w["tol"], w["max_iters"])
        b.record()
        b.synchronize()
        ms = a.elapsed_time(b)
        it = max(1, info["iters"])
        singles.append({"step": k, "iters": info["iters"], "ms": round(ms, 3), "us_per_iter": round(ms * 1e3 / it, 2),
                        "GBs": round(34 * vox * it / (ms / 1e3) / 1e9, 1)})
    # reproject_every (solver.py:236-239): one reprojection of a state of this grid, u and u_prev each projected
    # at the same defaults (float64 fields and CG vectors in the increment form)
    rng = np.random.default_rng(6)
    u0 = (1e-6 * rng.standard_normal(w["shape"] + (3,))).astype(np.float32)
    v0 = (1e-3 * rng.standard_normal(w["shape"] + (3,))).astype(np.float32)
    s.set_initial(u0, v0)
    s.reproject(w["tol"], w["max_iters"])  # warm
    s.set_initial(u0, v0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rinfo = s.reproject(w["tol"], w["max_iters"])
    torch.cuda.synchronize()
    t_rep = time.perf_counter() - t0
    reproj = {"ms": round(t_rep * 1e3, 2), "cg_iters": [i["iters"] for i in rinfo],
              "note": "u and u_prev = u - d projected in float64 (fields and CG vectors), d re-formed in float64"}
    del s, bz_d
    torch.cuda.empty_cache()
    cpu = None
    if cpu_steps > 0:  # the oracle's projection, the unit of a reprojection (two per reprojection)
        k = 20
        t0 = time.perf_counter()
        O.project_vector_field(u0[..., 0], u0[..., 1], u0[..., 2], h, h, h, w["tol"], k, dtype=np.float32)
        t_k = time.perf_counter() - t0
        reproj["cpu_ms"] = round(t_k / k * sum(reproj["cg_iters"]) * 1e3, 1)
        reproj["cpu_note"] = (f"float32 oracle: one projection of u0 capped at {k} CG iterations ({t_k:.2f} s), "
                              "extrapolated to the GPU run's iterations of both projections, one core")
    if cpu_steps > 0:  # the float32 oracle on the first cpu_steps slices
        info_cpu = []
        t0 = time.perf_counter()
        O.shear_fdtd(None, None, np.ascontiguousarray(bz[..., :cpu_steps]), *mats, cpu_steps, opts=opts,
                     dtype=np.float32, proj_info=info_cpu)
        t_cpu = (time.perf_counter() - t0) / cpu_steps
        cpu = {"ms_per_step": round(t_cpu * 1e3, 1), "cores": 1, "kind": "port",
               "cg_iters": [int(i.get("iters", -1)) for i in info_cpu],
               "sample": f"float32 oracle, the first {cpu_steps} steps (each projected), one core",
               "gpu_speedup": round(t_cpu / (runs["batched_windows"]["seconds"] / n_steps), 1)}
    return {
        "workload": "128^3, h=0.5 mm, rho 1000, mu 3 kPa with 3 inclusions (seed 6), eta 1, nT=100, bz a "
                    "(128,128,128,100) float32 Gaussian push (sigma 2 mm) moving 16 voxels along z, ramped "
                    "300 -> 0, projected every step (tol 1e-5, 200 iterations); host arrays in and out",
        "force_bytes": int(bz.nbytes),
        "calls": runs,
        "speedup_over_per_step_host_sliced": round(runs["per_step_host_sliced"]["seconds"] /
                                                   runs["batched_windows"]["seconds"], 3),
        "projection": {"batched_window": batched, "single_steps": singles,
                       "roofline": {"bound": "hbm", "kernel": "batched float32-vector CG iteration (dir + step)",
                                    "alg_bytes_per_voxel_iter": 34, "achieved": batched["GBs"],
                                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": round(batched["GBs"] / HBM_PEAK_GBS, 4)}},
        "reprojection": reproj,
        "cpu_baseline": cpu,
    }


def config1_cpu_vs_gpu():
    """Config 1 (the reference's CPU case, 64^3, nT=500) on CPU and GPU: the CPU-baseline leg for it.

    Runs the float64 NumPy oracle in full and the GPU drop-in in float64 and
    float32 (host arrays in and out), with relative L2 errors against the oracle.
    """
    import torch

    import shearwave_amd as sw
    from oracle import solver as O  # the checker
    from shearwave_amd import workloads as W

    w = W.plain_64()
    opts = {"env_t": w["env"], "eta": w["eta"], "poisson_tol": w["tol"], "poisson_max_iters": w["max_iters"]}
    args = (None, None, w["bz"], w["rho"], w["mu"], w["h"], w["h"], w["h"], w["dt"], w["n_steps"])
    vox_steps = int(np.prod(w["shape"])) * w["n_steps"]
    t0 = time.perf_counter()
    u_ref, v_ref, s_ref = O.shear_fdtd_snapshots(*args, opts, w["snapshot_stride"], dtype=np.float64)
    t_cpu = time.perf_counter() - t0
    rec = {"workload": "plain shear_fdtd 64^3, nT=500, eta=1, CG tol 1e-10, snapshot stride 10",
           "dt": w["dt"], "dt_cfl_unstable": w["dt_cfl"],
           "cpu_oracle_fp64": {"seconds": round(t_cpu, 3), "gvox_steps_per_s": vox_steps / t_cpu / 1e9,
                               "cores": 1, "numpy": np.__version__, "cpu_model": cpu_model()}}

    def rel(a, b):
        return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))

    for dtype in ("float64", "float32"):
        o = {**opts, "sw_dtype": dtype}
        sw.shear_fdtd_jax(*args, o, snapshot_stride=w["snapshot_stride"])  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u, v, snaps = sw.shear_fdtd_jax(*args, o, snapshot_stride=w["snapshot_stride"])
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        rec[f"gpu_{dtype}"] = {"seconds_host_to_host": round(t, 4), "gvox_steps_per_s": vox_steps / t / 1e9,
                               "rel_l2_u": rel(u, u_ref), "rel_l2_v": rel(v, v_ref), "rel_l2_snapshots": rel(snaps, s_ref)}
    return rec



def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the multi-GPU layout is "
              "a design, not implemented)", file=sys.stderr)
        sys.exit(2)
    # before anything touches the GPU: the CPU workers are spawned processes
    cpu_par = None
    if not args.no_cpu and not (args.no_batch and args.no_serve) and not args.only and not args.config1:
        cpu_par = cpu_parallel_throughput()
        progress("CPU process-parallel baseline done")
    import torch

    from shearwave_amd import workloads as W

    if args.config1:
        print(json.dumps(config1_cpu_vs_gpu()), flush=True)
        return
    if args.only:
        sections = {"radforce": bench_radforce, "static": bench_static, "adjoint": bench_adjoint,
                    "force4d": bench_force4d,
                    "batch": lambda: bench_batch(args.batch_solves),
                    "serve": lambda: bench_serve(tuple(float(r) for r in args.serve_rates.split(",")),
                                                 args.serve_requests)}
        print(json.dumps(sections[args.only]()), flush=True)
        return
    w = W.dropin_256(n=args.n, n_steps=args.nt)
    item = 4 if args.dtype == "float32" else 8
    vox = int(np.prod(w["shape"]))
    bench = DropIn(w, args.dtype, args.kernel, args.tile_y, args.x_chunk)
    progress(f"drop-in solve set up (two-step tile {bench.tuned})")
    for _ in range(args.warmup):
        bench.run()
    torch.cuda.synchronize()
    progress("warm-up solves done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bench.run()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    per_solve = t / args.steps
    value = vox * w["n_steps"] / per_solve / 1e9
    progress(f"timed solves done: {value:.1f} Gvoxel-steps/s")
    # the final state of the timed solve, kept for the float32-against-float64 check below
    final32 = (torch.stack(bench.solve.u_components()).double(), torch.stack(bench.solve.b_components()).double())

    # the projection by the plain CG (the reference's algorithm, poisson_cg): its iteration count is what the
    # CPU reference runs; its time shows what the preconditioner saves
    s = bench.solve
    os.environ["SW_CG_PRECOND"] = "0"
    try:
        torch.cuda.synchronize()
        tp0 = time.perf_counter()
        plain_info, _ = s.set_force(None, None, bench.bz, w["tol"], w["max_iters"])
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - tp0
    finally:
        del os.environ["SW_CG_PRECOND"]
    # projection alone and the loop alone (one extra solve, phase split), with the solver the drop-in uses
    torch.cuda.synchronize()
    tp0 = time.perf_counter()
    s.set_initial(None, None)
    s.set_force(None, None, bench.bz, w["tol"], w["max_iters"])
    torch.cuda.synchronize()
    t_proj = time.perf_counter() - tp0

    # the dominant kernel: the two-step pass in float32 (auto / march2), else the single step
    two = args.dtype == "float32" and args.kernel in ("auto", "march2")
    spl = 2 if two else 1
    kt = time_step_kernel(s, args.kernel_steps // spl, spl)
    per_voxel_step = W.alg_step_bytes(item, viscous=True, forced=False)  # 44 B (fp32) / 88 B (fp64)
    alg_bytes = per_voxel_step * spl * vox  # algorithmic bytes per launch
    achieved = alg_bytes / kt / 1e9
    traffic = traffic_from_profile() if (args.n == 256 and args.dtype == "float32") else None

    rec = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gvoxel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(per_solve * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.dtype == "float32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"single drop-in shear_fdtd solve, {args.n}^3 grid, h=0.25 mm, 8 seeded inclusions, "
                        f"per-voxel rho and eta, Gaussian push, nT={w['n_steps']}, CG tol 1e-8, strain invariants "
                        f"every 20 steps, {args.dtype} storage (increment form), inputs resident in HBM",
            "grid": list(w["shape"]),
            "n_steps": w["n_steps"],
            "dt": w["dt"],
            "dt_cfl": w["dt_cfl"],
            "seed": 2,
            "cg_vectors": args.dtype,
            "two_step_tile": None if bench.tuned is None else {
                "tile_y": bench.tuned[0], "x_chunk": bench.tuned[1], "chosen_by": "autotuner",
                "candidates_ps_per_voxel_step": (bench.solve_tuner_entry() or {}).get("ps_per_voxel_step")},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": ("step2_march_kernel (two force-free viscous steps per launch)" if two else
                       "step_march_kernel (force-free, viscous)" if args.kernel != "direct" else "step_direct_kernel"),
            "alg_bytes_per_voxel_step": per_voxel_step,
            "voxel_steps_per_launch": spl * vox,
            "avg_launch_us": round(kt * 1e6, 2),
            "traffic_GBs": None if traffic is None else round(traffic / kt / 1e9, 1),
        },
        "phases": {
            "projection_s": round(t_proj, 4),
            "cg_iters": bench.last_cg["iters"],
            "cg_relres": bench.last_cg["relres"],
            "cg_true_relres": bench.last_cg["true_relres"],
            "cg_solver": "spectral-preconditioned CG" if bench.last_cg["iters"] < plain_info["iters"] else "plain CG",
            "plain_cg_projection_s": round(t_plain, 4),
            "plain_cg_iters": plain_info["iters"],
            "plain_cg_true_relres": plain_info["true_relres"],
            "loop_gvox_steps_per_s": round(spl * vox / kt / 1e9, 2),
        },
    }
    if not args.no_fp64:
        b64 = DropIn(w, "float64", args.kernel, 0, 0)
        b64.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b64.run()
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        # the whole 2000-step solve at full size: float32 storage (float32 CG vectors) against the
        # float64 form, which the parity tests hold within 1e-10 of the float64 oracle
        final64 = (torch.stack(b64.solve.u_components()), torch.stack(b64.solve.b_components()))
        if args.dtype == "float32":
            def rl2(a, b):
                return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
            rec["fp32_vs_fp64"] = {"rel_l2_u": rl2(final32[0], final64[0]),
                                   "rel_l2_v": rl2(final32[1], final64[1]),
                                   "note": "final u and d = v dT of the timed solves; gate 1e-5"}
        del final64
        kt64 = time_step_kernel(b64.solve, args.kernel_steps // 2, 2)  # two-step passes, as in float32
        ab64 = W.alg_step_bytes(8, viscous=True, forced=False) * 2 * vox
        rec["fp64"] = {"value": round(vox * w["n_steps"] / t64 / 1e9, 3), "ms_per_solve": round(t64 * 1e3, 2),
                       "step_kernel_GBs": round(ab64 / kt64 / 1e9, 1), "frac": round(ab64 / kt64 / 1e9 / HBM_PEAK_GBS, 4)}
        del b64
        progress("float64 solve done")
    del final32
    if not args.no_batch:
        del bench
        torch.cuda.empty_cache()
        rec["batch"] = bench_batch(args.batch_solves)
        progress("batching done")
        if cpu_par:
            per_solve = rec["batch"]["voxel_steps"] / args.batch_solves
            rec["batch"]["cpu_parallel"] = {**cpu_par, "solves_per_s_est": round(cpu_par["gvox_steps_per_s"] * 1e9 /
                                                                                   per_solve, 5)}
    if not args.no_serve:
        rec["serve"] = bench_serve(tuple(float(r) for r in args.serve_rates.split(",")), args.serve_requests)
        progress("serving done")
        if cpu_par:
            per_req = rec["serve"]["rates"][0]["voxel_steps"] / args.serve_requests
            rec["serve"]["cpu_parallel"] = {**cpu_par, "solves_per_s_est": round(cpu_par["gvox_steps_per_s"] * 1e9 /
                                                                                   per_req, 5)}
    if not args.no_adjoint:
        rec["adjoint"] = bench_adjoint()
        progress("differentiation done")
    if not args.no_radforce:
        rec["radiation_force"] = bench_radforce()
        progress("radiation force and acoustic strain done")
    if not args.no_cpu:
        rec["dropin_host_to_host"] = dropin_host_to_host(w, args.dtype)
        progress("host-to-host drop-in done")
        rec["config1"] = config1_cpu_vs_gpu()
        progress("64^3 CPU-runnable case done")
    if not args.no_static:
        rec["static_solve"] = bench_static()
        progress("static solve done")
    if not args.no_force4d:
        rec["force4d"] = bench_force4d(cpu_steps=0 if args.no_cpu else 1)
        progress("non-separable forcing done")
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(w, args.cpu_steps, warm=2, n_steps64=3, proj_iters=3,
                                           full_iters=int(rec["phases"]["plain_cg_iters"]))
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
