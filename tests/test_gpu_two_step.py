"""The two-step march pass (two steps per read of the state) against single steps.

A pass evaluates each step with the same operations as the single-step
kernel (contraction off, fused multiply-adds written out), so both increments
are bit-identical to those of two single steps. The displacement it stores
takes one rounding per pass, u_{n+2} = u_n + (d_{n+1} + d_{n+2}), where two
single steps round twice; the reference below runs single steps and applies
that same sum at the end of each pass, so the comparison stays exact
(float32 and float64; partial tiles in y and z, with and without viscosity,
tile shapes and x chunks, around outputs and forced steps). Parity against the
float64 oracle is covered in test_gpu_parity (float64 within 1e-10, float32
within 1e-5).
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from tests.test_gpu_parity import H, _problem, _run_gpu  # noqa: E402


def _random_state(shape, seed):
    rng = np.random.default_rng(seed)
    u0 = 1e-6 * rng.standard_normal(shape + (3,))
    v0 = 1e-3 * rng.standard_normal(shape + (3,))
    return u0, v0


def _solve(sw, shape, seed, kernel, viscous, tile_y=0, x_chunk=0, dtype="float32"):
    from shearwave_amd import engine as E

    rng = np.random.default_rng(seed + 1)
    rho = 1000.0 * (1 + rng.uniform(-0.01, 0.01, shape))
    mu = rng.uniform(2e3, 2e4, shape)
    eta = rng.uniform(0.5, 3.0, shape) if viscous else None
    h = 0.5e-3
    dt = 0.2 * h / np.sqrt(2e4 / 990.0 * 3)
    s = E.Solve(shape, (h, h, h), dt, rho, mu, eta, dtype=dtype, kernel=kernel, tile_y=tile_y,
                x_chunk=x_chunk, vel_damp=0.003)
    u0, v0 = _random_state(shape, seed)
    s.set_initial(u0, v0)
    return s


def _flat(res):
    out = []
    for x in res:
        out.extend(x.values() if isinstance(x, dict) else [x])
    return [np.asarray(v) for v in out]


def _state(s):
    return [t.clone() for t in s.u_components()] + [t.clone() for t in s.b_components()]


def _plan(n_steps, out_after=lambda step: False, env=None, three=False):
    """The pass lengths sw_fdtd_run chooses: three force-free steps (march3,
    float32) or two steps wherever no output falls inside the pass, else one."""
    plan, it = [], 0
    while it < n_steps:
        free = env is None or all(env[it + j] == 0 for j in range(3) if it + j < n_steps)
        if three and it + 2 < n_steps and not out_after(it) and not out_after(it + 1) and free:
            k = 3
        elif it + 1 < n_steps and not out_after(it):
            k = 2
        else:
            k = 1
        plan.append(k)
        it += k
    return plan


def _run_rule(s, plan, env=None, after_pass=None):
    """Single steps on `s` (kernel "march"); a pass of k >= 2 steps then stores
    u = u_{k-2} + (d_{k-1} + d_k), as the multi-step passes do."""
    it = 0
    for k in plan:
        for j in range(k):
            if j == k - 2:
                u_keep = [t.clone() for t in s.u_components()]
            s.run(1, None if env is None else env[it + j: it + j + 1])
            if j == k - 2:
                d_keep = [t.clone() for t in s.b_components()]
        if k >= 2:
            d = s.b_components()
            for c, t in enumerate(s.u_components()):
                t.copy_(u_keep[c] + (d_keep[c] + d[c]))
        it += k
        if after_pass is not None:
            after_pass(it - 1)


def _assert_equal_states(got, ref):
    for a, b in zip(_state(got), _state(ref)):
        assert torch.equal(a, b), (a - b).abs().max().item()


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("shape", [(40, 37, 70), (24, 9, 130), (33, 20, 64), (9, 8, 32)])
@pytest.mark.parametrize("viscous", [True, False])
@pytest.mark.parametrize("n_steps", [7, 10])
def test_two_step_pass_rule(sw, shape, viscous, n_steps, dtype):
    ref = _solve(sw, shape, 5, "march", viscous, dtype=dtype)
    _run_rule(ref, _plan(n_steps))
    got = _solve(sw, shape, 5, "march2", viscous, dtype=dtype)
    got.run(n_steps)
    torch.cuda.synchronize()
    _assert_equal_states(got, ref)
    assert float(ref.u_components()[0].abs().max()) > 0


@pytest.mark.parametrize("tile_y,x_chunk", [(12, 0), (8, 5), (8, 17), (12, 3)])
def test_two_step_tile_shapes(sw, tile_y, x_chunk):
    shape = (37, 29, 140)
    ref = _solve(sw, shape, 6, "march", True)
    _run_rule(ref, _plan(8))
    got = _solve(sw, shape, 6, "march2", True, tile_y=tile_y, x_chunk=x_chunk)
    got.run(8)
    torch.cuda.synchronize()
    _assert_equal_states(got, ref)


def _forced_solve(sw, p, shape, kernel, dtype):
    from shearwave_amd import engine as E

    s = E.Solve(shape, H, p["dt"], p["rho"], p["mu"], p["opts"]["eta"], dtype=dtype, kernel=kernel,
                vel_damp=p["opts"]["vel_damp"])
    s.set_force(p["bx"], None, p["bz"], p["opts"]["poisson_tol"], p["opts"]["poisson_max_iters"])
    return s


def run_with_outputs_against_rule(sw, kernel, dtype, snap_stride, inv_stride, three=False):
    """A forced run with snapshots and in-loop strain against single steps under the pass rule."""
    from shearwave_amd import _abi
    from shearwave_amd import engine as E

    shape = (18, 16, 70)
    p = _problem(shape, seed=21, n_steps=60)
    env = np.asarray(p["opts"]["env_t"], dtype=np.float64)
    n = p["n_steps"]
    n_vox = int(np.prod(shape))
    tdt = E.torch_dtype(dtype)
    dev = E.device()

    got = _forced_solve(sw, p, shape, kernel, dtype)
    snaps = torch.zeros((n // snap_stride,) + shape + (3,), dtype=tdt, device=dev)
    inv = torch.zeros((n // inv_stride, 4) + shape, dtype=tdt, device=dev)
    got.run(n, env, snap_stride=snap_stride, snap_out=snaps, inv_stride=inv_stride, inv_out=inv)

    ref = _forced_solve(sw, p, shape, "march", dtype)
    snaps_ref = torch.zeros_like(snaps)
    inv_ref = torch.zeros_like(inv)

    def record(step):
        u = ref.u_components()
        if (step + 1) % snap_stride == 0:
            snaps_ref[(step + 1) // snap_stride - 1] = torch.stack(u, dim=-1)
        if (step + 1) % inv_stride == 0:
            o = inv_ref[(step + 1) // inv_stride - 1]
            _abi.check(_abi.lib().sw_strain_invariants_from_u(
                ref.code, u[0].data_ptr(), u[1].data_ptr(), u[2].data_ptr(), o[0].data_ptr(), o[1].data_ptr(),
                o[2].data_ptr(), o[3].data_ptr(), *shape, *H, E.stream_ptr()))

    def out_after(step):
        return (step + 1) % snap_stride == 0 or (step + 1) % inv_stride == 0

    plan = _plan(n, out_after, env, three=three and dtype == "float32")
    assert max(plan) == (3 if three and dtype == "float32" else 2) and 1 in plan
    _run_rule(ref, plan, env, after_pass=record)
    torch.cuda.synchronize()
    _assert_equal_states(got, ref)
    assert torch.equal(snaps, snaps_ref)
    # the in-loop invariants multiply by 1/(2h) where sw_strain_invariants_from_u divides (NumPy's order):
    # equal to a few roundings
    tol = 1e-6 if dtype == "float32" else 1e-14
    for k in range(inv.shape[0]):
        for j in range(4):
            a, b = inv[k, j].double(), inv_ref[k, j].double()
            assert float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) < tol, (k, j)
    assert float(snaps.abs().max()) > 0 and float(inv.abs().max()) > 0


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_two_step_around_force_and_outputs(sw, dtype):
    """Passes over forced and ramped steps; snapshots every 3 and strain every 4 steps break the pairing."""
    run_with_outputs_against_rule(sw, "march2", dtype, 3, 4)
    # the API's automatic kernel choice runs the same passes
    shape = (18, 16, 70)
    p = _problem(shape, seed=21, n_steps=48)  # a multiple of the snapshot stride (solver.py:469-507)
    kw = dict(snapshot_stride=3, return_strain=True)
    o = {"sw_strain_stride": 4, "sw_dtype": dtype}
    got = _run_gpu(sw, p, shape, {**o, "sw_kernel": "march2"}, **kw)
    auto = _run_gpu(sw, p, shape, o, **kw)
    assert len(got) == 4  # u, v, snapshots, strain dict
    for a, b in zip(_flat(got), _flat(auto)):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("dtype,shape,viscous,force", [("float32", (19, 13, 70), True, False),
                                                        ("float32", (24, 16, 64), False, True),
                                                        ("float64", (17, 11, 45), True, True)])
def test_pass_with_intermediate_state(sw, dtype, shape, viscous, force):
    """sw_fdtd_step2_mid: the pass's final state equals a two-step pass of sw_fdtd_run bit for bit, and the
    intermediate state it writes equals one single step (the adjoint's recomputation relies on both)."""
    from shearwave_amd import engine as E

    s = _solve(sw, shape, 7, "auto", viscous, dtype=dtype)
    if force:
        rng = np.random.default_rng(3)
        s.set_force(None, None, rng.standard_normal(shape), 1e-10, 5000)
    env = np.array([0.7, 0.3]) if force else np.zeros(2)
    u0, d0 = torch.stack(s.u_components()).clone(), torch.stack(s.b_components()).clone()
    um, dm, u2, d2 = (torch.full_like(u0, float("nan")) for _ in range(4))
    assert s.step2_into(u0, d0, um, dm, u2, d2, env)
    u1, d1 = torch.empty_like(u0), torch.empty_like(u0)
    s.step_into(u0, d0, u1, d1, env[0])  # one single step (kernel "march" for single steps of the auto kernel)
    assert torch.equal(um, u1) and torch.equal(dm, d1)
    s.run(2, env)  # one two-step pass of the loop from the same state
    assert torch.equal(u2, torch.stack(s.u_components())) and torch.equal(d2, torch.stack(s.b_components()))
    assert float(d2.abs().max()) > 0 and float(dm.abs().max()) > 0
    # grids the pass does not take: nothing is written
    small = _solve(sw, (10, 6, 20), 7, "auto", viscous, dtype=dtype)
    v0 = torch.stack(small.u_components())
    z = torch.zeros_like(v0)
    assert not small.step2_into(v0, v0, z, z, z, z)
    assert float(z.abs().max()) == 0


# ------------------------------------------------------- launch-shape tuning


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_trial_passes_leave_the_state_and_write_one_pass(sw, dtype):
    """sw_fdtd_trial_passes (the autotuner's timing pass) reads the state and writes only the next-step
    buffers: the state is unchanged, and the next-step buffers hold what one two-step pass of the loop
    writes, whatever the number of trial passes."""
    import ctypes as C

    from shearwave_amd import _abi
    from shearwave_amd import engine as E

    shape = (36, 21, 70)
    s = _solve(sw, shape, 31, "march2", True, dtype=dtype)
    before = _state(s)
    _abi.check(_abi.lib().sw_fdtd_trial_passes(C.byref(s.state), 3, E.stream_ptr()))
    for a, b in zip(before, _state(s)):
        torch.testing.assert_close(a, b, rtol=0, atol=0)
    nxt = [t.clone() for t in s._vec(s.state.u_next)] + [t.clone() for t in s._vec(s.state.b_next)]
    s.run(2)
    for a, b in zip(nxt, _state(s)):
        torch.testing.assert_close(a, b, rtol=0, atol=0)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_autotuned_launch_shape_changes_nothing(sw, dtype, tmp_path):
    """The autotuner times every candidate tile shape on the solve's own buffers, caches the fastest, and
    the solve then runs bit for bit as with the default launch shape (and as with every other candidate)."""
    from shearwave_amd.autotune import PASS_ROWS, Tuner, pass_candidates

    shape = (60, 40, 96)
    t = Tuner(path=str(tmp_path / "tune.json"))
    ref = _solve(sw, shape, 32, "auto", True, dtype=dtype)
    ref.run(30)
    want = _state(ref)
    s = _solve(sw, shape, 32, "auto", True, dtype=dtype)
    before = _state(s)
    ty, xc = s.autotune(t)
    assert (ty, xc) in pass_candidates(dtype, shape) and ty in PASS_ROWS[dtype]
    e = t.entry(t.pass_key(dtype, True, shape))
    assert len(e["ps_per_voxel_step"]) == len(pass_candidates(dtype, shape))
    assert min(e["ps_per_voxel_step"].values()) == e["ps_per_voxel_step"][f"{ty}x{xc}"] > 0
    for a, b in zip(before, _state(s)):  # tuning left the state alone
        torch.testing.assert_close(a, b, rtol=0, atol=0)
    s.run(30)
    for a, b in zip(want, _state(s)):
        torch.testing.assert_close(a, b, rtol=0, atol=0)
    for cty, cxc in pass_candidates(dtype, shape)[::3]:  # and a sample of the other candidates
        o = _solve(sw, shape, 32, "march2", True, tile_y=cty, x_chunk=cxc, dtype=dtype)
        o.run(30)
        for a, b in zip(want, _state(o)):
            torch.testing.assert_close(a, b, rtol=0, atol=0)


def test_dropin_autotune_option(sw, tmp_path, monkeypatch):
    """shear_fdtd tunes the launch shape of a long enough loop (opts sw_autotune, default on) and returns
    the same arrays as with tuning off."""
    import shearwave_amd.autotune as A
    from shearwave_amd import api

    monkeypatch.setattr(A, "_DEFAULT", A.Tuner(path=str(tmp_path / "tune.json")))
    monkeypatch.setattr(api, "AUTOTUNE_MIN_VOXEL_STEPS", 1)
    shape = (40, 36, 70)
    p = _problem(shape, 5, n_steps=24)
    u1, v1 = _run_gpu(sw, p, shape, {"sw_autotune": False})
    assert not A._DEFAULT._table
    u2, v2 = _run_gpu(sw, p, shape)
    assert any("|pass|float32|" in k for k in A._DEFAULT._table)
    np.testing.assert_array_equal(u1, u2)
    np.testing.assert_array_equal(v1, v2)
