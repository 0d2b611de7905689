"""Batched solves (one persistent launch per step for many uneven solves).

Gates: every solve of a batch within 1e-5 relative L2 of the float64 oracle
(float32 storage); a solve's result is bit-identical alone and in the mix;
hipGraph replay equals direct launches bit for bit.
"""

import math

import numpy as np
import pytest

from oracle import solver as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make_spec(shape, seed, n_steps, viscous=True, damp=0.0):
    from shearwave_amd.batch import SolveSpec

    rng = np.random.default_rng(seed)
    h = (0.3e-3, 0.3e-3, 0.3e-3)
    rho = 1000.0 * (1 + rng.uniform(-0.01, 0.01, shape))
    mu = np.full(shape, 2500.0)
    ax = [np.arange(n) * h[0] for n in shape]
    for _ in range(3):
        r = rng.uniform(0.15, 0.3) * min(shape) * h[0]
        c = [rng.uniform(r, n * h[0] - r) for n in shape]
        m = (ax[0][:, None, None] - c[0]) ** 2 + (ax[1][None, :, None] - c[1]) ** 2 + (ax[2][None, None, :] - c[2]) ** 2
        mu[m < r * r] = rng.uniform(4e3, 15e3)
    eta = rng.uniform(0.5, 3.0, shape) if viscous else None
    cmax = math.sqrt(mu.max() / rho.min())
    dt = 0.3 * h[0] / (cmax * math.sqrt(3))
    if viscous:
        dt = min(dt, 0.5 * rho.min() * h[0] ** 2 / (6 * eta.max()))
    centre = [rng.uniform(0.3, 0.7) * (n - 1) for n in shape]
    g = np.meshgrid(*(np.arange(n) - cc for n, cc in zip(shape, centre)), indexing="ij")
    bz = 300.0 * np.exp(-(g[0] ** 2 + g[1] ** 2 + g[2] ** 2) / (2 * 3.0**2))
    env = np.zeros(n_steps)
    env[: max(1, n_steps // 5)] = 1.0
    return SolveSpec(shape=shape, h=h, dt=dt, rho=rho, mu=mu, eta=eta, force=(None, None, bz), env=env,
                     n_steps=n_steps, vel_damp=damp, tol=1e-10, max_iters=5000)


def oracle_of(sp):
    opts = {"env_t": sp.env, "vel_damp": sp.vel_damp, "poisson_tol": 1e-12, "poisson_max_iters": 5000}
    if sp.eta is not None:
        opts["eta"] = sp.eta
    return O.shear_fdtd(None, None, sp.force[2], sp.rho, sp.mu, *sp.h, sp.dt, sp.n_steps, opts=opts, dtype=np.float64)


SHAPES = [((24, 20, 70), 60, True, 0.0), ((30, 18, 40), 45, False, 0.01), ((16, 24, 66), 80, True, 0.02),
          ((40, 36, 44), 30, True, 0.0), ((20, 9, 33), 55, True, 0.0)]


@pytest.fixture(scope="module")
def specs():
    return [make_spec(sh, 100 + k, nt, visc, damp) for k, (sh, nt, visc, damp) in enumerate(SHAPES)]


@pytest.fixture(scope="module")
def mixed(sw, specs):
    from shearwave_amd.batch import Batch

    b = Batch(specs, graph_steps=0)  # direct launches
    info = b.project()
    assert all(i["reason"] == "converged" for i in info), info
    b.run()
    # the per-solve axis order permutes some of these grids in HBM
    assert any(s.perm != (0, 1, 2) for s in b.solves)
    return [b.result(k) for k in range(len(specs))]


def test_batch_matches_oracle(specs, mixed):
    for sp, (u, v) in zip(specs, mixed):
        u_ref, v_ref = oracle_of(sp)
        assert rel(u, u_ref) < 1e-5, (sp.shape, rel(u, u_ref))
        assert rel(v, v_ref) < 1e-5, (sp.shape, rel(v, v_ref))


def test_alone_equals_in_mix(sw, specs, mixed):
    from shearwave_amd.batch import Batch

    for k in (0, 3):
        b = Batch([specs[k]])
        b.project()
        b.run()
        u, v = b.result(0)
        np.testing.assert_array_equal(u, mixed[k][0])
        np.testing.assert_array_equal(v, mixed[k][1])


@pytest.mark.parametrize("graph_steps,collect", [(7, False), (16, True)])
def test_graph_replay_equals_direct(sw, specs, mixed, graph_steps, collect):
    """Graph replays equal direct launches bit for bit; with collect the results exported and copied to the
    host in one go after the loop (permuted back to the given axis order there) are the same."""
    from shearwave_amd.batch import Batch

    b = Batch(specs, graph_steps=graph_steps)
    b.project()
    b.run(collect=collect)
    for k in range(len(specs)):
        u, v = b.result(k)
        assert u.shape == specs[k].shape + (3,) and u.dtype == np.float32
        np.testing.assert_array_equal(u, mixed[k][0])
        np.testing.assert_array_equal(v, mixed[k][1])
    if collect:
        for k in range(len(specs)):
            ud, vd = b.result(k, to_host=False)  # exported again, on the main stream
            np.testing.assert_array_equal(ud.cpu().numpy(), b.result(k)[0])
            np.testing.assert_array_equal(vd.cpu().numpy(), b.result(k)[1])


def test_axis_order_changes_only_rounding(sw, specs, mixed):
    """Stored in the given axis order, every solve agrees with its permuted storage to float32 rounding (the
    stencil adds the axes' terms in another order), and with the oracle."""
    from shearwave_amd.batch import Batch

    b = Batch(specs, axes="given")
    assert all(s.perm == (0, 1, 2) for s in b.solves)
    b.project()
    b.run()
    for k, sp in enumerate(specs):
        u, v = b.result(k)
        assert rel(u, mixed[k][0]) < 1e-6 and rel(v, mixed[k][1]) < 1e-5, (sp.shape, rel(u, mixed[k][0]))


def test_batch_agrees_with_single_solve_dropin(sw, specs, mixed):
    k = 2
    sp = specs[k]
    opts = {"env_t": sp.env, "vel_damp": sp.vel_damp, "poisson_tol": sp.tol, "poisson_max_iters": sp.max_iters,
            "eta": sp.eta}
    u, v = sw.shear_fdtd(None, None, sp.force[2], sp.rho, sp.mu, *sp.h, sp.dt, sp.n_steps, opts)
    assert rel(u, mixed[k][0]) < 1e-6 and rel(v, mixed[k][1]) < 1e-6


def test_two_step_passes_against_single_steps(sw, specs, mixed):
    """Pass plans that pair steps (forced ones included) against one step per pass.

    The increments of a pair equal two single steps' bit for bit; the pair
    stores u_{n+2} = u_n + (d_{n+1} + d_{n+2}) with one rounding where single
    steps round twice, so the results agree to float32 rounding (the exact
    rule is checked on the same tile code in test_gpu_two_step).
    """
    from shearwave_amd.batch import Batch

    b = Batch(specs, two_step=False)
    assert all(s.n_passes == s.spec.n_steps for s in b.solves)
    b.project()
    b.run()
    for k in range(len(specs)):
        u, v = b.result(k)
        assert rel(u, mixed[k][0]) < 1e-6 and rel(v, mixed[k][1]) < 1e-5, (rel(u, mixed[k][0]), rel(v, mixed[k][1]))


def test_zero_force_solve_in_a_batch(sw, specs, mixed):
    """A solve with no force projects by the early exit (max|div| < 1e-12) and stays zero; its
    neighbour in the batch is unchanged, bit for bit."""
    from shearwave_amd.batch import Batch, SolveSpec

    sp = specs[0]
    zero = SolveSpec(shape=sp.shape, h=sp.h, dt=sp.dt, rho=sp.rho, mu=sp.mu, eta=sp.eta,
                     force=(None, None, np.zeros(sp.shape)), env=sp.env, n_steps=sp.n_steps, vel_damp=sp.vel_damp,
                     tol=sp.tol, max_iters=sp.max_iters)
    b = Batch([zero, specs[1]])
    info = b.project()
    assert info[0]["reason"] == "early_exit" and info[0]["iters"] == 0
    b.run()
    u0, v0 = b.result(0)
    assert not np.any(u0) and not np.any(v0)
    u1, v1 = b.result(1)
    np.testing.assert_array_equal(u1, mixed[1][0])
    np.testing.assert_array_equal(v1, mixed[1][1])


def test_batched_cg_stops_after_odd_and_even_iterations(sw):
    """Solves of one batch stop after 5 and 6 iterations: the deferred phi term of the odd one is flushed."""
    from shearwave_amd.batch import Batch

    sps = [make_spec((24, 20, 70), 300, 10), make_spec((26, 19, 66), 301, 10)]
    for sp, k in zip(sps, (5, 6)):
        sp.tol, sp.max_iters = 1e-30, k
    b = Batch(sps)
    info = b.project(method="batched")
    for k, sp in enumerate(sps):
        assert info[k]["reason"] == "max_iters" and info[k]["iters"] == sp.max_iters, info[k]
        z = np.zeros(sp.shape, np.float32)
        want = sw.project_vector_field(z, z, sp.force[2].astype(np.float32), *sp.h, 1e-30, sp.max_iters,
                                       cg_vectors="float32")
        got = b.solves[k].given_vec(b.solves[k].f).cpu().numpy()
        for c in range(3):
            if np.any(want[c]):
                assert rel(got[c], want[c]) < 1e-6, (k, c, rel(got[c], want[c]))


def test_batched_cg_item_length_changes_nothing(sw):
    """Items of 48, 96 or 192 x planes: the kernels sum each dot product per 48-plane
    sub-chunk, so the projected forces and iteration counts agree bit for bit (grids of
    several sub-chunks, the last one ragged, and one shorter than a sub-chunk)."""
    from shearwave_amd.batch import Batch

    sps = [make_spec((150, 20, 70), 310, 4), make_spec((97, 17, 66), 311, 4), make_spec((40, 12, 64), 312, 4)]
    got = []
    for xc in (48, 96, 192):
        b = Batch(sps)
        info = b.project(xchunk=xc, method="batched")
        assert all(i["reason"] == "converged" for i in info), info
        got.append(([i["iters"] for i in info], [[t.cpu().numpy() for t in s.force] for s in b.solves]))
    for iters, forces in got[1:]:
        assert iters == got[0][0]
        for fa, fb in zip(forces, got[0][1]):
            for a, c in zip(fa, fb):
                np.testing.assert_array_equal(a, c)
    assert any(np.any(f) for f in got[0][1][0])
    with pytest.raises(ValueError, match="multiple of 48"):
        Batch(sps[:1]).project(xchunk=100, method="batched")


def test_batched_cg_rejects_bad_descriptors(sw, monkeypatch):
    """The library checks the descriptors before any launch: an x chunk that is not a
    multiple of 48 (and shorter than nx) is an argument error, not a stalled reduction."""
    from shearwave_amd import batch as B

    b = B.Batch([make_spec((150, 20, 70), 313, 4)], axes="given")  # x stays the 150-plane axis
    monkeypatch.setattr(B, "cg_xchunk", lambda shapes, n_cu: 100)
    with pytest.raises(RuntimeError, match="multiple of 48"):
        b.project(method="batched")


def test_projection_methods_agree(sw, specs):
    """Each solve projected in turn (spectral-preconditioned CG) and all of them by one batched plain CG: the
    projected loop forces agree to the CG tolerance (1e-10)."""
    from shearwave_amd.batch import Batch

    got = {}
    for method in ("each", "batched"):
        b = Batch(specs)
        info = b.project(method=method)
        assert all(i["reason"] == "converged" for i in info), (method, info)
        got[method] = ([i["iters"] for i in info], [s.given_vec(s.force).cpu().numpy() for s in b.solves])
    assert max(got["each"][0]) <= 12 < min(got["batched"][0])
    for a, c in zip(got["each"][1], got["batched"][1]):
        assert rel(a, c) < 1e-7


def test_tuner_times_the_mix(sw, tmp_path):
    """The autotuner's measurement on the device: every (tile rows, cap) pair timed on the tuning mix, the
    fastest cached."""
    from shearwave_amd.autotune import Tuner, batch_candidates

    t = Tuner(path=str(tmp_path / "tune.json"))
    rows, cap = t.batch_shape()
    entry = t.entry(t.key())
    assert (rows, cap) in batch_candidates()
    assert set(entry["ps_per_voxel_step"]) == {f"{r}x{c}" for r, c in batch_candidates()}
    times = entry["ps_per_voxel_step"]
    assert all(v > 0 for v in times.values()) and times[f"{rows}x{cap}"] == min(times.values())


@pytest.mark.parametrize("rows", [4, 12])
def test_tile_rows_change_nothing(sw, specs, rows):
    """The batched step's tile rows (4, 8, 12; the autotuner picks one per batch) change the work items and
    the launch, never a voxel's arithmetic: results equal the 8-row tiles' bit for bit (axis order fixed,
    since the lane fill that picks it depends on the rows), graph replays included."""
    from shearwave_amd.batch import Batch

    got = {}
    for by in (8, rows):
        b = Batch(specs, tile_y=by, axes="given", graph_steps=5)
        b.project()
        b.run()
        got[by] = [b.result(k) for k in range(len(specs))]
    for (u8, v8), (u, v) in zip(got[8], got[rows]):
        np.testing.assert_array_equal(u, u8)
        np.testing.assert_array_equal(v, v8)
