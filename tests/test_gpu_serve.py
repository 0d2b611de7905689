"""Serving: continuous batching of solve requests that arrive while others run.

Each result is the same alone and in the mix (bit for bit), and within 1e-5
relative L2 of the float64 oracle; requests really overlap on the device.
"""

import copy

import numpy as np
import pytest

from oracle import solver as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def trace():
    from shearwave_amd.serve import poisson_trace

    return poisson_trace(rate=2000.0, n_requests=6, seed=11, sizes=(24, 48), steps=(40, 90))


@pytest.fixture(scope="module")
def served(sw, trace):
    from shearwave_amd.serve import Server

    srv = Server(tick_passes=7)
    reqs, summary = srv.run(copy.deepcopy(trace))
    return reqs, summary


def test_all_served_and_overlapping(served):
    reqs, summary = served
    assert summary["requests"] == len(reqs) and all(np.isfinite(r.latency) for r in reqs)
    assert summary["p99_latency_s"] >= summary["p50_latency_s"] > 0
    # at least two requests were on the GPU at the same time
    spans = sorted((r.start_pass, r.start_pass + r.n_passes) for r in reqs)
    assert any(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1))
    assert len({r.start_pass for r in reqs}) > 1  # and they joined at different ticks


def test_same_alone_and_in_the_mix(sw, trace, served):
    from shearwave_amd.serve import Server

    reqs, _ = served
    for k in (0, 3, 5):
        alone = copy.deepcopy(trace[k])
        alone.arrival = 0.0
        (r,), _ = Server(tick_passes=5).run([alone])
        np.testing.assert_array_equal(r.u, reqs[k].u)
        np.testing.assert_array_equal(r.v, reqs[k].v)


def test_served_results_match_oracle(served):
    reqs, _ = served
    for r in reqs[:2]:
        sp = r.spec
        opts = {"env_t": sp.env, "eta": sp.eta.astype(np.float64), "poisson_tol": 1e-12, "poisson_max_iters": 5000}
        u, v = O.shear_fdtd(None, None, sp.force[2].astype(np.float64), sp.rho.astype(np.float64),
                            sp.mu.astype(np.float64), *sp.h, sp.dt, sp.n_steps, opts=opts, dtype=np.float64)
        assert rel(r.u, u) < 1e-5 and rel(r.v, v) < 1e-5


def test_tight_admission_serves_everything_identically(sw, trace, served):
    """With at most two active solves and a memory budget of about one solve (with its CG vectors),
    requests queue instead of being dropped, start in arrival order, and give the same results."""
    from shearwave_amd.batch import solve_bytes
    from shearwave_amd.serve import Server

    budget = max(solve_bytes(r.spec, with_cg=True) for r in trace) + 1
    srv = Server(tick_passes=7, max_active=2, budget_bytes=budget)
    reqs, summary = srv.run(copy.deepcopy(trace))
    assert summary["requests"] == len(trace)
    starts = [r.start_pass for r in sorted(reqs, key=lambda r: r.arrival)]
    assert starts == sorted(starts)
    ref = {r.arrival: r for r in served[0]}
    for r in reqs:
        np.testing.assert_array_equal(r.u, ref[r.arrival].u)
        np.testing.assert_array_equal(r.v, ref[r.arrival].v)
