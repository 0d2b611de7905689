"""The oracle's float32 mode against vectors the reference itself wrote.

``tests/golden/ref_fp32.npz`` holds the reference's outputs for the seeded
inputs of ``tests/golden/cases.py`` (written by ``tests/golden/make_golden.py``).
Same NumPy version and the same order of operations give bit-identical
float32 results, including the CG stopping iteration and NumPy 2 promotion
with a float64 ``dT``.
"""

import os

import numpy as np
import pytest

from oracle import solver as O
from tests.golden.cases import CASES, build_case

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_fp32.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", CASES)
def test_oracle_fp32_matches_reference(name, gold):
    c = build_case(name)
    if c["kind"] == "project":
        out = O.project_vector_field(c["fx"].copy(), c["fy"].copy(), c["fz"].copy(), *c["h"], c["tol"], c["max_iters"])
        for k, a in zip(("fx", "fy", "fz"), out):
            assert a.dtype == np.float32
            np.testing.assert_array_equal(a, gold[f"{name}/{k}"])
        return
    stride = c.get("stride", 0)
    snaps = []

    def cb(it, u):
        if stride and (it + 1) % stride == 0:
            snaps.append(u.copy())

    u, v = O.shear_fdtd(c["bx"], c["by"], c["bz"], c["rho"], c["mu"], *c["h"], c["dT"], c["n_steps"],
                        opts=c["opts"], callback=cb, dtype=np.float32)
    assert u.dtype == np.float32 and v.dtype == gold[f"{name}/v"].dtype
    np.testing.assert_array_equal(u, gold[f"{name}/u"])
    np.testing.assert_array_equal(v, gold[f"{name}/v"])
    if stride:
        np.testing.assert_array_equal(np.stack(snaps), gold[f"{name}/snaps"])


def test_golden_cases_are_not_trivial(gold):
    for name in CASES:
        key = f"{name}/fz" if name == "project" else f"{name}/u"
        assert np.abs(gold[key]).max() > 0, name
