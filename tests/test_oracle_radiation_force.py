"""Pins of the radiation-force oracle (closed forms; the reference has no tests for this module)."""

import math

import numpy as np
import pytest

from oracle import radiation_force as R

SPP = 8  # samples per period: sum of sin^2 over whole periods is exactly N/2


def burst(shape=(5, 4, 6), nt=40, on=(8, 32), seed=0, dtype=np.float64, phase_shift=None, amp2=None):
    """p = A sin(w t + phi) (and v = B sin(w t + phi + delta)) inside `on`, zero outside."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 2.0, shape)
    phi = rng.uniform(0, 2 * np.pi, shape)
    t = np.arange(nt)
    w = 2 * np.pi / SPP
    active = (t >= on[0]) & (t < on[1])
    p = a[..., None] * np.sin(w * t + phi[..., None]) * active
    out = [p.astype(dtype)]
    if phase_shift is not None:
        for k, d in enumerate(phase_shift):
            v = amp2[k] * np.sin(w * t + phi[..., None] + d) * active
            out.append(v.astype(dtype))
    return a, out


def maps(shape, seed=1, dtype=np.float64):
    rng = np.random.default_rng(seed)
    c = rng.uniform(1450.0, 1600.0, shape).astype(dtype)
    rho = rng.uniform(950.0, 1100.0, shape).astype(dtype)
    alpha = rng.uniform(0.3, 1.2, shape).astype(dtype)
    return c, rho, alpha


def test_alpha_conversion_is_the_neper_definition():
    # 1 Np = 20 / ln(10) dB; 1 dB/cm at 1 MHz = 100 / 8.6858896 Np/m
    got = R.alpha_np_per_m(np.array([1.0]), 1e6)[0]
    assert got == pytest.approx(100.0 * math.log(10.0) / 20.0, rel=1e-9)
    assert R.alpha_np_per_m(np.array([2.0]), 3e6)[0] == pytest.approx(6 * got, rel=1e-12)


def test_auto_gate_is_the_burst():
    _, (p,) = burst()
    gate = R.auto_gate(p)
    np.testing.assert_array_equal(gate, np.arange(8, 32))
    # no energy at all: every sample reaches 10% of a zero peak, so the gate is all samples
    np.testing.assert_array_equal(R.auto_gate(np.zeros((2, 2, 2, 5))), np.arange(5))


@pytest.mark.parametrize("gate", [None, slice(8, 32), np.arange(8, 32)])
def test_force_closed_form(gate):
    a, (p,) = burst()
    c, rho, alpha = maps(a.shape)
    b = R.compute_radiation_force(p, c, rho, alpha, 2.5e6, gate)
    want = 2.0 * (alpha * 2.5) / R.DB_PER_NEPER * 100.0 * (a * a / 2.0) / (rho * c) / c
    np.testing.assert_allclose(b, want, rtol=1e-12)


def test_force_velocity_closed_form():
    shifts = (0.0, 0.7, 2.0)
    amps = (1.3, 0.4, 2.2)
    a, (p, vu, vw, vv) = burst(phase_shift=shifts, amp2=amps)
    c, rho, alpha = maps(a.shape)
    bx, by, bz = R.compute_radiation_force_velocity(p, vu, vw, vv, c, rho, alpha, 1e6)
    coeff = 2.0 * alpha / R.DB_PER_NEPER * 100.0 / c
    for b, d, bb in zip((bx, by, bz), shifts, amps):
        np.testing.assert_allclose(b, coeff * a * bb * math.cos(d) / 2.0, rtol=1e-11, atol=1e-14 * np.abs(coeff).max())


def test_float32_inputs_give_float32():
    a, (p,) = burst(dtype=np.float32)
    c, rho, alpha = maps(a.shape, dtype=np.float32)
    b = R.compute_radiation_force(p, c, rho, alpha, 1e6)
    assert b.dtype == np.float32
    b64 = R.compute_radiation_force(p.astype(np.float64), c.astype(np.float64), rho.astype(np.float64),
                                    alpha.astype(np.float64), 1e6)
    np.testing.assert_allclose(b, b64, rtol=2e-6)
