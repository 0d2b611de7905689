"""Pins of the acoustic-strain oracle (closed forms and brute force; the reference has no tests here)."""

import math

import numpy as np
import pytest

from oracle import acoustic_strain as A


def _maps(shape, seed=0, homogeneous=False):
    rng = np.random.default_rng(seed)
    if homogeneous:
        return np.full(shape, 1500.0), np.full(shape, 1000.0)
    return rng.uniform(1450, 1600, shape), rng.uniform(950, 1100, shape)


def test_volumetric_is_minus_p_over_rho_c2():
    rng = np.random.default_rng(1)
    p = rng.standard_normal((3, 4, 5, 9)) * 1e5
    c, rho = _maps((3, 4, 5))
    eps = A.compute_acoustic_strain_volumetric(p, c, rho, slice(2, 7))
    for idx in np.ndindex(3, 4, 5):
        for k, t in enumerate(range(2, 7)):
            assert eps[idx + (k,)] == -p[idx + (t,)] / (rho[idx] * c[idx] ** 2)


def test_peak_rms_mean_of_a_whole_period_sinusoid():
    shape, nt, spp = (4, 3, 5), 24, 8
    amp = np.random.default_rng(2).uniform(1e4, 1e5, shape)
    t = np.arange(nt)
    p = amp[..., None] * np.sin(2 * np.pi * t / spp)  # sin = 1 at t = 2 (mod 8)
    c, rho = _maps(shape)
    K = rho * c * c
    st = A.compute_acoustic_strain_peak(p, c, rho, slice(None))
    np.testing.assert_allclose(st["peak"], amp / K, rtol=1e-15)
    np.testing.assert_allclose(st["rms"], amp / math.sqrt(2) / K, rtol=1e-14)
    assert np.abs(st["mean"]).max() < 1e-13 * (amp / K).max()


def test_gradient_of_a_linear_field_is_exact():
    shape, nt = (6, 5, 7), 4
    dX, dY, dZ = 0.3e-3, 0.5e-3, 0.4e-3
    x, y, z = np.meshgrid(np.arange(shape[0]) * dX, np.arange(shape[1]) * dY, np.arange(shape[2]) * dZ, indexing="ij")
    w = np.array([1.0, -2.0, 0.5, 3.0])
    a, b, cc, d = 2e6, -1e6, 3e6, 5e3
    p = (a * x + b * y + cc * z + d)[..., None] * w
    c, rho = _maps(shape, homogeneous=True)
    K = 1000.0 * 1500.0**2
    gx, gy, gz = A.compute_acoustic_strain_gradient(p, c, rho, dX, dY, dZ, slice(None))
    inner = (slice(1, -1),) * 3
    for g, slope in ((gx, a), (gy, b), (gz, cc)):
        np.testing.assert_allclose(g[inner], np.broadcast_to(-slope * w / K, g[inner].shape), rtol=1e-9)
        edge = np.ones(shape, bool)
        edge[inner] = False
        assert np.all(g[edge] == 0)


def test_gradient_peak_brute_force():
    rng = np.random.default_rng(3)
    shape, nt = (5, 4, 6), 7
    p = (rng.standard_normal(shape + (nt,)) * 1e5).astype(np.float32)
    c, rho = (m.astype(np.float32) for m in _maps(shape, 4))
    h = (0.2e-3, 0.3e-3, 0.25e-3)
    gate = np.array([0, 2, 3, 6])
    st = A.compute_acoustic_strain_gradient_peak(p, c, rho, *h, gate)
    K = (rho * c**2).astype(np.float64)
    s = -p.astype(np.float64) / K[..., None]
    want = {k: np.zeros(shape) for k in ("peak_magnitude", "peak_grad_x", "peak_grad_y", "peak_grad_z", "sumsq")}
    for i, j, k in np.ndindex(*shape):
        if not (0 < i < shape[0] - 1 and 0 < j < shape[1] - 1 and 0 < k < shape[2] - 1):
            continue
        for t in gate:
            gx = (s[i + 1, j, k, t] - s[i - 1, j, k, t]) / (2 * h[0])
            gy = (s[i, j + 1, k, t] - s[i, j - 1, k, t]) / (2 * h[1])
            gz = (s[i, j, k + 1, t] - s[i, j, k - 1, t]) / (2 * h[2])
            m = math.sqrt(gx * gx + gy * gy + gz * gz)
            want["peak_magnitude"][i, j, k] = max(want["peak_magnitude"][i, j, k], m)
            want["peak_grad_x"][i, j, k] = max(want["peak_grad_x"][i, j, k], abs(gx))
            want["peak_grad_y"][i, j, k] = max(want["peak_grad_y"][i, j, k], abs(gy))
            want["peak_grad_z"][i, j, k] = max(want["peak_grad_z"][i, j, k], abs(gz))
            want["sumsq"][i, j, k] += m * m
    for key in ("peak_magnitude", "peak_grad_x", "peak_grad_y", "peak_grad_z"):
        np.testing.assert_allclose(st[key], want[key], rtol=1e-14, atol=0)
    np.testing.assert_allclose(st["rms_magnitude"], np.sqrt(want["sumsq"] / len(gate)), rtol=1e-14)
    assert all(v.dtype == np.float64 for v in st.values())


@pytest.mark.parametrize("fn", ["compute_acoustic_strain_peak", "compute_acoustic_strain_volumetric"])
def test_auto_gate_matches_radiation_force(fn):
    from oracle.radiation_force import auto_gate

    rng = np.random.default_rng(5)
    t = np.arange(30)
    p = rng.standard_normal((4, 4, 4, 30)) * np.where((t > 8) & (t < 20), 1.0, 0.01)
    c, rho = _maps((4, 4, 4))
    got = getattr(A, fn)(p, c, rho)
    want = getattr(A, fn)(p, c, rho, auto_gate(p))
    for g, w in zip(got.values() if isinstance(got, dict) else [got], want.values() if isinstance(want, dict) else [want]):
        np.testing.assert_array_equal(g, w)
