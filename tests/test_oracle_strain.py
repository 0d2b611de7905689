"""Pins for the strain oracle: the closed forms of the reference's tests/test_strain.py."""

import numpy as np
import pytest

from oracle import strain as S

H = (1e-3, 1.3e-3, 0.8e-3)


def test_constant_displacement_has_zero_strain():
    n = 12
    u = [np.full((n, n, n), c) for c in (3.0, -1.0, 0.5)]
    eps = S.compute_strain_tensor(*u, *H)
    assert eps.shape == (n, n, n, 6)
    np.testing.assert_allclose(eps, 0.0, atol=1e-14)


def test_linear_fields_give_exact_components():
    n = (9, 10, 11)
    x = np.arange(n[0])[:, None, None] * H[0]
    y = np.arange(n[1])[None, :, None] * H[1]
    z = np.arange(n[2])[None, None, :] * H[2]
    one = np.ones(n)
    ux = (0.01 * x + 0.002 * y - 0.003 * z) * one
    uy = (0.004 * x - 0.02 * y + 0.005 * z) * one
    uz = (-0.006 * x + 0.007 * y + 0.03 * z) * one
    eps = S.compute_strain_tensor(ux, uy, uz, *H)
    want = [0.01, -0.02, 0.03, 0.005 + 0.007, -0.003 - 0.006, 0.002 + 0.004]
    for k, w in enumerate(want):
        np.testing.assert_allclose(eps[1:-1, 1:-1, 1:-1, k], w, rtol=1e-9)
        assert np.all(eps[0, :, :, k] == 0) and np.all(eps[:, :, -1, k] == 0)


def test_invariants_closed_forms():
    e = 0.03
    eps = np.zeros((4, 4, 4, 6))
    eps[..., 0] = e
    inv = S.compute_strain_invariants(eps)
    np.testing.assert_allclose(inv["von_mises"], 2 * e / 3, rtol=1e-12)
    np.testing.assert_allclose(inv["volumetric"], e / 3, rtol=1e-12)
    eps = np.zeros((4, 4, 4, 6))
    eps[..., :3] = 0.01
    inv = S.compute_strain_invariants(eps)
    np.testing.assert_allclose(inv["deviatoric"], 0.0, atol=1e-14)
    np.testing.assert_allclose(inv["max_shear"], 0.0, atol=1e-14)
    eps = np.zeros((4, 4, 4, 6))
    eps[..., 5] = 0.02  # engineering gxy: tensor shear 0.01, principal +-0.01
    inv = S.compute_strain_invariants(eps)
    np.testing.assert_allclose(inv["volumetric"], 0.0, atol=1e-14)
    np.testing.assert_allclose(inv["deviatoric"], np.sqrt(2) * 0.01, rtol=1e-12)


def test_invariant_identities_on_random_tensors():
    rng = np.random.default_rng(1)
    eps = rng.standard_normal((7, 5, 6, 6))
    inv = S.compute_strain_invariants(eps)
    # deviatoric Frobenius norm from the full symmetric tensor
    t = np.zeros(eps.shape[:3] + (3, 3))
    for (i, j), k in (((0, 0), 0), ((1, 1), 1), ((2, 2), 2)):
        t[..., i, j] = eps[..., k]
    for (i, j), k in (((1, 2), 3), ((0, 2), 4), ((0, 1), 5)):
        t[..., i, j] = t[..., j, i] = 0.5 * eps[..., k]
    tr = np.trace(t, axis1=-2, axis2=-1)
    dev = t - tr[..., None, None] / 3 * np.eye(3)
    fro = np.sqrt(np.sum(dev**2, axis=(-2, -1)))
    np.testing.assert_allclose(inv["deviatoric"], fro, rtol=1e-12)
    np.testing.assert_allclose(inv["volumetric"], tr / 3, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(inv["von_mises"], np.sqrt(2 / 3) * fro, rtol=1e-12)
    np.testing.assert_allclose(inv["max_shear"], fro / np.sqrt(2), rtol=1e-12)


def test_errors():
    z = np.zeros((5, 5, 5))
    with pytest.raises(ValueError, match="same shape"):
        S.compute_strain_tensor(z, np.zeros((5, 5, 6)), z, *H)
    with pytest.raises(ValueError, match="positive"):
        S.compute_strain_tensor(z, z, z, -1e-3, 1e-3, 1e-3)
    with pytest.raises(ValueError, match="3-D"):
        S.compute_strain_tensor(np.zeros((5, 5)), z, z, *H)
    with pytest.raises(ValueError, match="Voigt"):
        S.compute_strain_invariants(np.zeros((3, 3, 3, 3)))
