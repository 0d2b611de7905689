"""CPU oracle for strain post-processing: a NumPy restatement of the reference.

THIS MODULE IS TEST INFRASTRUCTURE (same rule as ``oracle/solver.py``): only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import it, as the checker.

Restates ``src/shearwave/strain.py``. Pinned by the closed forms of the
reference's ``tests/test_strain.py`` (constant, linear, pure-shear
displacements; dilation, pure shear and uniaxial invariants) plus the
isotropy identity ``max_shear = dev / sqrt(2)`` and ``von_mises =
sqrt(2/3) dev`` on random tensors.
"""

from __future__ import annotations

import numpy as np

from .solver import gradient_center


def _validate_inputs(ux, uy, uz, dX, dY, dZ):
    """strain.py:138-159."""
    if ux.ndim != 3:
        raise ValueError(f"Displacement arrays must be 3-D, got ndim={ux.ndim}")
    if ux.shape != uy.shape or ux.shape != uz.shape:
        raise ValueError(
            "All displacement components must share the same shape: "
            f"ux={ux.shape}, uy={uy.shape}, uz={uz.shape}"
        )
    for name, val in (("dX", dX), ("dY", dY), ("dZ", dZ)):
        if val <= 0:
            raise ValueError(f"Grid spacing {name} must be positive, got {val}")


def compute_strain_tensor(ux, uy, uz, dX, dY, dZ):
    """Voigt strain ``[exx, eyy, ezz, gyz, gxz, gxy]`` with engineering shear (strain.py:16-70)."""
    _validate_inputs(ux, uy, uz, dX, dY, dZ)
    dtype = np.result_type(ux.dtype, uy.dtype, uz.dtype)
    gx = gradient_center(ux, dX, dY, dZ)
    gy = gradient_center(uy, dX, dY, dZ)
    gz = gradient_center(uz, dX, dY, dZ)
    eps = np.zeros(ux.shape + (6,), dtype=dtype)
    eps[..., 0] = gx[0]
    eps[..., 1] = gy[1]
    eps[..., 2] = gz[2]
    eps[..., 3] = gy[2] + gz[1]
    eps[..., 4] = gx[2] + gz[0]
    eps[..., 5] = gx[1] + gy[0]
    return eps


def compute_strain_invariants(epsilon):
    """Volumetric, deviatoric norm, max shear and von Mises strain (strain.py:73-135)."""
    if epsilon.shape[-1] != 6:
        raise ValueError(f"Expected last dimension == 6 (Voigt notation), got {epsilon.shape[-1]}")
    exx, eyy, ezz, gyz, gxz, gxy = (epsilon[..., k] for k in range(6))
    vol = (exx + eyy + ezz) / 3.0
    dxx, dyy, dzz = exx - vol, eyy - vol, ezz - vol
    q = dxx**2 + dyy**2 + dzz**2 + 0.5 * (gyz**2 + gxz**2 + gxy**2)
    dev = np.sqrt(q)
    von_mises = np.sqrt(2.0 / 3.0) * dev
    j2 = 0.5 * (dxx**2 + dyy**2 + dzz**2 + 0.5 * (gyz**2 + gxz**2 + gxy**2))
    max_shear = np.sqrt(np.maximum(j2, 0.0))
    return {"volumetric": vol, "deviatoric": dev, "max_shear": max_shear, "von_mises": von_mises}
