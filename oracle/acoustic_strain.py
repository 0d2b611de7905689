"""CPU oracle for the acoustic strain statistics: a NumPy restatement of the reference.

THIS MODULE IS TEST INFRASTRUCTURE (same rule as ``oracle/solver.py``): only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import it, as the checker.

Restates ``src/shearwave/acoustic_strain.py``. The reference has no tests for
this module; the restatement is pinned by closed forms
(tests/test_oracle_acoustic_strain.py): a pressure that is linear in space
has a known strain gradient at every interior voxel, a sinusoidal burst over
whole periods has known peak, RMS and mean, and brute-force loops over a
small record check the gradient statistics element by element.
"""

from __future__ import annotations

import numpy as np

from .radiation_force import auto_gate
from .solver import gradient_center


def bulk_modulus(c_map, rho_map):
    """K = rho c^2 (acoustic_strain.py:33-38)."""
    return rho_map * c_map**2


def resolve_gate(pressure_4d, gate):
    """The caller's gate, or the 10%-of-peak auto-gate (acoustic_strain.py:41-67)."""
    return auto_gate(pressure_4d) if gate is None else gate


def compute_acoustic_strain_volumetric(pressure_4d, c_map, rho_map, gate=None):
    """epsilon = -p / K over the gated samples (acoustic_strain.py:75-104)."""
    gate = resolve_gate(pressure_4d, gate)
    K = bulk_modulus(c_map, rho_map)
    return -pressure_4d[:, :, :, gate] / K[..., np.newaxis]


def compute_acoustic_strain_peak(pressure_4d, c_map, rho_map, gate=None):
    """peak = max|p| / K, rms = sqrt(mean p^2) / K, mean = -mean(p) / K (acoustic_strain.py:107-151)."""
    gate = resolve_gate(pressure_4d, gate)
    K = bulk_modulus(c_map, rho_map)
    p_gated = pressure_4d[:, :, :, gate]
    peak = np.max(np.abs(p_gated), axis=3) / K
    rms = np.sqrt(np.mean(p_gated**2, axis=3)) / K
    mean = -np.mean(p_gated, axis=3) / K
    return {"peak": peak, "rms": rms, "mean": mean}


def compute_acoustic_strain_gradient(pressure_4d, c_map, rho_map, dX, dY, dZ, gate=None):
    """gradient_center of -p_t / K for every gated sample t (acoustic_strain.py:154-203)."""
    gate = resolve_gate(pressure_4d, gate)
    K = bulk_modulus(c_map, rho_map)
    p_gated = pressure_4d[:, :, :, gate]
    nt = p_gated.shape[3]
    shape = K.shape + (nt,)
    grad_x = np.zeros(shape, dtype=p_gated.dtype)
    grad_y = np.zeros_like(grad_x)
    grad_z = np.zeros_like(grad_x)
    for t in range(nt):
        gx, gy, gz = gradient_center(-p_gated[:, :, :, t] / K, dX, dY, dZ)
        grad_x[:, :, :, t] = gx
        grad_y[:, :, :, t] = gy
        grad_z[:, :, :, t] = gz
    return grad_x, grad_y, grad_z


def compute_acoustic_strain_gradient_peak(pressure_4d, c_map, rho_map, dX, dY, dZ, gate=None):
    """Running max / sum over t of the strain gradient, in float64 (acoustic_strain.py:206-277)."""
    gate = resolve_gate(pressure_4d, gate)
    K = bulk_modulus(c_map, rho_map)
    p_gated = pressure_4d[:, :, :, gate]
    nt = p_gated.shape[3]
    peak_mag = np.zeros(K.shape, dtype=np.float64)
    peak_gx = np.zeros_like(peak_mag)
    peak_gy = np.zeros_like(peak_mag)
    peak_gz = np.zeros_like(peak_mag)
    sum_mag_sq = np.zeros_like(peak_mag)
    for t in range(nt):
        strain_t = -p_gated[:, :, :, t].astype(np.float64) / K
        gx, gy, gz = gradient_center(strain_t, dX, dY, dZ)
        mag = np.sqrt(gx**2 + gy**2 + gz**2)
        np.maximum(peak_mag, mag, out=peak_mag)
        np.maximum(peak_gx, np.abs(gx), out=peak_gx)
        np.maximum(peak_gy, np.abs(gy), out=peak_gy)
        np.maximum(peak_gz, np.abs(gz), out=peak_gz)
        sum_mag_sq += mag**2
    return {
        "peak_magnitude": peak_mag,
        "peak_grad_x": peak_gx,
        "peak_grad_y": peak_gy,
        "peak_grad_z": peak_gz,
        "rms_magnitude": np.sqrt(sum_mag_sq / max(nt, 1)),
    }
