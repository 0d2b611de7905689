"""CPU oracle for the acoustic radiation force: a NumPy restatement of the reference.

THIS MODULE IS TEST INFRASTRUCTURE (same rule as ``oracle/solver.py``): only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import it, as the checker.

Restates ``src/shearwave/radiation_force.py``. The reference has no tests or
stored vectors for this module; the restatement is pinned by closed forms
(tests/test_oracle_radiation_force.py): a sinusoidal burst sampled at whole
periods has mean(p^2) = A^2/2 and mean(p v) = A B cos(phase)/2 exactly, the
auto-gate of a burst is the burst's sample range, and the unit conversion
dB/cm/MHz -> Np/m is checked against its definition (1 Np = 20/ln 10 dB).
"""

from __future__ import annotations

import numpy as np

DB_PER_NEPER = 8.685889638  # radiation_force.py:79, 149 (20 / ln 10, as written in the reference)


def auto_gate(pressure_4d):
    """Time indices where the spatial mean of p^2 reaches 10% of its peak (radiation_force.py:66-72).

    Returns ``slice(None)`` when no sample qualifies, as the reference does.
    The mean and the comparison run in the pressure's dtype.
    """
    p_sq = pressure_4d**2
    env = p_sq.mean(axis=(0, 1, 2))
    threshold = 0.1 * env.max()
    gate_mask = env >= threshold
    return slice(None) if not gate_mask.any() else np.nonzero(gate_mask)[0]


def alpha_np_per_m(alpha_coeff_map, f0):
    """dB/cm/MHz -> Np/m at f0 (radiation_force.py:78-79)."""
    f0_mhz = f0 / 1e6
    return (alpha_coeff_map * f0_mhz) / DB_PER_NEPER * 100.0


def compute_radiation_force(pressure_4d, c_map, rho_map, alpha_coeff_map, f0, gate=None):
    """b = 2 alpha_Np Isppa / c with Isppa = mean_gate(p^2) / (rho c) (radiation_force.py:27-83)."""
    if gate is None:
        gate = auto_gate(pressure_4d)
    z_map = rho_map * c_map
    isppa = np.mean(pressure_4d[:, :, :, gate] ** 2, axis=3) / z_map
    return 2.0 * alpha_np_per_m(alpha_coeff_map, f0) * isppa / c_map


def compute_radiation_force_velocity(pressure_4d, velocity_u_4d, velocity_w_4d, velocity_v_4d, c_map, rho_map,
                                     alpha_coeff_map, f0, gate=None):
    """f_i = 2 alpha_Np mean_gate(p v_i) / c for the three velocity components (radiation_force.py:86-157)."""
    del rho_map  # kept for interface parity, as in the reference (radiation_force.py:91)
    if gate is None:
        gate = auto_gate(pressure_4d)
    p_gated = pressure_4d[:, :, :, gate]
    ix = np.mean(p_gated * velocity_u_4d[:, :, :, gate], axis=3)
    iy = np.mean(p_gated * velocity_w_4d[:, :, :, gate], axis=3)
    iz = np.mean(p_gated * velocity_v_4d[:, :, :, gate], axis=3)
    coeff = 2.0 * alpha_np_per_m(alpha_coeff_map, f0) / c_map
    return coeff * ix, coeff * iy, coeff * iz
