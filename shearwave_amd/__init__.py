"""shearwave_amd: shearwave's shear_fdtd hot path, MI355X-native.

The viscoelastic shear FDTD time loop, its Helmholtz-Hodge force projection,
the strain post-processing, the acoustic radiation force that produces the
body force, the acoustic strain statistics and the static elasticity solve, from pinton-lab/shearwave,
re-implemented as
hand-written gfx950 HIP kernels behind a C ABI (``include/shearwave.h``) and a
Python drop-in with the reference's API. Build the library with
``python -m shearwave_amd.build``.
"""

from ._abi import NativeLibraryMissing, SwError, lib
from .api import (
    compute_acoustic_strain_gradient,
    compute_acoustic_strain_gradient_peak,
    compute_acoustic_strain_peak,
    compute_acoustic_strain_volumetric,
    compute_radiation_force,
    compute_radiation_force_velocity,
    compute_strain_invariants,
    compute_strain_tensor,
    divergence_center,
    elastic_solve_static,
    gradient_center,
    hann1d,
    laplacian_center,
    project_body_force_to_shear,
    project_vector_field,
    shear_fdtd,
    shear_fdtd_jax,
)
from .engine import Solve

__version__ = "0.1.0"


def require_native():
    """Load the HIP library; raise NativeLibraryMissing if it is not built."""
    return lib()


__all__ = [
    "NativeLibraryMissing",
    "Solve",
    "SwError",
    "__version__",
    "compute_acoustic_strain_gradient",
    "compute_acoustic_strain_gradient_peak",
    "compute_acoustic_strain_peak",
    "compute_acoustic_strain_volumetric",
    "compute_radiation_force",
    "compute_radiation_force_velocity",
    "compute_strain_invariants",
    "compute_strain_tensor",
    "divergence_center",
    "elastic_solve_static",
    "gradient_center",
    "hann1d",
    "laplacian_center",
    "project_body_force_to_shear",
    "project_vector_field",
    "require_native",
    "shear_fdtd",
    "shear_fdtd_jax",
]
