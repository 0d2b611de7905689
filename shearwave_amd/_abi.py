"""ctypes binding of the C ABI in ``include/shearwave.h``.

The HIP library is built in-tree (``python -m shearwave_amd.build``) into
``shearwave_amd/libshearwave_hip.so``. Nothing here computes: every function
forwards plain device pointers, sizes and a stream to the library. If the
library is missing, :func:`lib` raises instead of falling back to anything.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

LIB_PATH = os.environ.get("SHEARWAVE_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                             "libshearwave_hip.so")

SW_OK = 0
SW_ERR_UNSUPPORTED = 4
SW_F32, SW_F64 = 0, 1
SW_FORM_INCREMENT, SW_FORM_LEAPFROG = 0, 1
SW_KERNEL_AUTO, SW_KERNEL_DIRECT, SW_KERNEL_MARCH, SW_KERNEL_MARCH2, SW_KERNEL_MARCH3 = 0, 1, 2, 3, 4
CG_REASONS = {0: "running", 1: "converged", 2: "denom", 3: "nonfinite", 4: "max_iters", 5: "early_exit"}

_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int
_d = C.c_double


class NativeLibraryMissing(RuntimeError):
    """The gfx950 HIP library is not built or cannot be loaded."""


class SwError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"shearwave HIP error {code}: {msg}")
        self.code = code


class CgInfo(C.Structure):
    _fields_ = [("iters", _i64), ("relres", _d), ("max_div", _d), ("reason", C.c_int32), ("pad", C.c_int32),
                ("true_relres", _d)]


class StaticInfo(C.Structure):
    _fields_ = [("iters", _i64), ("relres", C.c_float), ("reason", C.c_int32), ("pad", C.c_int32)]


class FdtdState(C.Structure):
    _fields_ = [
        ("dtype", C.c_int32),
        ("form", C.c_int32),
        ("nx", _i64),
        ("ny", _i64),
        ("nz", _i64),
        ("dx", _d),
        ("dy", _d),
        ("dz", _d),
        ("dt", _d),
        ("vel_damp", _d),
        ("mu_over_rho", _p),
        ("eta_over_rho", _p),
        ("force", _p * 3),
        ("u", _p * 3),
        ("b", _p * 3),
        ("u_next", _p * 3),
        ("b_next", _p * 3),
        ("kernel", C.c_int32),
        ("tile_y", C.c_int32),
        ("x_chunk", C.c_int32),
        ("tile_z", C.c_int32),
    ]


class FdtdOutputs(C.Structure):
    _fields_ = [
        ("step0", _i64),
        ("snap_stride", _i64),
        ("snap_out", _p),
        ("inv_stride", _i64),
        ("inv_out", _p),
        ("trace_idx", _p),
        ("n_traces", _i64),
        ("trace_out", _p),
        ("trace_ld", _i64),
        ("finite_stride", _i64),
        ("nonfinite_step", _p),
        ("inv_mode", _i64),
    ]


class BatchSolve(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("xchunk", C.c_int32),
        ("cx", C.c_float),
        ("cy", C.c_float),
        ("cz", C.c_float),
        ("vx", C.c_float),
        ("vy", C.c_float),
        ("vz", C.c_float),
        ("damp", C.c_float),
        ("pad0", C.c_float),
        ("start", _i64),
        ("n_steps", _i64),
        ("mu_over_rho", _p),
        ("eta_over_rho", _p),
        ("force", _p * 3),
        ("env", _p),
        ("u", (_p * 3) * 2),
        ("d", (_p * 3) * 2),
        ("plan", _p),
        ("n_passes", _i64),
    ]


class BatchCg(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("xchunk", C.c_int32),
        ("dx", _d),
        ("dy", _d),
        ("dz", _d),
        ("tol", _d),
        ("max_iters", _i64),
        ("f", _p * 3),
        ("out", _p * 3),
        ("inv_rho", _p),
        ("scale", _d),
        ("phi", _p),
        ("r", _p),
        ("p0", _p),
        ("p1", _p),
        ("part", _p),
        ("item0", C.c_int32),
        ("n_items", C.c_int32),
    ]


# name -> (restype, argtypes); one entry per declaration of include/shearwave.h
SIGNATURES = {
    "sw_last_error": (C.c_char_p, []),
    "sw_abi_version": (_i32, []),
    "sw_device_info": (_i32, [C.POINTER(_i32), C.c_char_p, _i32]),
    "sw_laplacian_center": (_i32, [_i32, _p, _p, _i64, _i64, _i64, _d, _d, _d, _p]),
    "sw_divergence_center": (_i32, [_i32, _p, _p, _p, _p, _i64, _i64, _i64, _d, _d, _d, _p]),
    "sw_gradient_center": (_i32, [_i32, _p, _p, _p, _p, _i64, _i64, _i64, _d, _d, _d, _p]),
    "sw_project_workspace_bytes": (_i64, [_i64, _i64, _i64]),
    "sw_project_vector_field": (
        _i32,
        [_i32, _i32, _p, _p, _p, _i64, _i64, _i64, _d, _d, _d, _d, _i64, _p, _i64, C.POINTER(_p), _p, _d,
         C.POINTER(CgInfo), _p],
    ),
    "sw_prepare_coefficients": (_i32, [_i32, _p, _p, _p, _p, _p, _p, _i64, _p]),
    "sw_fdtd_run": (_i32, [C.POINTER(FdtdState), _i64, C.POINTER(_d), C.POINTER(FdtdOutputs), _p]),
    "sw_strain_tensor": (_i32, [_i32, _p, _p, _p, _p, _i64, _i64, _i64, _d, _d, _d, _p]),
    "sw_strain_invariants": (_i32, [_i32, _p, _i64, _p, _p, _p, _p, _p]),
    "sw_strain_invariants_from_u": (_i32, [_i32, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _d, _d, _d, _p]),
    "sw_pressure_energy_workspace": (_i64, [_i64, _i64]),
    "sw_pressure_energy": (_i32, [_i32, _p, _i64, _i64, _p, _p, _i64, _p]),
    "sw_radiation_force": (_i32, [_i32, _p, _i64, _i64, _p, _i64, _i64, _p, _p, _p, _d, _p, _p]),
    "sw_radiation_force_velocity": (_i32, [_i32, _p, _p, _p, _p, _i64, _i64, _p, _i64, _i64, _p, _p, _d, _p, _p, _p,
                                           _p]),
    "sw_acoustic_strain_volumetric": (_i32, [_i32, _p, _i64, _i64, _p, _i64, _i64, _p, _p, _p, _p]),
    "sw_acoustic_strain_peak": (_i32, [_i32, _p, _i64, _i64, _p, _i64, _i64, _p, _p, _p, _p, _p, _p]),
    "sw_acoustic_strain_gradient": (_i32, [_i32, _p, _i64, _i64, _i64, _i64, _p, _i64, _i64, _p, _p, _d, _d, _d, _p,
                                           _p, _p, _p]),
    "sw_acoustic_strain_gradient_peak": (_i32, [_i32, _p, _i64, _i64, _i64, _i64, _p, _i64, _i64, _p, _p, _d, _d, _d,
                                                _p, _p, _p, _p, _p, _p]),
    "sw_static_workspace_bytes": (_i64, [_i64, _i64, _i64]),
    "sw_elastic_solve_static": (_i32, [_p, _p, _d, _i64, _i64, _i64, _d, _d, _d, _d, _i32, _i64, _p, _p, _i64, _p,
                                       C.POINTER(StaticInfo), _p]),
    "sw_aos3_to_soa": (_i32, [_i32, _p, _p, _p, _p, _i64, _p]),
    "sw_soa_to_aos3": (_i32, [_i32, _p, _p, _p, _p, _i64, _p]),
    "sw_export_state": (_i32, [_i32, _i32, C.POINTER(_p), C.POINTER(_p), _d, _p, _p, _i64, _p]),
    "sw_import_state": (_i32, [_i32, _i32, _p, _p, _d, C.POINTER(_p), C.POINTER(_p), _i64, _i64, _i64, _p]),
    "sw_combine3": (_i32, [_i32, C.POINTER(_p), C.POINTER(_p), _d, _d, C.POINTER(_p), _i64, _p]),
    "sw_project_preconditioned": (_i32, [_i32, _i64, _i64, _i64, _d, _i64]),
    "sw_project_noalloc_max_voxels": (_i64, []),
    "sw_project_vector_field_noalloc": (_i32, [_i32, _i32, _p, _p, _p, _i64, _i64, _i64, _d, _d, _d, _d, _i64,
                                                C.POINTER(_p), _p, _d, C.POINTER(CgInfo), _p]),
    "sw_gather_traces": (_i32, [_i32, _p, _i64, _p, _i64, _p, _i64, _p]),
    "sw_batch_step": (_i32, [_p, _p, _i64, _p, _i32, _i32, _i64, _p]),
    "sw_batch_graph_create": (_i32, [_p, _p, _i64, _p, _i32, _i32, _i64, C.POINTER(_p)]),
    "sw_graph_launch": (_i32, [_p, _p]),
    "sw_graph_destroy": (_i32, [_p]),
    "sw_cg_state_bytes": (_i64, []),
    "sw_batch_project": (_i32, [_p, _i32, _p, _i64, _p, C.POINTER(CgInfo), _p]),
    "sw_adjoint_step": (_i32, [_i32, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p), C.POINTER(_p), C.POINTER(_p),
                               C.POINTER(_p), _p, _p, _p, _p, _i64, _i64, _i64, _d, _d, _d, _d, _d, _p]),
    "sw_inject_traces": (_i32, [_i32, _p, _i64, _p, _i64, _p, _i64, _p]),
    "sw_fdtd_step2_mid": (_i32, [C.POINTER(FdtdState), C.POINTER(_d), _p, _p, _p]),
    "sw_fdtd_trial_passes": (_i32, [C.POINTER(FdtdState), _i64, _p]),
}

_lock = threading.Lock()
_lib = None


# SW_ABI_VERSION of include/shearwave.h that these ctypes declarations follow.
ABI_VERSION = 5


def lib():
    """Load the HIP library once; raise NativeLibraryMissing if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is not built; run `python -m shearwave_amd.build` (hipcc, gfx950)"
            )
        try:
            handle = C.CDLL(LIB_PATH)
        except OSError as e:
            raise NativeLibraryMissing(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        if handle.sw_abi_version() != ABI_VERSION:  # a stale build would misread the structs
            raise NativeLibraryMissing(f"{LIB_PATH} has ABI version {handle.sw_abi_version()}, this package "
                                       f"needs {ABI_VERSION}; rebuild with `python -m shearwave_amd.build`")
        _lib = handle
    return _lib


def check(rc):
    if rc != SW_OK:
        raise SwError(rc, lib().sw_last_error().decode(errors="replace"))


def ptrs(*tensors):
    """A ctypes array of device pointers (None -> NULL)."""
    arr = (_p * len(tensors))()
    for k, t in enumerate(tensors):
        arr[k] = None if t is None else t.data_ptr()
    return arr
