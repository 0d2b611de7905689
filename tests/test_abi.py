"""The C-ABI library loads and exports every symbol include/shearwave.h declares (no GPU calls)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shearwave.h")
LIB = os.path.join(ROOT, "shearwave_amd", "libshearwave_hip.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\*\s]+?\b(sw_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("sw_fdtd_run", "sw_project_vector_field", "sw_strain_tensor", "sw_strain_invariants",
              "sw_laplacian_center", "sw_divergence_center", "sw_gradient_center", "sw_last_error"):
        assert s in syms


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    lib.sw_abi_version.restype = ctypes.c_int
    assert lib.sw_abi_version() == 5


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_python_binding_covers_header():
    from shearwave_amd import _abi

    assert set(declared_symbols()) == set(_abi.SIGNATURES)
    lib = _abi.lib()
    # argument checking happens before any device call
    rc = lib.sw_laplacian_center(7, None, None, 4, 4, 4, 1.0, 1.0, 1.0, None)
    assert rc == 1 and b"dtype" in lib.sw_last_error()
    rc = lib.sw_fdtd_run(None, 1, None, None, None)
    assert rc == 1 and b"null state" in lib.sw_last_error()
    assert lib.sw_project_workspace_bytes(8, 8, 8) > 4 * 8 * 512
    # the CG vector storage is SW_F32 or SW_F64, like the field storage
    info = _abi.CgInfo()
    rc = lib.sw_project_vector_field(_abi.SW_F32, 5, None, None, None, 8, 8, 8, 1.0, 1.0, 1.0, 1e-8, 10, None, 0,
                                     None, None, 1.0, ctypes.byref(info), None)
    assert rc == 1 and b"dtype" in lib.sw_last_error()
    rc = lib.sw_project_vector_field(_abi.SW_F32, _abi.SW_F32, None, None, None, 8, 8, 8, 1.0, 1.0, 1.0, 1e-8, 10,
                                     None, 0, None, None, 1.0, ctypes.byref(info), None)
    assert rc == 1 and b"null force" in lib.sw_last_error()
    # the trace kernels take the field's size, so a bad index cannot reach outside it (checked in-kernel)
    rc = lib.sw_gather_traces(_abi.SW_F32, None, -1, None, 4, None, 1, None)
    assert rc == 1 and b"n" in lib.sw_last_error()
    rc = lib.sw_inject_traces(_abi.SW_F32, None, -5, None, 4, None, 1, None)
    assert rc == 1 and b"n" in lib.sw_last_error()
    # a non-finite check needs its device flag
    st = _abi.FdtdState()
    st.dtype, st.form, st.nx, st.ny, st.nz = _abi.SW_F32, _abi.SW_FORM_INCREMENT, 8, 8, 8
    st.dx = st.dy = st.dz = st.dt = 1.0
    st.mu_over_rho = 1
    for c in range(3):
        st.u[c] = st.b[c] = st.u_next[c] = st.b_next[c] = 1
    out = _abi.FdtdOutputs()
    out.finite_stride = 10
    rc = lib.sw_fdtd_run(ctypes.byref(st), 1, None, ctypes.byref(out), None)
    assert rc == 1 and b"nonfinite_step" in lib.sw_last_error()


def test_struct_layouts_match_header():
    """The ctypes structures have the sizes the header's C layout gives them."""
    from shearwave_amd import _abi

    assert ctypes.sizeof(_abi.CgInfo) == 40  # int64, 2 double, 2 int32, double
    assert ctypes.sizeof(_abi.FdtdOutputs) == 12 * 8


def test_product_does_not_import_the_oracle():
    """The product package never reaches the oracle (test infrastructure only)."""
    pkg = os.path.join(ROOT, "shearwave_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle", src, flags=re.M), f
