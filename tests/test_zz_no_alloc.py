"""sw_project_vector_field_noalloc: the projection with the library's static scratch instead of a workspace.

It must give what sw_project_vector_field gives with a caller's workspace,
bit for bit (same kernels, same launch geometry), on two calls in a row, the
second on a smaller grid (the scratch keeps the first call's contents), and
reject a grid beyond its maximum before any device call.
"""

import ctypes

import numpy as np
import pytest

from shearwave_amd import _abi


def _gauss(shape, sigma):
    c = [(n - 1) / 2 for n in shape]
    g = np.meshgrid(*(np.arange(n) - cc for n, cc in zip(shape, c)), indexing="ij")
    return np.exp(-(g[0] ** 2 + g[1] ** 2 + g[2] ** 2) / (2 * sigma**2))


def _inputs(shape, dtype, seed):
    import torch

    rng = np.random.default_rng(seed)
    f = np.stack([rng.standard_normal(shape), rng.standard_normal(shape), 300.0 * _gauss(shape, 4.0)])
    inv_rho = 1.0 / (1000.0 * (1 + rng.uniform(-0.01, 0.01, shape)))
    tdt = getattr(torch, dtype)
    return torch.from_numpy(f).to("cuda", tdt), torch.from_numpy(inv_rho).to("cuda", tdt)


def _project(lib, f, inv_rho, dtype, vec, tol, workspace):
    """One projection through the C ABI; returns (f_S, scaled force, info)."""
    import torch

    from shearwave_amd import engine as E

    f = f.clone()
    scaled = torch.empty_like(f)
    info = _abi.CgInfo()
    code = _abi.SW_F32 if dtype == "float32" else _abi.SW_F64
    vcode = _abi.SW_F32 if vec == "float32" else _abi.SW_F64
    shape = tuple(f.shape[1:])
    h = (0.3e-3, 0.25e-3, 0.35e-3)
    so = _abi.ptrs(*scaled)
    if workspace is None:
        rc = lib.sw_project_vector_field_noalloc(code, vcode, f[0].data_ptr(), f[1].data_ptr(), f[2].data_ptr(),
                                                 *shape, *h, tol, 5000, so, inv_rho.data_ptr(), 4e-12,
                                                 ctypes.byref(info), E.stream_ptr())
    else:
        rc = lib.sw_project_vector_field(code, vcode, f[0].data_ptr(), f[1].data_ptr(), f[2].data_ptr(), *shape, *h,
                                         tol, 5000, workspace.data_ptr(), workspace.numel(), so, inv_rho.data_ptr(),
                                         4e-12, ctypes.byref(info), E.stream_ptr())
    _abi.check(rc)
    torch.cuda.synchronize()
    rep = (info.iters, info.relres, info.max_div, info.reason, info.true_relres)
    return f.cpu().numpy(), scaled.cpu().numpy(), rep


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,vec,tol", [("float32", "float32", 1e-8), ("float64", "float64", 1e-10)])
def test_noalloc_projection_equals_workspace_projection(sw, dtype, vec, tol):
    import torch

    lib = _abi.lib()
    # the first call is the largest: the second finds its contents in the scratch
    for shape, seed in (((72, 88, 130), 1), ((20, 18, 40), 2)):
        assert np.prod(shape) <= lib.sw_project_noalloc_max_voxels()
        f, inv_rho = _inputs(shape, dtype, seed)
        ws = torch.empty(lib.sw_project_workspace_bytes(*shape), dtype=torch.uint8, device="cuda")
        want = _project(lib, f, inv_rho, dtype, vec, tol, ws)
        got = _project(lib, f, inv_rho, dtype, vec, tol, None)
        # converged (float32 vectors: the spectral-preconditioned CG, a few iterations; float64: the plain CG)
        assert want[2][3] == 1 and want[2][0] >= 1 and want[2][4] < tol, want[2]
        np.testing.assert_array_equal(got[0], want[0])
        np.testing.assert_array_equal(got[1], want[1])
        assert got[2] == want[2]


def test_noalloc_projection_rejects_large_grids():
    """Beyond 2^20 voxels the call fails with SW_ERR_ARG (1) before touching a device."""
    lib = _abi.lib()
    assert lib.sw_project_noalloc_max_voxels() == 1 << 20
    info = _abi.CgInfo()
    rc = lib.sw_project_vector_field_noalloc(_abi.SW_F32, _abi.SW_F32, 1, 1, 1, 128, 128, 65, 1.0, 1.0, 1.0, 1e-8,
                                             10, None, None, 1.0, ctypes.byref(info), None)
    assert rc == 1 and b"no-allocation" in lib.sw_last_error()
    rc = lib.sw_project_vector_field_noalloc(_abi.SW_F32, _abi.SW_F32, 1, 1, 1, 0, 8, 8, 1.0, 1.0, 1.0, 1e-8, 10,
                                             None, None, 1.0, ctypes.byref(info), None)
    assert rc == 1
