"""The drop-in's argument checks and messages (solver.py:88-171, 281-312, 1060-1079; strain.py:98-100, 138-159).

All of these are raised before any device work, so they run without a GPU.
"""

import numpy as np
import pytest

import shearwave_amd as sw

B = np.zeros((6, 7, 8), np.float32)
ARGS = (1000.0, 3000.0, 1e-3, 1e-3, 1e-3, 1e-6)


def run(*a, **k):
    return sw.shear_fdtd(*a, **k)


def test_shear_fdtd_errors():
    with pytest.raises(ValueError, match="At least one body-force component must be provided."):
        run(None, None, None, *ARGS, 3)
    with pytest.raises(ValueError, match="reproject_every must be a non-negative integer."):
        run(B, None, None, *ARGS, 3, {"reproject_every": -1})
    with pytest.raises(ValueError, match=r"env_t length \(2\) must equal n_steps \(3\)\."):
        run(B, None, None, *ARGS, 3, {"env_t": [1.0, 0.0]})
    with pytest.raises(ValueError, match="Material field has incompatible shape."):
        run(B, None, None, np.ones((6, 7, 9)), 3000.0, 1e-3, 1e-3, 1e-3, 1e-6, 3, {"env_t": [1, 1, 1]})
    with pytest.raises(ValueError, match="Material field has incompatible shape."):
        run(B, None, None, *ARGS, 3, {"env_t": [1, 1, 1], "eta": np.ones((2, 2, 2))})
    with pytest.raises(ValueError, match="Body-force component has incompatible shape."):
        run(B, np.zeros((6, 7, 9)), None, *ARGS, 3)
    with pytest.raises(ValueError, match="Body-force component has incompatible shape."):
        run(np.zeros((6, 7, 8, 4)), None, None, *ARGS, 3)
    with pytest.raises(ValueError, match=r"initial_u must match shape \(nx,ny,nz,3\)"):
        run(B, None, None, *ARGS, 3, {"env_t": [1, 1, 1], "initial_u": np.zeros((6, 7, 8))})
    with pytest.raises(ValueError, match=r"initial_v must match shape \(nx,ny,nz,3\)"):
        run(B, None, None, *ARGS, 3, {"env_t": [1, 1, 1], "initial_v": np.zeros((6, 7, 8, 2))})


def test_jax_contract_errors():
    opts = {"env_t": np.ones(24, np.float32)}
    with pytest.raises(ValueError, match=r"multiple of snapshot_stride"):
        sw.shear_fdtd_jax(None, None, B, *ARGS, 24, opts, snapshot_stride=5)
    with pytest.raises(ValueError, match=r"positive integer"):
        sw.shear_fdtd_jax(None, None, B, *ARGS, 24, opts, snapshot_stride=0)
    with pytest.raises(ValueError, match=r"not both supported"):
        sw.shear_fdtd_jax(None, None, B, *ARGS, 24, opts, snapshot_stride=6, return_traces=True,
                          trace_indices=([0], [0], [0]))
    with pytest.raises(ValueError, match="trace_indices must be provided"):
        sw.shear_fdtd_jax(None, None, B, *ARGS, 24, opts, return_traces=True)
    with pytest.raises(ValueError, match="callback is not supported"):
        sw.shear_fdtd_jax(None, None, B, *ARGS, 24, opts, callback=lambda it, u: None)
    with pytest.raises(NotImplementedError, match="reproject_every"):
        sw.shear_fdtd_jax(None, None, B, *ARGS, 24, {**opts, "reproject_every": 5})
    with pytest.raises(ValueError, match="3-tuple"):
        sw.shear_fdtd_jax(None, None, B, *ARGS, 24, opts, trace_indices=([0], [0]), return_traces=True)
    with pytest.raises(ValueError, match="same shape"):
        sw.shear_fdtd_jax(None, None, B, *ARGS, 24, opts, trace_indices=([0], [0, 1], [0]), return_traces=True)
    with pytest.raises(ValueError, match="1D"):
        sw.shear_fdtd_jax(None, None, B, *ARGS, 24, opts, trace_indices=([[0]], [[0]], [[0]]), return_traces=True)


def test_strain_errors():
    z = np.zeros((5, 5, 5))
    with pytest.raises(ValueError, match="same shape"):
        sw.compute_strain_tensor(z, np.zeros((5, 5, 6)), z, 1e-3, 1e-3, 1e-3)
    with pytest.raises(ValueError, match="must be positive"):
        sw.compute_strain_tensor(z, z, z, -1e-3, 1e-3, 1e-3)
    with pytest.raises(ValueError, match="3-D"):
        sw.compute_strain_tensor(np.zeros((5, 5)), z, z, 1e-3, 1e-3, 1e-3)
    with pytest.raises(ValueError, match="Voigt"):
        sw.compute_strain_invariants(np.zeros((3, 3, 3, 3)))


def test_hann1d_matches_numpy():
    for n in (0, 1, 2, 7, 64):
        w = sw.hann1d(n)
        assert w.dtype == np.float32 and w.shape == (n,)
        if n > 1:
            np.testing.assert_allclose(w, np.hanning(n), rtol=1e-6)
