"""GPU parity of the acoustic strain statistics (acoustic_strain.py) against the oracle.

Gates: the volumetric strain, the gradient, the peak and the float64 gradient
statistics are bit-identical to the oracle in float32 and float64 (same
operations, contraction off, the gradient statistics summed in sample order).
rms and mean use float64 sums where NumPy sums in the storage type: they are
checked against the oracle run on float64 copies of the same inputs, to 1e-6
(float32) or 1e-12 (float64) of the rms scale.
"""

import numpy as np
import pytest

from oracle import acoustic_strain as A

pytestmark = pytest.mark.gpu

H = (0.3e-3, 0.25e-3, 0.35e-3)


def record(shape, nt, seed, dtype, burst=True):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(shape + (nt,)) * 1e5
    if burst:
        t = np.arange(nt)
        p *= np.where((t >= nt // 5) & (t < nt - nt // 4), 1.0, 0.05)
    c = rng.uniform(1450, 1600, shape)
    rho = rng.uniform(950, 1100, shape)
    return p.astype(dtype), c.astype(dtype), rho.astype(dtype)


GATES = [None, slice(3, 17), np.array([5, 5, 2, 11]), slice(None, None, 4)]
# contiguous gates on 16-byte boundaries (start and length multiples of four samples) take the kernels that
# read (and, for the volumetric strain and gradient, write) 16 bytes per lane
ALIGNED_GATES = [slice(8, 40), slice(0, 40), slice(4, 28)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("gate", GATES + ALIGNED_GATES)
def test_volumetric_and_gradient_bit_exact(sw, dtype, gate):
    p, c, rho = record((9, 8, 11), 24, 1, dtype)
    eps = sw.compute_acoustic_strain_volumetric(p, c, rho, gate)
    ref = A.compute_acoustic_strain_volumetric(p, c, rho, gate)
    assert eps.dtype == ref.dtype and eps.shape == ref.shape
    np.testing.assert_array_equal(eps, ref)
    got = sw.compute_acoustic_strain_gradient(p, c, rho, *H, gate)
    want = A.compute_acoustic_strain_gradient(p, c, rho, *H, gate)
    for g, w in zip(got, want):
        assert g.dtype == w.dtype and g.shape == w.shape
        np.testing.assert_array_equal(g, w)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("gate", GATES)
def test_gradient_statistics_bit_exact(sw, dtype, gate):
    p, c, rho = record((10, 7, 12), 30, 2, dtype)
    got = sw.compute_acoustic_strain_gradient_peak(p, c, rho, *H, gate)
    want = A.compute_acoustic_strain_gradient_peak(p, c, rho, *H, gate)
    assert set(got) == set(want)
    for k in want:
        assert got[k].dtype == np.float64
        np.testing.assert_array_equal(got[k], want[k])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("gate", ALIGNED_GATES + [slice(8, 44)])
def test_gradient_statistics_aligned_gates(sw, dtype, gate):
    """Contiguous gates on 16-byte boundaries (the staging reads each voxel's window of samples in 16-byte
    runs; slice(8, 44) ends in a partial window), on a grid with partial tiles in y and z."""
    p, c, rho = record((10, 13, 37), 48, 3, dtype)
    got = sw.compute_acoustic_strain_gradient_peak(p, c, rho, *H, gate)
    want = A.compute_acoustic_strain_gradient_peak(p, c, rho, *H, gate)
    for k in want:
        np.testing.assert_array_equal(got[k], want[k])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("gate", [slice(None), slice(8, 40)])
def test_gradient_statistics_propagate_nan(sw, dtype, gate):
    """np.maximum propagates NaN: a NaN sample, an infinite one (inf - inf in a gradient) and a zero bulk
    modulus (0/0 and p/0 strains) give NaN and inf statistics exactly where NumPy does."""
    p, c, rho = record((9, 10, 12), 48, 4, dtype)
    p[3, 4, 5, 20] = np.nan
    p[6, 2, 7, 30] = np.inf
    p[6, 3, 7, 30] = np.inf
    c[2, 7, 3] = 0.0
    p[2, 7, 3, :10] = 0.0
    with np.errstate(all="ignore"):
        got = sw.compute_acoustic_strain_gradient_peak(p, c, rho, *H, gate)
        want = A.compute_acoustic_strain_gradient_peak(p, c, rho, *H, gate)
    for k in want:
        assert np.isnan(want[k]).any()
        np.testing.assert_array_equal(got[k], want[k])


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-6), (np.float64, 1e-12)])
@pytest.mark.parametrize("gate", GATES + ALIGNED_GATES)
def test_peak_rms_mean(sw, dtype, tol, gate):
    p, c, rho = record((12, 9, 10), 40, 3, dtype)
    got = sw.compute_acoustic_strain_peak(p, c, rho, gate)
    ref = A.compute_acoustic_strain_peak(p, c, rho, gate)
    np.testing.assert_array_equal(got["peak"], ref["peak"])
    gate64 = A.resolve_gate(p, gate)  # the gate the storage-type record selects
    ref64 = A.compute_acoustic_strain_peak(p.astype(np.float64), c.astype(np.float64), rho.astype(np.float64), gate64)
    scale = np.abs(ref64["rms"]).max()
    for k in ("rms", "mean"):
        assert got[k].dtype == dtype
        assert np.abs(got[k].astype(np.float64) - ref64[k]).max() <= tol * scale, k


def test_device_tensors_in_and_out(sw):
    import torch

    p, c, rho = record((6, 6, 6), 16, 4, np.float32)
    tp, tc, tr = (torch.from_numpy(a).cuda() for a in (p, c, rho))
    eps = sw.compute_acoustic_strain_volumetric(tp, tc, tr, slice(2, 9))
    assert isinstance(eps, torch.Tensor) and eps.is_cuda
    np.testing.assert_array_equal(eps.cpu().numpy(), A.compute_acoustic_strain_volumetric(p, c, rho, slice(2, 9)))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gradient_statistics_bit_exact_large(sw, dtype):
    """Millions of quotients: the kernel divides by K and 2h through a correction step from the reciprocal."""
    p, c, rho = record((40, 36, 44), 64, 9, dtype, burst=False)
    got = sw.compute_acoustic_strain_gradient_peak(p, c, rho, *H, slice(None))
    want = A.compute_acoustic_strain_gradient_peak(p, c, rho, *H, slice(None))
    for k in want:
        np.testing.assert_array_equal(got[k], want[k])


def test_example_call_pattern(sw):
    """As tips_transcranial_shear_3d.py:1136-1160 calls it: a transposed record view, gate=slice(None),
    spacings positional, and the returned dicts' arrays masked in place."""
    rng = np.random.default_rng(13)
    shape, nt = (9, 10, 8), 20
    p = (rng.standard_normal((nt,) + shape) * 1e5).astype(np.float32).transpose(1, 2, 3, 0)
    _, c, rho = record(shape, 2, 14, np.float32)
    st = sw.compute_acoustic_strain_peak(p, c, rho, gate=slice(None))
    np.testing.assert_array_equal(st["peak"], A.compute_acoustic_strain_peak(p, c, rho, gate=slice(None))["peak"])
    mask = np.zeros(shape, bool)
    mask[0] = True
    for key in ("peak", "rms", "mean"):
        st[key][mask] = 0.0
    gp = sw.compute_acoustic_strain_gradient_peak(p, c, rho, *H, gate=slice(None))
    want = A.compute_acoustic_strain_gradient_peak(p, c, rho, *H, gate=slice(None))
    for key in want:
        np.testing.assert_array_equal(gp[key], want[key])
        gp[key][mask] = 0.0


def test_empty_gate_as_numpy(sw):
    """An empty gate: empty sample axes, zero gradient statistics, and NumPy's error for the peak."""
    p, c, rho = record((6, 7, 8), 12, 15, np.float32)
    gate = np.array([], dtype=np.int64)
    eps = sw.compute_acoustic_strain_volumetric(p, c, rho, gate)
    assert eps.shape == (6, 7, 8, 0)
    for g in sw.compute_acoustic_strain_gradient(p, c, rho, *H, gate):
        assert g.shape == (6, 7, 8, 0)
    got = sw.compute_acoustic_strain_gradient_peak(p, c, rho, *H, gate)
    want = A.compute_acoustic_strain_gradient_peak(p, c, rho, *H, gate)
    for k in want:
        np.testing.assert_array_equal(got[k], want[k])
    with pytest.raises(ValueError):
        A.compute_acoustic_strain_peak(p, c, rho, gate)
    with pytest.raises(ValueError, match="zero-size array"):
        sw.compute_acoustic_strain_peak(p, c, rho, gate)
