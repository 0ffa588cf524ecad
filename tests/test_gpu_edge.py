"""Edge cases of the device path: empty batches, more fields than one launch's grid.z holds,
non-square and odd grids, zero and non-finite inputs (GPU)."""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from hcipy_amd import pointwise
from oracle import hcipy_oracle as o

from helpers import dev, host, og, peak_rel_field, say

pytestmark = pytest.mark.gpu


def _rand(shape, seed, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


@pytest.mark.parametrize('precision,dtype', [(None, np.complex128), ('f16x2', np.complex64), ('fp32', np.complex64)])
def test_empty_batch(precision, dtype):
    pupil = ha.make_pupil_grid(16)
    focal = ha.make_focal_grid(2, 4)
    mft = ha.MatrixFourierTransform(pupil, focal, precision=precision)
    x = torch.zeros((0, pupil.size), dtype=torch.complex128 if dtype == np.complex128 else torch.complex64, device='cuda')
    y = mft.forward(ha.Field(x, pupil))
    assert tuple(y.shape) == (0, focal.size)
    z = mft.backward(ha.Field(torch.zeros((0, focal.size), dtype=x.dtype, device='cuda'), focal))
    assert tuple(z.shape) == (0, pupil.size)
    filt = ha.FourierFilter(pupil, lambda g: ha.Field(torch.ones(g.size, dtype=torch.complex128, device='cuda'), g), 2)
    assert tuple(filt.forward(ha.Field(x, pupil)).shape) == (0, pupil.size)
    assert tuple(pointwise.multiply(ha.Field(x, pupil), dev(np.ones(pupil.size))).shape) == (0, pupil.size)
    assert tuple(pointwise.intensity(ha.Field(x, pupil)).shape) == (0, pupil.size)


def test_batch_beyond_one_launch():
    """70 000 fields: more than the 65 535 slices of one launch's grid.y / grid.z, so the stage-2,
    filter, matrix-filter, pointwise and deformable-mirror launches run in chunks."""
    pupil = ha.make_pupil_grid(8)
    focal = ha.make_focal_grid(2, 2)
    n = 70000
    x = _rand((n, pupil.size), 1)
    mft = ha.MatrixFourierTransform(pupil, focal)
    y = host(mft.forward(ha.Field(dev(x), pupil)))
    sel = [0, 1, 65534, 65535, 65536, n - 1]
    ref = o.mft_forward(x[sel], og(pupil), og(focal))
    say('batch %d: max error of sampled fields %.2e' % (n, peak_rel_field(y[sel], ref)))
    assert peak_rel_field(y[sel], ref) < 1e-12
    filt = ha.FourierFilter(pupil, lambda g: ha.Field(dev(_rand(g.size, 2)), g), 2)
    f = host(filt.forward(ha.Field(dev(x), pupil)))
    tf = _rand(o.make_fft_grid(og(pupil), 2).size, 2)
    ref = o.fourier_filter(x[sel], og(pupil), tf, q=2)
    assert peak_rel_field(f[sel], ref) < 1e-12
    # a 2 x 2 matrix transfer function over the same 70 000 fields per component: the batch is
    # split into calls of 65 535 and the row passes of each call into launches of 65 535 fields
    xm = np.stack([x, _rand((n, pupil.size), 5)])
    tfm = _rand((2, 2, o.make_fft_grid(og(pupil), 2).size), 6)
    filt_m = ha.FourierFilter(pupil, dev(tfm), 2)
    fm = host(filt_m.forward(ha.Field(dev(xm), pupil)))
    refm = o.fourier_filter(xm[:, sel], og(pupil), tfm, q=2)
    assert peak_rel_field(fm[:, sel], refm) < 1e-12
    fb = host(filt_m.backward(ha.Field(dev(xm), pupil)))
    refb = o.fourier_filter(xm[:, sel], og(pupil), tfm, q=2, adjoint=True)
    assert peak_rel_field(fb[:, sel], refb) < 1e-12
    # pointwise kernels: broadcast and per-field operands, intensity, FP64 power
    a = _rand(pupil.size, 3)
    m = host(pointwise.multiply(ha.Field(dev(x), pupil), dev(a)))
    assert np.abs(m[sel] - x[sel] * a).max() < 1e-13
    per = _rand((n, pupil.size), 4)
    m = host(pointwise.sub_mul(ha.Field(dev(x), pupil), None, dev(per), conjugate=True))
    assert np.abs(m[sel] - x[sel] * np.conj(per[sel])).max() < 1e-13
    p = host(pointwise.power_sum(dev(x)))
    np.testing.assert_allclose(p[sel], (np.abs(x[sel]) ** 2).sum(axis=-1), rtol=1e-13)
    # deformable-mirror surfaces of more pokes than one launch holds
    IF = ha.make_gaussian_influence_functions(pupil, 4, 0.25)
    acts = np.random.default_rng(5).standard_normal((n, 16))
    surf = host(IF.linear_combination(dev(acts)))
    ref = o.dm_surface(o.gaussian_influence_functions(og(pupil), 4, 0.25), acts[sel])
    assert np.abs(surf[sel] - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize('dtype,tol', [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_non_square_odd_grids_and_tensor_fields(dtype, tol):
    pupil = ha.make_uniform_grid([7, 12], [0.7, 1.2])
    focal = ha.make_uniform_grid([9, 5], [18.0, 10.0])
    x = _rand((2, 3, pupil.size), 3, dtype)
    mft = ha.MatrixFourierTransform(pupil, focal)
    y = host(mft.forward(ha.Field(dev(x), pupil)))
    assert y.shape == (2, 3, focal.size)
    ref = o.mft_forward(x.astype(np.complex128), og(pupil), og(focal))
    assert peak_rel_field(y, ref) < tol
    b = host(mft.backward(ha.Field(dev(y.astype(dtype)), focal)))
    refb = o.mft_backward(y.astype(np.complex128), og(pupil), og(focal))
    assert peak_rel_field(b, refb) < tol


def test_zero_and_non_finite_inputs():
    pupil = ha.make_pupil_grid(32)
    focal = ha.make_focal_grid(4, 4)
    for precision, dt in ((None, torch.complex128), ('f16x2', torch.complex64), ('fp32', torch.complex64)):
        mft = ha.MatrixFourierTransform(pupil, focal, precision=precision)
        zero = torch.zeros(pupil.size, dtype=dt, device='cuda')
        assert host(mft.forward(ha.Field(zero, pupil))).tolist() == [0] * focal.size
        bad = zero.clone()
        bad[5] = float('nan')
        out = host(mft.forward(ha.Field(bad, pupil)))
        # a NaN sample reaches every output sample (each output sums over every input)
        assert np.isnan(out).all(), precision
    x = torch.zeros(pupil.size, dtype=torch.complex64, device='cuda')
    x[3] = float('inf')
    assert np.isinf(host(pointwise.intensity(ha.Field(x, pupil)))[3])
