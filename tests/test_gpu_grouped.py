"""Grouped MFT execution (one launch per stage for many plans, sizes and batches) against the
same transforms run one by one.  Tolerances: FP64 1e-12 of the peak, split-FP16 5e-7 of the
peak (the individual calls may use other tile shapes, i.e. another summation order)."""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from hcipy_amd import _lib

from helpers import dev, host, peak_rel_field, say

pytestmark = pytest.mark.gpu


def _rand(shape, seed, dtype):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


@pytest.mark.parametrize('precision,dtype,tol', [(None, np.complex128, 1e-12), ('f16x2', np.complex64, 5e-7),
                                                 ('bf16x3', np.complex64, 5e-7),
                                                 ('fp32', np.complex64, 3e-6)])
def test_grouped_matches_individual(precision, dtype, tol):
    mfts, calls, refs = [], [], []
    k = 0
    for npix in (48, 64):
        gi = ha.make_pupil_grid(npix)
        for wl in (0.95, 1.0, 1.05):
            go = ha.make_focal_grid(4, 6).scaled(2 * np.pi / wl)
            mft = ha.MatrixFourierTransform(gi, go, precision=precision)
            for direction, batch in ((_lib.FORWARD, 3), (_lib.BACKWARD, 2)):
                n_in = gi.size if direction == _lib.FORWARD else go.size
                x = dev(_rand((batch, n_in), k, dtype))
                k += 1
                alpha = complex(0.3, -0.7 * wl)
                calls.append((mft, direction, x, alpha, _lib.OUTPUT_FIELD))
                ref = mft._execute_plain(x.contiguous(), direction, alpha)
                refs.append(host(ref))
            mfts.append(mft)
    # a 1-D transform in the same group
    g1 = ha.make_uniform_grid(50, 1.0)
    o1 = ha.make_uniform_grid(30, 40.0)
    m1 = ha.MatrixFourierTransform(g1, o1, precision=precision)
    x1 = dev(_rand((4, 50), 99, dtype))
    calls.append((m1, _lib.FORWARD, x1, 1.0, _lib.OUTPUT_FIELD))
    refs.append(host(m1._execute_plain(x1, _lib.FORWARD, 1.0)))
    group = ha.MftGroup(64)
    outs = group.execute(calls)
    say('grouped', precision, len(calls), 'items')
    for o, r in zip(outs, refs):
        assert peak_rel_field(host(o), r) < tol
    # a second call with the same group reuses its table and workspaces
    outs2 = group.execute(calls)
    for o, r in zip(outs2, refs):
        assert peak_rel_field(host(o), r) < tol


def test_grouped_broadband_intensity_and_precision_check():
    pupil = ha.make_pupil_grid(64)
    focal = ha.make_focal_grid(4, 8)
    prop = ha.FraunhoferPropagator(pupil, focal, precision='f16x2')
    E = _rand(pupil.size, 3, np.complex64)
    wfs = [ha.Wavefront(ha.Field(dev(E), pupil), wl) for wl in np.linspace(0.95, 1.05, 5)]
    I = prop.forward_multiple(wfs, output='intensity')
    for wf, Ig in zip(wfs, I):
        Iref = host(prop.forward_intensity(wf))
        assert np.abs(host(Ig) - Iref).max() <= 1e-6 * Iref.max()
    U = prop.forward_multiple(wfs)
    for wf, u in zip(wfs, U):
        assert peak_rel_field(host(u.electric_field), host(prop.forward(wf).electric_field)) < 5e-7
    # mixing precisions in one group is refused
    p64 = ha.FraunhoferPropagator(pupil, focal)
    m32 = prop.instance(1.0).fourier_transform
    m64 = p64.instance(1.0).fourier_transform
    with pytest.raises(ValueError):
        ha.MftGroup(4).execute([(m32, _lib.FORWARD, dev(E), 1.0, _lib.OUTPUT_FIELD),
                                (m64, _lib.FORWARD, dev(E.astype(np.complex128)), 1.0, _lib.OUTPUT_FIELD)])
