"""Pyramid wavefront-sensor optics and the tip-tilt mirror on the device against the oracle (GPU).

Tolerances: complex128 1e-12 of the output peak (two FP64 MFTs and FP64 phases); complex64 with
the split-FP16 MFT 2e-6 (two transforms each at about 1e-7 of their peak); the reference's
slow/fast modulation agreement 2e-2."""
import numpy as np
import pytest

import hcipy_amd as ha
from oracle import hcipy_oracle as o

from helpers import dev, host, og, peak_rel_field, say

pytestmark = pytest.mark.gpu


def _aberrated_pupil(pupil, seed=1, d=1.0):
    x, y = og(pupil).coords
    c = np.random.default_rng(seed).standard_normal(5)
    ab = 0.4 * (c[0] * x + c[1] * y + c[2] * (x * x + y * y) + c[3] * x * y + c[4] * (x * x - y * y))
    return o.evaluate_supersampled(lambda g: o.circular_aperture(g, d), og(pupil), 4) * np.exp(1j * ab)


@pytest.mark.parametrize('dtype,tol', [(np.complex128, 1e-12), (np.complex64, 2e-6)])
def test_pyramid_optics_match_oracle(dtype, tol):
    """PyramidWavefrontSensorOptics (pyramid.py:95-200): forward and backward against the oracle at
    64^2 -> 128^2 (q = 2, focal grid 128^2), separation 0.8, at a wavelength other than the
    reference one."""
    pupil = ha.make_pupil_grid(64, 1.0)
    wfs = ha.make_pupil_grid(128, 2.0)
    opt = ha.PyramidWavefrontSensorOptics(pupil, wfs, 0.8)
    geom = o.PyramidGeometry(og(pupil), og(wfs), 0.8)
    assert opt.focal_grid.dims.tolist() == list(geom.focal_grid.dims)
    E = _aberrated_pupil(pupil)
    wl = 1.1
    out = host(opt.forward(ha.Wavefront(ha.Field(dev(E.astype(dtype)), pupil), wl)).electric_field)
    ref = o.pyramid_wfs_forward(E, geom, wl)
    e = peak_rel_field(out, ref)
    say('pyramid %s forward %.2e' % (dtype.__name__, e))
    assert e < tol
    back = host(opt.backward(ha.Wavefront(ha.Field(dev(ref.astype(dtype)), wfs), wl)).electric_field)
    refb = o.pyramid_wfs_forward(ref, geom, wl, adjoint=True)
    assert peak_rel_field(back, refb) < tol


def test_pyramid_arguments():
    pupil = ha.make_pupil_grid(32, 1.0)
    wfs = ha.make_pupil_grid(64, 2.0)
    with pytest.raises(ValueError):
        ha.PyramidWavefrontSensorOptics(pupil, wfs, q=1.5)
    opt = ha.PyramidWavefrontSensorOptics(pupil, wfs, q=3, num_airy=8)
    assert opt.focal_grid.dims.tolist() == [48, 48]


@pytest.mark.parametrize('fast', [False, True])
def test_modulated_pyramid_matches_oracle(fast):
    """ModulatedPyramidWavefrontSensorOptics, slow (tilted inputs through the phase-ramp kernel)
    and fast (shifted surfaces as one phase operand per step), all steps in one batch, against
    the oracle per step; a batch of two input fields gives each field's own result."""
    pupil = ha.make_pupil_grid(64, 1.0)
    wfs = ha.make_pupil_grid(128, 2.0)
    pyr = ha.PyramidWavefrontSensorOptics(pupil, wfs, 1.0, q=3)
    mod = ha.ModulatedPyramidWavefrontSensorOptics(pyr, 1.5, num_steps=8, fast_modulation_method=fast)
    geom = o.PyramidGeometry(og(pupil), og(wfs), 1.0, q=3)
    E = _aberrated_pupil(pupil)
    outs = mod.forward(ha.Wavefront(ha.Field(dev(E), pupil), 0.9))
    refs = o.modulated_pyramid_wfs_forward(E, geom, 1.5, 8, fast=fast, wavelength=0.9)
    assert len(outs) == 8
    e = max(peak_rel_field(host(w.electric_field), r) for w, r in zip(outs, refs))
    say('modulated pyramid fast=%s: %.2e' % (fast, e))
    assert e < 1e-12
    E2 = np.stack([E, _aberrated_pupil(pupil, seed=2)])
    outs2 = mod.forward(ha.Wavefront(ha.Field(dev(E2), pupil), 0.9))
    refs2 = o.modulated_pyramid_wfs_forward(E2[1], geom, 1.5, 8, fast=fast, wavelength=0.9)
    for w, r, r0 in zip(outs2, refs2, refs):
        t = host(w.electric_field)
        assert t.shape == (2, wfs.size)
        assert peak_rel_field(t[0], r0) < 1e-12 and peak_rel_field(t[1], r) < 1e-12
    with pytest.raises(RuntimeError):
        mod.backward(outs[0])


def test_modulated_pyramid_reference_slow_fast_agreement():
    """tests/test_wavefront_sensing.py:29-52 on the device: 128^2 pupil, 256^2 sensor grid, q = 3,
    modulation 2, 12 steps; mean power of slow and fast within 2e-2 of the peak."""
    pupil = ha.make_pupil_grid(128)
    wfs = ha.make_pupil_grid(256, 2)
    pyr = ha.PyramidWavefrontSensorOptics(pupil, wfs, 1.0, q=3)
    slow = ha.ModulatedPyramidWavefrontSensorOptics(pyr, 2, num_steps=12)
    fast = ha.ModulatedPyramidWavefrontSensorOptics(pyr, 2, num_steps=12, fast_modulation_method=True)
    wf = ha.Wavefront(ha.Field(dev(_aberrated_pupil(pupil)), pupil))
    ts = sum(host(w.power) for w in slow(wf)) / 12
    tf = sum(host(w.power) for w in fast(wf)) / 12
    e = np.abs(ts - tf).max() / ts.max()
    say('slow/fast modulation %.2e' % e)
    assert e < 2e-2


def test_tip_tilt_mirror():
    """TipTiltMirror (tip_tilt_mirror.py:6-23): exp(2 i k (a0 x + a1 y)) forward, its conjugate
    backward, the surface a0 x + a1 y."""
    g = ha.make_pupil_grid(48, 1.2)
    ttm = ha.TipTiltMirror(g)
    ttm.actuators = [3e-7, -1.2e-7]
    rng = np.random.default_rng(4)
    E = rng.standard_normal(g.size) + 1j * rng.standard_normal(g.size)
    wl = 7e-7
    x, y = og(g).coords
    ph = 2 * 2 * np.pi / wl * (3e-7 * x - 1.2e-7 * y)
    out = host(ttm.forward(ha.Wavefront(ha.Field(dev(E), g), wl)).electric_field)
    assert np.abs(out - E * np.exp(1j * ph)).max() < 1e-12 * np.abs(E).max()
    back = host(ttm.backward(ha.Wavefront(ha.Field(dev(out), g), wl)).electric_field)
    assert np.abs(back - E).max() < 1e-12 * np.abs(E).max()
    assert np.abs(host(ttm.surface) - (3e-7 * x - 1.2e-7 * y)).max() < 1e-20
