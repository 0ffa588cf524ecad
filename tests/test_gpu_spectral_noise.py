"""FFT spectral noise on the device against the oracle (GPU).

The white noise is drawn on the host with numpy's default_rng(seed) in both, so a realization is
the same set of Fourier coefficients; the device then forms C times the noise, the FP64 phase
ramps of the shifts and the FP64 inverse transform.  Tolerance 1e-12 of the screen's peak."""
import numpy as np
import pytest

import hcipy_amd as ha
from hcipy_amd import pointwise
from oracle import hcipy_oracle as o

from helpers import dev, host, og, say

pytestmark = pytest.mark.gpu


def _psd(exponent):
    f = o.power_law_psd(exponent)
    return lambda grid: f(og(grid))


@pytest.mark.parametrize('dims,extent,oversample', [([32, 32], [2.0, 2.0], 1), ([24, 16], [2.4, 1.6], 2),
                                                    ([40], [4.0], 1)])
def test_spectral_noise_matches_oracle(dims, extent, oversample):
    g = ha.make_uniform_grid(dims, extent)
    fac = ha.SpectralNoiseFactoryFFT(_psd(-2.5), g, oversample=oversample)
    noise = fac.make_random(7)
    scr = host(noise())
    ref = o.spectral_noise_fft(o.power_law_psd(-2.5), og(g), 7, oversample=oversample)
    scale = np.abs(ref).max()
    say('spectral noise', dims, oversample, 'error', np.abs(scr - ref).max() / scale)
    assert np.abs(scr - ref).max() < 1e-12 * scale
    shift = (0.37 * extent[0], -0.21 * extent[-1]) if len(dims) == 2 else (0.37 * extent[0],)
    sh = host(noise.shifted(shift)())
    refs = o.spectral_noise_fft(o.power_law_psd(-2.5), og(g), 7, oversample=oversample, shift=shift)
    assert np.abs(sh - refs).max() < 1e-12 * scale
    # shifted() leaves the realization alone; shift() moves it in place
    assert np.abs(host(noise()) - ref).max() < 1e-12 * scale
    noise.shift(shift)
    assert np.abs(host(noise()) - refs).max() < 1e-12 * scale


def test_phase_ramp_kernel():
    g = ha.make_uniform_grid([9, 7], [1.8, 0.7], center=[0.3, -0.1])
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, g.size)) + 1j * rng.standard_normal((3, g.size))
    gx, gy = g.coords
    for dt, tol in ((np.complex128, 1e-15), (np.complex64, 1e-6)):
        out = host(pointwise.phase_ramp(dev(x.astype(dt)), g, 2.5, -1.25))
        ref = x.astype(dt) * np.exp(1j * (2.5 * gx - 1.25 * gy))
        assert np.abs(out - ref).max() < tol * 4


def test_power_law_error_matches_oracle_and_reference_ptv():
    """tests/test_optics.py:823-838 on the device: peak-to-valley 0.1 over the 2 D aperture."""
    g = ha.make_pupil_grid(64, 2.0)
    large = host(ha.evaluate_supersampled(ha.make_circular_aperture(2.0), g, 4))
    for exponent in (-2.5, -1.0):
        scr = host(ha.make_power_law_error(g, 0.1, 2.0, exponent=exponent, seed=11))
        ref = o.power_law_error(og(g), 0.1, 2.0, exponent=exponent, seed=11)
        assert np.abs(scr - ref).max() < 1e-12 * 0.1
        assert abs((np.ptp(scr[large > 0]) - 0.1) / 0.1) < 1e-10


def _smooth(grid):
    return (np.hypot(*grid.coords) ** 2 + 100.0) ** -1.0


@pytest.mark.parametrize('dims,extent,ovs,kind', [([32, 32], [2.0, 2.0], 4, 'power'), ([24, 16], [2.4, 1.6], 3, 'smooth')])
def test_spectral_noise_multiscale_matches_oracle(dims, extent, ovs, kind):
    """SpectralNoiseFactoryMultiscale (spectral_noise.py:203-330): the FFT part and the MFT part
    on the 1/ovs finer grid, both backward transforms in FP64 on the device, against the oracle
    with the same seed; shifted() and shift() in place.  1e-12 of the screen peak."""
    g = ha.make_uniform_grid(dims, extent)
    if kind == 'power':
        psd_dev, psd_ref = _psd(-2.5), o.power_law_psd(-2.5)
    else:
        psd_dev, psd_ref = _smooth, _smooth
    fac = ha.SpectralNoiseFactoryMultiscale(psd_dev, g, ovs)
    noise = fac.make_random(21)
    scr = host(noise())
    ref = o.spectral_noise_multiscale(psd_ref, og(g), 21, ovs)
    scale = np.abs(ref).max()
    say('multiscale noise', dims, ovs, kind, 'error', np.abs(scr - ref).max() / scale)
    assert np.abs(scr - ref).max() < 1e-12 * scale
    shift = (0.37 * extent[0], -0.21 * extent[1])
    refs = o.spectral_noise_multiscale(psd_ref, og(g), 21, ovs, shift=shift)
    assert np.abs(host(noise.shifted(shift)()) - refs).max() < 1e-12 * scale
    assert np.abs(host(noise()) - ref).max() < 1e-12 * scale
    noise.shift(shift)
    assert np.abs(host(noise()) - refs).max() < 1e-12 * scale
