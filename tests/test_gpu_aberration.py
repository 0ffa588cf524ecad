"""Surface apodizers and surface aberrations on the device against the oracle (GPU).

Tolerances: 1e-12 of the field or screen peak (FP64 phases and FP64 transforms), plus 1e-15 of
the largest phase in radians where the phase reaches 1e5 rad; round trips of unit-modulus
elements to 1e-12."""
import numpy as np
import pytest

import hcipy_amd as ha
from oracle import hcipy_oracle as o

from helpers import dev, host, og, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('n', [64, 256])
def test_high_pass_power_law_error_matches_oracle(n):
    """make_high_pass_power_law_error (device FourierFilter, q = 2) against the oracle with the
    same seed, with the default aperture and with a given one; the reference's flat-spectrum check
    (tests/test_optics.py:844-850) on the device screen."""
    D = 1.0
    g = ha.make_pupil_grid(n, 2.0)
    cutoff = 2 * np.pi * 10 / D
    ap = o.evaluate_supersampled(lambda gg: o.circular_aperture(gg, D), og(g), 4)
    fft = o.FFT(og(g))
    fmask = o.circular_aperture(fft.output_grid, 2 * cutoff)
    for aperture in (None, ap):
        scr = host(ha.make_high_pass_power_law_error(g, 0.1, 2 * D, cutoff, aperture=None if aperture is None
                                                     else ha.Field(dev(aperture), g), seed=11))
        ref = o.high_pass_power_law_error(og(g), 0.1, 2 * D, cutoff, aperture=aperture, seed=11)
        e = np.abs(scr - ref).max() / np.abs(ref).max()
        say('high-pass power-law error n=%d aperture=%s: %.2e' % (n, aperture is not None, e))
        assert e < 1e-12
        fk = np.abs(fft.forward(scr.astype(complex))) ** 2
        assert np.ptp(fk[fmask > 0]) < 1e-6


def test_surface_apodizer_matches_closed_form_and_round_trips():
    """E exp(i (n - 1) k sag) with scalar and wavelength-dependent n (apodization.py:64-102), and
    the reference's round trips (tests/test_optics.py:1114-1131)."""
    g = ha.make_pupil_grid(64)
    rng = np.random.default_rng(2)
    sag = rng.standard_normal(g.size) * 1e-7
    E = rng.standard_normal(g.size) + 1j * rng.standard_normal(g.size)
    wl = 1.3e-6
    for n_of in (1.5, lambda w: 1.5 + (w - 1e-6) * 1e4):
        sa = ha.SurfaceApodizer(ha.Field(dev(sag), g), n_of)
        n = n_of(wl) if callable(n_of) else n_of
        out = host(sa.forward(ha.Wavefront(ha.Field(dev(E), g), wl)).electric_field)
        ref = E * np.exp(1j * (n - 1) * sag * 2 * np.pi / wl)
        assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()
        wf = ha.Wavefront(ha.Field(dev(np.ones(g.size, complex)), g), 1e-6)
        back = host(sa.backward(sa.forward(wf)).electric_field)
        assert np.abs(back - 1).max() < 1e-12


def test_complex_surface_apodizer():
    """a E exp(i (n - 1) k s) forward, conj(a) E exp(-i (n - 1) k s) backward
    (apodization.py:104-150), with callable amplitude and index; unit-modulus round trips as in
    tests/test_optics.py:1134-1149."""
    g = ha.make_pupil_grid(64)
    rng = np.random.default_rng(3)
    surf = 0.1 * np.ones(g.size)
    E = rng.standard_normal(g.size) + 1j * rng.standard_normal(g.size)
    amp = np.exp(1j * rng.standard_normal(g.size)) * (0.5 + rng.random(g.size))
    csa = ha.ComplexSurfaceApodizer(ha.Field(dev(amp), g), ha.Field(dev(surf), g), 1.5)
    wl = 1e-6
    # the phase (n - 1) k s is 3.1e5 rad here: one ulp of it is 6e-11 rad, and the device and
    # numpy round the product in different orders, so the bound grows with the phase
    phase = 0.5 * surf * 2 * np.pi / wl
    tol = 1e-12 + 1e-15 * np.abs(phase).max()
    out = host(csa.forward(ha.Wavefront(ha.Field(dev(E), g), wl)).electric_field)
    ref = E * amp * np.exp(1j * phase)
    assert np.abs(out - ref).max() < tol * np.abs(ref).max()
    outb = host(csa.backward(ha.Wavefront(ha.Field(dev(E), g), wl)).electric_field)
    refb = E * np.conj(amp) * np.exp(-1j * phase)
    assert np.abs(outb - refb).max() < tol * np.abs(refb).max()
    csa2 = ha.ComplexSurfaceApodizer(lambda w: np.exp(1j * (w / 1e-6)), ha.Field(dev(surf), g),
                                     lambda w: 1.5 + (w - 1e-6) * 1e4)
    wf = ha.Wavefront(ha.Field(dev(np.ones(g.size, complex)), g), 1.2e-6)
    back = host(csa2.backward(csa2.forward(wf)).electric_field)
    assert np.abs(back - 1).max() < 1e-12


def test_surface_aberration_at_distance_matches_oracle():
    """SurfaceAberrationAtDistance (aberration.py:123-183): Fresnel to the surface, the mirror
    surface (n = -1: exp(-2 i k s)), Fresnel back; forward and the reference's backward (the
    surface's backward between the same two propagations), against the oracle's Fresnel
    filter.  The surface alone is SurfaceAberration with the same seed."""
    n, D, wl, z = 128, 0.01, 600e-9, 0.2
    g = ha.make_pupil_grid(n, 1.5 * D)
    surf = ha.SurfaceAberration(g, 50e-9, D, seed=6)
    s = o.power_law_error(og(g), 50e-9, D, seed=6)
    assert np.abs(host(surf.surface_sag) - s).max() < 1e-12 * np.abs(s).max()
    el = ha.SurfaceAberrationAtDistance(g, surf, z)
    E = o.evaluate_supersampled(lambda gg: o.circular_aperture(gg, D), og(g), 4).astype(complex)
    k = 2 * np.pi / wl
    for adjoint in (False, True):
        wf = ha.Wavefront(ha.Field(dev(E), g), wl)
        out = host((el.backward(wf) if adjoint else el.forward(wf)).electric_field)
        a = o.fresnel_forward(E, og(g), wl, z)
        a = a * np.exp((2j if adjoint else -2j) * k * s)
        ref = o.fresnel_forward(a, og(g), wl, z, adjoint=True)
        e = np.abs(out - ref).max() / np.abs(ref).max()
        say('surface aberration at %.2f m, adjoint=%s: %.2e' % (z, adjoint, e))
        assert e < 1e-12
