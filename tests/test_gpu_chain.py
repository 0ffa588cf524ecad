"""Pointwise kernels and the Lyot chain against the oracle (GPU)."""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from hcipy_amd import pointwise, workloads
from oracle import hcipy_oracle as o

from helpers import dev, host, og, peak_rel_field, say

pytestmark = pytest.mark.gpu


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
def test_sub_mul_and_conj(dtype):
    x = _rand((3, 1000), 0).astype(dtype)
    y = _rand((3, 1000), 1).astype(dtype)
    a = _rand(1000, 2)
    r = host(pointwise.sub_mul(dev(x), dev(y), dev(a), conjugate=False))
    tol = 1e-6 if dtype == np.complex64 else 1e-14
    assert np.abs(r - (x - y) * a).max() < tol * 10
    r = host(pointwise.multiply(dev(x), dev(a.real), conjugate=True))
    assert np.abs(r - x * a.real).max() < tol * 10
    r = host(pointwise.multiply(dev(x), dev(a), conjugate=True))
    assert np.abs(r - x * np.conj(a)).max() < tol * 10
    # per-field operand (stride n)
    r = host(pointwise.multiply(dev(x), dev(y)))
    assert np.abs(r - x * y).max() < tol * 10


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
def test_phase_intensity_power(dtype):
    x = _rand((2, 4097), 3).astype(dtype)
    phi = np.random.default_rng(4).standard_normal(4097) * 50
    r = host(pointwise.multiply_phase(dev(x), dev(phi), -2.0))
    tol = 1e-5 if dtype == np.complex64 else 1e-13
    assert np.abs(r - x * np.exp(-2j * phi)).max() < tol
    I = host(pointwise.intensity(dev(x)))
    assert np.allclose(I, np.abs(x) ** 2, rtol=tol)
    mask = (np.arange(4097) % 3 == 0).astype(float)
    p = host(pointwise.power_sum(dev(x), 0.5, mask=dev(mask)))
    ref = 0.5 * np.sum(np.abs(x.astype(np.complex128)) ** 2 * mask, axis=-1)
    assert np.allclose(p, ref, rtol=1e-12)


def test_wavefront_total_power():
    g = ha.make_pupil_grid(64, 2.0)
    E = _rand(g.size, 5)
    wf = ha.Wavefront(ha.Field(dev(E), g), 1.0)
    assert wf.total_power == pytest.approx(np.sum(np.abs(E) ** 2) * g.weights, rel=1e-12)
    wf.total_power = 3.0
    assert wf.total_power == pytest.approx(3.0, rel=1e-12)


def test_lyot_chain_small_vs_oracle():
    """LyotCoronagraph.forward and .backward, complex128, against the oracle restatement
    (hcipy/coronagraphy/lyot.py:66-122); suppression bar of tests/test_coronagraphy.py:272-304."""
    pupil = ha.make_pupil_grid(128, 1.1)
    ap = ha.evaluate_supersampled(ha.make_circular_aperture(1.0), pupil, 8)
    stop = ha.evaluate_supersampled(ha.make_circular_aperture(0.95), pupil, 8)
    fpm_grid = ha.make_focal_grid(32, 3)
    fpm = 1 - ha.evaluate_supersampled(ha.make_circular_aperture(5), fpm_grid, 8)
    cor = ha.LyotCoronagraph(pupil, fpm, stop)
    wf = ha.Wavefront(ap, 1.0)
    wf.total_power = 1
    L = host(cor(wf).electric_field)
    E = host(wf.electric_field)
    Lr = o.lyot_forward(E, og(pupil), og(fpm_grid), host(fpm), host(stop))
    assert peak_rel_field(L, Lr) < 1e-12
    B = host(cor.backward(ha.Wavefront(ha.Field(dev(Lr), pupil), 1.0)).electric_field)
    Br = o.lyot_backward(Lr, og(pupil), og(fpm_grid), host(fpm), host(stop))
    assert peak_rel_field(B, Br) < 1e-12
    focal = ha.make_focal_grid(3, 25)
    prop = ha.FraunhoferPropagator(pupil, focal)
    norm = host(prop(wf).power).max()
    assert host(prop(cor(wf)).power).max() / norm < 5e-3


@pytest.mark.parametrize('precision,tol', [('f16x2', 1e-9), ('bf16x3', 1e-9), ('fp32', 1e-7)])
def test_lyot_1024_chain_contrast_error(precision, tol):
    """Full 1024^2 Lyot chain (256^2 mask plane, 512^2 final plane) in complex64: intensity
    error as a fraction of the unocculted peak, against the FP64 oracle on the same input."""
    pupil, fpm_grid, focal, E, fpm, stop, ap = workloads.lyot_1024()
    E64 = E.astype(np.complex64)
    cor = ha.LyotCoronagraph(pupil, ha.Field(dev(fpm), fpm_grid), ha.Field(dev(stop), pupil), precision=precision)
    prop = ha.FraunhoferPropagator(pupil, focal, precision=precision)
    wf = ha.Wavefront(ha.Field(dev(E64), pupil), 1.0)
    img = host(prop(cor(wf)).electric_field)
    Ec = E64.astype(np.complex128)
    L = o.lyot_forward(Ec, og(pupil), og(fpm_grid), fpm, stop)
    ref = o.fraunhofer_forward(L, og(pupil), og(focal))
    peak = np.max(np.abs(o.fraunhofer_forward(Ec, og(pupil), og(focal))) ** 2)
    err = np.abs(np.abs(img) ** 2 - np.abs(ref) ** 2).max() / peak
    say('lyot 1024 %s: raw contrast %.3e, error / unocculted peak %.3e' % (precision, (np.abs(ref) ** 2).max() / peak, err))
    assert err < tol


