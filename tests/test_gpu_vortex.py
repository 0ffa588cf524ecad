"""Multi-scale (vortex) coronagraph against the oracle, the reference's rejection thresholds
(tests/test_coronagraphy.py:4-28) and its own adjoint (GPU).

Tolerances: complex128 runs the FP64 MFMA transforms and FP64 FFTs, so the Lyot-plane field
matches the oracle to 1e-10 of the input peak; complex64 (split FP16 transforms, FP32 FFTs)
to 2e-6, the FP32 rounding of the level-0 FFT filter."""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from oracle import hcipy_oracle as o

from helpers import dev, host, og, peak_rel_field, say

pytestmark = pytest.mark.gpu


def _setup(n=256):
    pupil = ha.make_pupil_grid(n)
    focal = ha.make_focal_grid(4, 32)
    ap = host(ha.evaluate_supersampled(ha.make_circular_aperture(1), pupil, 8))
    lyot = host(ha.evaluate_supersampled(ha.make_circular_aperture(0.99), pupil, 8)) > 1 - 1e-5
    E = ap.astype(complex)
    E = E / np.sqrt(np.sum(np.abs(E) ** 2) * pupil.weights)
    return pupil, focal, E, lyot


@pytest.mark.parametrize('dtype,tol', [(np.complex128, 1e-10), (np.complex64, 2e-6)])
def test_vortex_matches_oracle(dtype, tol):
    pupil, focal, E, lyot = _setup()
    grids, masks = o.multiscale_masks(og(pupil), o.vortex_mask(2))
    ref = o.multiscale_forward(E, og(pupil), grids, masks)
    vc = ha.VortexCoronagraph(pupil, 2)
    assert [tuple(g.dims) for g in vc.focal_grids] == [g.dims for g in grids]
    for m, r in zip(vc.focal_masks, masks):
        assert peak_rel_field(host(m), r) < 1e-12
    wf = ha.Wavefront(ha.Field(dev(E.astype(dtype)), pupil))
    L = host(vc.forward(wf).electric_field)
    err = peak_rel_field(L, ref)
    say('vortex', dtype.__name__, 'Lyot plane error vs oracle', err)
    assert err < tol
    # the backward direction against the oracle's conj(M) chain
    refb = o.multiscale_forward(E, og(pupil), grids, masks, adjoint=True)
    Lb = host(vc.backward(wf).electric_field)
    assert peak_rel_field(Lb, refb) < tol


@pytest.mark.parametrize('dtype', [np.complex128, np.complex64])
def test_vortex_rejection_reference_thresholds(dtype):
    pupil, focal, E, lyot = _setup()
    prop = ha.FraunhoferPropagator(pupil, focal)
    wf0 = ha.Wavefront(ha.Field(dev(E.astype(dtype)), pupil))
    ref = host(prop.forward(wf0).electric_field)
    stop = ha.Field(dev(lyot.astype(float)), pupil)
    for charge in (2, 4, 6, 8):
        vc = ha.VortexCoronagraph(pupil, charge, lyot_stop=stop)
        img = host(prop.forward(vc.forward(wf0)).electric_field)
        power = np.sum(np.abs(img) ** 2) * focal.weights
        ratio = np.abs(img).max() ** 2 / np.abs(ref).max() ** 2
        say('charge', charge, dtype.__name__, 'power', power, 'peak ratio', ratio)
        assert power < 1e-6
        assert ratio < 1e-8


def test_vortex_adjoint_and_gradient_path():
    """<V E, U> = <E, V^dagger U> on the GPU, and the recorded (autograd) path equals the grouped
    one."""
    pupil, focal, E, lyot = _setup(128)
    vc = ha.VortexCoronagraph(pupil, 2, q=64)
    rng = np.random.default_rng(7)
    U = rng.standard_normal(pupil.size) + 1j * rng.standard_normal(pupil.size)
    Et = ha.Field(dev(E), pupil)
    Ut = ha.Field(dev(U), pupil)
    a = np.vdot(U, host(vc.forward(ha.Wavefront(Et)).electric_field))
    b = np.vdot(host(vc.backward(ha.Wavefront(Ut)).electric_field), E)
    assert abs(a - b) < 1e-11 * abs(a)
    x = dev(E).clone().requires_grad_(True)
    L = vc.forward(ha.Wavefront(ha.Field(x, pupil))).electric_field
    plain = host(vc.forward(ha.Wavefront(Et)).electric_field)
    assert peak_rel_field(host(L), plain) < 1e-13
    # d/dx of Re<U, V x> is V^dagger U (PyTorch's conjugate convention for real losses)
    loss = torch.real(torch.sum(torch.conj(Ut.as_subclass(torch.Tensor)) * L.as_subclass(torch.Tensor)))
    loss.backward()
    g = host(x.grad)
    assert peak_rel_field(g, host(vc.backward(ha.Wavefront(Ut)).electric_field)) < 1e-12
