"""Polarization on the device: the Stokes and Jones-product kernels against the oracle, the
Jones-matrix elements, and the vector vortex coronagraph against the oracle and the reference's
tests (tests/test_coronagraphy.py:81-147) (GPU).

Tolerances: the Stokes kernel works in FP64 (1e-13 relative for complex128 fields, 1e-6 for
complex64 fields whose float32 outputs round at 6e-8); the Jones product is one complex
multiply-add per component (1e-15 / 1e-6).  The vector vortex runs the same transforms as the
scalar vortex, so its Lyot-plane field matches the oracle to 1e-10 (complex128) and 2e-6
(complex64) of the input peak."""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from hcipy_amd import pointwise
from oracle import hcipy_oracle as o

from helpers import dev, host, og, peak_rel_field, say

pytestmark = pytest.mark.gpu

_STATES = [(1, 0, 0, 0), (1, 1, 0, 0), (1, 0, -1, 0), (1, 0, 0, 1), (0.7, 0.1, -0.3, 0.4)]


def _crand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize('dtype,tol', [(np.complex128, 1e-13), (np.complex64, 1e-6)])
def test_stokes_kernel(dtype, tol):
    n = 1000
    J = _crand((2, 2, n), 1).astype(dtype)
    for s in _STATES:
        ref = o.stokes_vector(J.astype(np.complex128), s)
        allc = host(pointwise.stokes(dev(J), s))
        assert allc.shape == (4, n)
        scale = np.abs(ref).max()
        assert np.abs(allc - ref).max() < tol * scale, s
        for c in range(4):
            one = host(pointwise.stokes(dev(J), s, component=c))
            assert np.abs(one - ref[c]).max() < tol * scale
    e = _crand((2, n), 2).astype(dtype)
    ref = o.stokes_vector(e.astype(np.complex128))
    assert np.abs(host(pointwise.stokes(dev(e))) - ref).max() < tol * np.abs(ref).max()
    # Wavefront properties: scalar, Jones vector, partially polarized
    g = ha.make_pupil_grid([50, 20])
    wf = ha.Wavefront(ha.Field(dev(J), g), 1, (1, 0, 0, 1))
    sv = o.stokes_vector(J.astype(np.complex128), (1, 0, 0, 1))
    for name, k in (('I', 0), ('Q', 1), ('U', 2), ('V', 3)):
        assert np.abs(host(getattr(wf, name)) - sv[k]).max() < tol * np.abs(sv).max()
    assert np.abs(host(wf.stokes_vector) - sv).max() < tol * np.abs(sv).max()
    wv = ha.Wavefront(ha.Field(dev(e), g))
    assert np.abs(host(wv.V) - o.stokes_vector(e.astype(np.complex128))[3]).max() < tol * np.abs(ref).max()
    ws = ha.Wavefront(ha.Field(dev(e[0]), g))
    np.testing.assert_array_equal(host(ws.Q), np.zeros(n))
    st = host(ws.stokes_vector)
    assert np.abs(st[0] - np.abs(e[0].astype(np.complex128)) ** 2).max() < tol * st[0].max()
    # total power of a partially polarized wavefront: the FP64 sum of I times the weights
    p = wf.total_power
    w = float(np.asarray(g.weights).ravel()[0])
    assert abs(p - sv[0].sum() * w) < 1e-12 * abs(p) + tol * np.abs(sv[0]).sum() * w


@pytest.mark.parametrize('jdt,edt,tol', [(np.complex128, np.complex128, 1e-15), (np.complex128, np.complex64, 1e-6),
                                         (np.complex64, np.complex64, 1e-6)])
def test_field_dot(jdt, edt, tol):
    n = 777
    for p, q, rest in ((2, 2, ()), (2, 2, (2,)), (3, 2, (4,)), (2, 3, (3, 2))):
        J = _crand((p, q, n), 3).astype(jdt)
        E = _crand((q,) + rest + (n,), 4).astype(edt)
        out = host(pointwise.field_dot(dev(J), dev(E)))
        ref = o.field_dot(J.astype(np.complex128), E.astype(np.complex128))
        assert out.shape == ref.shape and out.dtype == edt
        assert peak_rel_field(out, ref) < 4 * tol, (p, q, rest)
        F = _crand((p,) + rest + (n,), 5).astype(edt)
        back = host(pointwise.field_dot(dev(J), dev(F), adjoint=True))
        refb = o.field_dot(np.conj(np.swapaxes(J, 0, 1)).astype(np.complex128), F.astype(np.complex128))
        assert peak_rel_field(back, refb) < 4 * tol


def test_jones_elements_and_adjoint():
    """LinearRetarder / CircularRetarder (polarization.py:121-220) on scalar and polarized
    wavefronts; backward is the adjoint of forward, and the autograd gradient is J^H g."""
    g = ha.make_pupil_grid(32)
    n = g.size
    E = _crand(n, 6)
    theta = np.linspace(0, np.pi, n)
    hwp = ha.LinearRetarder(np.pi, ha.Field(dev(theta), g))
    out = hwp.forward(ha.Wavefront(ha.Field(dev(E), g)))
    J = o.phase_retarder_jones(np.pi, theta, 0)
    assert out.is_partially_polarized and tuple(out.electric_field.shape) == (2, 2, n)
    ref = o.field_dot(J, np.eye(2)[:, :, None] * E)
    assert peak_rel_field(host(out.electric_field), ref) < 1e-15
    # a polarized wavefront: U of a Jones-vector input
    ev = _crand((2, n), 7)
    qwp = ha.CircularRetarder(np.pi / 2)
    Jc = o.phase_retarder_jones(np.pi / 2, np.pi / 4, np.pi / 2)
    w = qwp.forward(ha.Wavefront(ha.Field(dev(ev), g)))
    assert peak_rel_field(host(w.electric_field), Jc @ ev) < 1e-15
    # adjoint: <J E, U> = <E, J^H U>
    U = _crand((2, n), 8)
    a = np.vdot(U, host(hwp.forward(ha.Wavefront(ha.Field(dev(ev), g))).electric_field))
    b = np.vdot(host(hwp.backward(ha.Wavefront(ha.Field(dev(U), g))).electric_field), ev)
    assert abs(a - b) < 1e-12 * abs(a)
    x = dev(ev).requires_grad_(True)
    y = hwp.forward(ha.Wavefront(ha.Field(x, g))).electric_field.as_subclass(torch.Tensor)
    torch.real(torch.sum(torch.conj(dev(U)) * y)).backward()
    assert peak_rel_field(host(x.grad), host(hwp.backward(ha.Wavefront(ha.Field(dev(U), g))).electric_field)) < 1e-14


def _setup(n=256, obstructed=False):
    pupil = ha.make_pupil_grid(n)
    focal = ha.make_focal_grid(4, 32)
    if obstructed:
        ap = ha.make_obstructed_circular_aperture(1, 0.2)
        st = ha.make_obstructed_circular_aperture(0.99, 0.2 / 0.99)
    else:
        ap = ha.make_circular_aperture(1)
        st = ha.make_circular_aperture(0.99)
    ap = host(ha.evaluate_supersampled(ap, pupil, 8))
    lyot = host(ha.evaluate_supersampled(st, pupil, 8)) > 1 - 1e-5
    E = ap.astype(complex)
    E = E / np.sqrt(np.sum(np.abs(E) ** 2) * pupil.weights)
    return pupil, focal, E, lyot


@pytest.mark.parametrize('dtype,tol', [(np.complex128, 1e-10), (np.complex64, 2e-6)])
@pytest.mark.parametrize('n', [256, 96])
def test_vector_vortex_matches_oracle(dtype, tol, n):
    """n = 96: the level-0 filter grid is not a power of two (element filters on the FFT class)."""
    pupil, focal, E, lyot = _setup(n)
    grids, masks = o.multiscale_masks(og(pupil), o.vector_vortex_mask(2))
    E2 = np.eye(2)[:, :, None] * E
    ref = o.multiscale_forward(E2, og(pupil), grids, masks)
    vvc = ha.VectorVortexCoronagraph(2)
    wf = ha.Wavefront(ha.Field(dev(E.astype(dtype)), pupil))
    out = vvc.forward(wf)
    assert out.is_partially_polarized
    L = host(out.electric_field)
    coro = vvc.get_instance_data(pupil, None, 1).coronagraph
    for m, r in zip(coro.focal_masks, masks):
        assert peak_rel_field(host(m), r) < 1e-12
    err = peak_rel_field(L, ref)
    say('vector vortex', dtype.__name__, 'Lyot plane error vs oracle', err)
    assert err < tol
    refb = o.multiscale_forward(E2, og(pupil), grids, masks, adjoint=True)
    Lb = host(vvc.backward(wf).electric_field)
    assert peak_rel_field(Lb, refb) < tol


@pytest.mark.parametrize('dtype', [np.complex128, np.complex64])
def test_vector_vortex_reference_thresholds(dtype):
    """tests/test_coronagraphy.py:81-118: total power < 1e-6 and peak < 1e-8 of the unocculted
    peak for charges 2..8; the backward pass of a 2-tensor image with only the (0, 0) component
    equals that of the scalar image."""
    pupil, focal, E, lyot = _setup()
    prop = ha.FraunhoferPropagator(pupil, focal)
    wf0 = ha.Wavefront(ha.Field(dev(E.astype(dtype)), pupil))
    ref = host(prop.forward(wf0).intensity)
    stop = ha.Field(dev(lyot.astype(float)), pupil)
    for charge in (2, 4, 6, 8):
        vvc = ha.VectorVortexCoronagraph(charge)
        wf = vvc.forward(wf0)
        wf = ha.Apodizer(stop).forward(wf)
        img = prop.forward(wf)
        power = img.total_power
        ratio = host(img.intensity).max() / ref.max()
        say('vector vortex charge', charge, dtype.__name__, 'power', power, 'peak ratio', ratio)
        assert power < 1e-6
        assert ratio < 1e-8
        Ei = img.electric_field.as_subclass(torch.Tensor).clone()
        Ei[0, 1] = 0
        Ei[1, 0] = 0
        Ei[1, 1] = 0
        b = vvc.backward(prop.backward(ha.Wavefront(ha.Field(Ei, focal), 1, (1, 0, 0, 0))))
        c = vvc.backward(prop.backward(ha.Wavefront(ha.Field(Ei[0, 0].contiguous(), focal))))
        bb, cc = host(b.electric_field)[0, 0], host(c.electric_field)[0, 0]
        assert np.abs(bb - cc).max() <= 1e-6 * np.abs(cc).max()


def test_vector_vortex_circular_polarization():
    """tests/test_coronagraphy.py:120-176: circular input (Stokes (1, 0, 0, 1)) leaves a charge 2
    vector vortex in the orthogonal state (I = -V) and returns to I = V backward; the scalar
    vortex keeps the state (I = V both ways)."""
    pupil, focal, E, lyot = _setup(obstructed=True)
    prop = ha.FraunhoferPropagator(pupil, focal)
    stop = ha.Apodizer(ha.Field(dev(lyot.astype(float)), pupil))
    for coro, sign in ((ha.VectorVortexCoronagraph(2), -1), (ha.VortexCoronagraph(pupil, 2), 1)):
        wf = ha.Wavefront(ha.Field(dev(E), pupil), 1, (1, 0, 0, 1))
        wf.total_power = 1
        img = prop.forward(stop.forward(coro.forward(wf)))
        I, V = host(img.intensity), host(img.V)
        say(type(coro).__name__, 'max |I - sign V|', np.abs(I - sign * V).max(), 'peak', I.max())
        np.testing.assert_allclose(I, sign * V, rtol=1e-5, atol=1e-8)
        back = coro.backward(prop.backward(img))
        np.testing.assert_allclose(host(back.intensity), host(back.V), rtol=1e-5, atol=1e-8)
    # against the oracle's Stokes parameters of the same chain
    vvc = ha.VectorVortexCoronagraph(2)
    wf = ha.Wavefront(ha.Field(dev(E), pupil), 1, (1, 0, 0, 1))
    img = prop.forward(stop.forward(vvc.forward(wf)))
    grids, masks = o.multiscale_masks(og(pupil), o.vector_vortex_mask(2))
    L = o.multiscale_forward(np.eye(2)[:, :, None] * E, og(pupil), grids, masks) * lyot
    st = o.stokes_vector(o.fraunhofer_forward(L, og(pupil), og(focal)), (1, 0, 0, 1))
    assert np.abs(host(img.stokes_vector) - st).max() < 1e-10 * st[0].max()
