"""Gradients through the propagation chain (torch.autograd with the library's adjoint kernels).

Checks:
  * MFT vector-Jacobian product == the weighted adjoint given by MatrixFourierTransform.backward
    (the reference's backward() is the adjoint in the grid-weighted inner product), 1e-12;
  * FourierFilter vector-Jacobian product == FourierFilter.backward (conjugate transfer
    function), 1e-12;
  * pointwise adjoints (phase screen incl. d/dphi, intensity) == plain PyTorch autograd of the
    same ops (fp32 reference for complex64: 1e-5 relative; complex128: 1e-12);
  * d(dark-zone mean intensity)/d(DM actuators) through a DM + Lyot coronagraph chain ==
    central finite differences in FP64, 1e-5 relative (complex128), and the split-FP16
    (complex64) gradient == the same finite differences, 2e-4 of the gradient's peak (limited by
    the complex64 / 22-bit storage of the chain's planes relative to the dark-zone field).
"""
import numpy as np
import pytest
import torch

import hcipy_amd as ha

from helpers import dev, host, say

pytestmark = pytest.mark.gpu


def _crand(shape, seed, dtype=torch.complex128):
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(shape, generator=g, dtype=torch.float64) + 1j * torch.randn(shape, generator=g,
                                                                                   dtype=torch.float64)).to(dtype).cuda()


def test_mft_vjp_is_weighted_adjoint():
    gi = ha.make_pupil_grid(48, 1.0)
    go = ha.make_focal_grid(4, 5).scaled(2 * np.pi)
    mft = ha.MatrixFourierTransform(gi, go)
    E = _crand(gi.size, 1).requires_grad_(True)
    U = _crand(go.size, 2)
    y = mft.forward(ha.Field(E, gi)).as_subclass(torch.Tensor)
    (torch.conj(U) * y).real.sum().backward()
    ratio = gi.weights / (go.weights / (2 * np.pi) ** 2)
    ref = host(mft.backward(ha.Field(U, go))) * ratio
    got = host(E.grad)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # and the backward direction's VJP is the forward kernel with the conjugate scale
    F = _crand(go.size, 3).requires_grad_(True)
    V = _crand(gi.size, 4)
    x = mft.backward(ha.Field(F, go)).as_subclass(torch.Tensor)
    (torch.conj(V) * x).real.sum().backward()
    ref2 = host(mft.forward(ha.Field(V, gi))) / ratio
    assert np.abs(host(F.grad) - ref2).max() <= 1e-12 * np.abs(ref2).max()


def test_fourier_filter_vjp_is_adjoint():
    grid = ha.make_pupil_grid(32, 1e-2)
    prop = ha.AngularSpectrumPropagator(grid, 0.5)
    x = _crand(grid.size, 5).requires_grad_(True)
    v = _crand(grid.size, 6)
    wf = ha.Wavefront(ha.Field(x, grid), 5e-7)
    y = prop.forward(wf).electric_field.as_subclass(torch.Tensor)
    (torch.conj(v) * y).real.sum().backward()
    ref = host(prop.backward(ha.Wavefront(ha.Field(v, grid), 5e-7)).electric_field)
    assert np.abs(host(x.grad) - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize('cdt,tol', [(torch.complex64, 1e-5), (torch.complex128, 1e-12)])
def test_pointwise_adjoints_match_torch(cdt, tol):
    n = 4096
    rdt = torch.float32 if cdt == torch.complex64 else torch.float64
    x0 = _crand((3, n), 7, cdt)
    phi0 = torch.randn(n, dtype=rdt, generator=torch.Generator().manual_seed(8)).cuda()
    w = torch.rand((3, n), dtype=rdt, generator=torch.Generator().manual_seed(9)).cuda()
    s = 0.8
    x = x0.clone().requires_grad_(True)
    phi = phi0.clone().requires_grad_(True)
    I = ha.pointwise.intensity(ha.pointwise.multiply_phase(x, phi, s))
    (I * w).sum().backward()
    xr = x0.clone().requires_grad_(True)
    pr = phi0.clone().requires_grad_(True)
    Ir = (xr * torch.exp(1j * s * pr)).abs() ** 2
    (Ir * w).sum().backward()
    say('pointwise adjoint', cdt)
    assert (x.grad - xr.grad).abs().max().item() <= tol * xr.grad.abs().max().item()
    assert (phi.grad - pr.grad).abs().max().item() <= tol * max(pr.grad.abs().max().item(), 1e-30) + tol


def _dark_hole_setup(cdt, precision):
    pupil = ha.make_pupil_grid(64, 1.0)
    aperture = ha.make_circular_aperture(1.0)(pupil)
    fpm_grid = ha.make_focal_grid(8, 4)
    fpm = 1 - ha.make_circular_aperture(4.0)(fpm_grid)
    lyot = ha.make_circular_aperture(0.95)(pupil)
    focal = ha.make_focal_grid(4, 12)
    coro = ha.LyotCoronagraph(pupil, fpm, lyot, focal_plane_mask_grid=fpm_grid, precision=precision)
    prop = ha.FraunhoferPropagator(pupil, focal, precision=precision)
    r = np.hypot(focal.x, focal.y)
    dz = torch.as_tensor((r > 3) & (r < 8) & (focal.x > 1), device='cuda')
    n_act = 8
    IF = ha.make_gaussian_influence_functions(pupil, n_act, 1.0 / n_act)
    E0 = ha.Field(torch.as_tensor(host(aperture), device='cuda').to(cdt), pupil)
    wl = 1.0

    def loss(act):
        dm = ha.DeformableMirror(IF)
        dm.actuators = act
        wf = ha.Wavefront(E0, wl)
        wf = dm.forward(wf)
        wf = coro.forward(wf)
        img = prop.forward(wf).intensity
        return ha.get_mean_intensity_in_roi(img, dz)

    return loss, n_act


def test_dark_hole_gradient_matches_finite_differences():
    loss, n_act = _dark_hole_setup(torch.complex128, None)
    a0 = torch.as_tensor(np.random.default_rng(11).standard_normal(n_act ** 2) * 2e-3, device='cuda')
    a = a0.clone().requires_grad_(True)
    L = loss(a)
    L.backward()
    g = host(a.grad)
    h = 1e-6
    idx = [9, 18, 27, 36, 45]
    fd = []
    for k in idx:
        e = torch.zeros_like(a0)
        e[k] = h
        with torch.no_grad():
            fd.append((float(loss(a0 + e)) - float(loss(a0 - e))) / (2 * h))
    fd = np.array(fd)
    say('grad', g[idx], 'fd', fd)
    assert np.abs(g[idx] - fd).max() <= 1e-5 * np.abs(fd).max()
    # split-FP16 (complex64) gradient against the FP64 finite differences.  Every plane of the
    # chain is held as complex64 / 22-bit hi-lo values, i.e. to ~1e-7 of the unocculted peak
    # amplitude, while the dark-zone field is ~1e-3 of it: the dark-zone gradient is therefore
    # good to ~1e-4 relative, and that is the bar here.
    loss32, _ = _dark_hole_setup(torch.complex64, 'f16x2')
    b = a0.clone().requires_grad_(True)
    loss32(b).backward()
    g32 = host(b.grad)
    say('f16x2 grad', g32[idx])
    assert np.abs(g32[idx] - fd).max() <= 2e-4 * np.abs(fd).max()
    assert np.abs(g32 - g).max() <= 2e-4 * np.abs(g).max()


def test_roi_mean_gradient():
    img = torch.rand(500, dtype=torch.float64, generator=torch.Generator().manual_seed(3)).cuda().requires_grad_(True)
    mask = torch.rand(500, generator=torch.Generator().manual_seed(4)).cuda() > 0.5
    m = ha.get_mean_intensity_in_roi(img, mask)
    m.backward()
    ref = host(mask).astype(float) / host(mask).sum()
    np.testing.assert_allclose(host(img.grad), ref, rtol=1e-15, atol=0)
