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


@pytest.mark.parametrize('precision', ['f16x2', 'bf16x3'])
def test_grouped_parity_form(precision):
    """Grouped launches in the parity form (every item's plan runs it): one pupil field sent to
    several focal grids (one shared parity split), FFT-style focal and pupil axes whose unpaired
    row the first stage-1 row tile forms itself (a single call forms it in the split), both
    directions, batches, field and intensity outputs; against the same transforms called one by
    one (5e-7 of the peak, 1e-6 for intensities) and against the oracle."""
    from oracle import hcipy_oracle as o
    from helpers import og
    calls, refs, oracle_refs = [], [], []
    k = 200
    grids = [(ha.make_pupil_grid(64), ha.make_focal_grid(4, 8)),                     # focal h = 33
             (ha.make_pupil_grid(96), ha.make_focal_grid(4, 8)),
             (ha.make_uniform_grid([128, 128], 1.0, has_center=True), ha.make_focal_grid(3, 10))]  # pupil h = 65
    for gi, go0 in grids:
        E = _rand((2, gi.size), k, np.complex64)
        k += 1
        for wl in (0.9, 1.0, 1.1):
            go = go0.scaled(2 * np.pi / wl)
            w_in = float(np.asarray(gi.weights).ravel()[0])
            w_out = float(np.asarray(go.weights).ravel()[0]) / (2 * np.pi) ** 2
            mft = ha.MatrixFourierTransform(gi, go, precision=precision)
            assert mft.plan_for(torch.complex64, torch.device('cuda', 0)).parity
            alpha = complex(0.2, -0.5 * wl)
            x = dev(E)
            for output in (_lib.OUTPUT_FIELD, _lib.OUTPUT_INTENSITY):
                calls.append((mft, _lib.FORWARD, x, alpha, output))
                refs.append(host(mft._execute_plain(x.contiguous(), _lib.FORWARD, alpha, output)))
                # the oracle's transform carries the input weight, the plain execution alpha only
                f = o.mft_forward(E.astype(np.complex128), og(gi), og(go)) * (alpha / w_in)
                oracle_refs.append(f if output == _lib.OUTPUT_FIELD else np.abs(f) ** 2)
            F = _rand((3, go.size), k, np.complex64)
            k += 1
            calls.append((mft, _lib.BACKWARD, dev(F), alpha, _lib.OUTPUT_FIELD))
            refs.append(host(mft._execute_plain(dev(F), _lib.BACKWARD, alpha)))
            oracle_refs.append(o.mft_backward(F.astype(np.complex128), og(gi), og(go)) * (alpha / w_out))
    group = ha.MftGroup(64)
    for rep in range(2):
        outs = group.execute(calls)
        worst = 0.0
        for (_, _, _, _, output), out, r, ro in zip(calls, outs, refs, oracle_refs):
            a = host(out)
            if output == _lib.OUTPUT_FIELD:
                e, eo = peak_rel_field(a, r), peak_rel_field(a, ro)
                assert e < 5e-7 and eo < 5e-7
            else:
                a = a.astype(np.float64)
                e = np.abs(a - r).max() / r.max()
                eo = np.abs(a - ro).max() / ro.max()
                assert e < 1e-6 and eo < 1e-6
            worst = max(worst, e, eo)
        say('grouped parity', precision, len(calls), 'items, pass', rep, 'worst error %.2e' % worst)


@pytest.mark.parametrize('precision,cdt,tol', [(None, torch.complex128, 1e-12), ('f16x2', torch.complex64, 5e-7),
                                               ('fp32', torch.complex64, 3e-6)])
def test_stacked_wavelengths_match_single_calls(precision, cdt, tol):
    """Rows of a stacked field at several wavelengths: FraunhoferPropagator.forward_stacked /
    backward_stacked and LyotCoronagraph.forward_stacked equal forward / backward at each
    wavelength (grouped launches take other tile shapes: the split-FP16 bar is the grouped one),
    and their vector-Jacobian products equal the single calls' (autograd through the grouped
    adjoints); forward_multiple / backward_multiple give the same as lists."""
    pupil = ha.make_pupil_grid(96, 1.0)
    focal = ha.make_focal_grid(4, 8, spatial_resolution=500e-9)
    fpm_grid = ha.make_focal_grid(8, 3, spatial_resolution=500e-9)
    wls = list(500e-9 * np.array([0.95, 1.0, 1.05]))
    rng = np.random.default_rng(31)
    x = torch.as_tensor(_rand((len(wls), pupil.size), 32, np.complex128), device='cuda').to(cdt)
    prop = ha.FraunhoferPropagator(pupil, focal, precision=precision)
    fpm = ha.Field(torch.as_tensor((np.hypot(fpm_grid.x, fpm_grid.y) > 1.5e-6).astype(float), device='cuda'), fpm_grid)
    stop = ha.Field(torch.as_tensor((np.hypot(pupil.x, pupil.y) < 0.45).astype(float), device='cuda'), pupil)
    coro = ha.LyotCoronagraph(pupil, fpm, stop, focal_plane_mask_grid=fpm_grid, precision=precision)

    def peak_err(a, b):
        return (a - b).abs().max().item() / b.abs().max().item()

    F = prop.forward_stacked(x, wls)
    F1 = torch.stack([prop.forward(ha.Wavefront(ha.Field(x[i], pupil), wl)).electric_field.as_subclass(torch.Tensor)
                      for i, wl in enumerate(wls)])
    B = prop.backward_stacked(F1, wls)
    B1 = torch.stack([prop.backward(ha.Wavefront(ha.Field(F1[i], focal), wl)).electric_field.as_subclass(torch.Tensor)
                      for i, wl in enumerate(wls)])
    L = coro.forward_stacked(x, wls)
    L1 = torch.stack([coro.forward(ha.Wavefront(ha.Field(x[i], pupil), wl)).electric_field.as_subclass(torch.Tensor)
                      for i, wl in enumerate(wls)])
    Lm = torch.stack([w.electric_field.as_subclass(torch.Tensor) for w in
                      coro.forward_multiple([ha.Wavefront(ha.Field(x[i], pupil), wl) for i, wl in enumerate(wls)])])
    errs = [peak_err(F, F1), peak_err(B, B1), peak_err(L, L1), peak_err(Lm, L1)]
    say('stacked wavelengths', precision, 'forward %.2e backward %.2e lyot %.2e lyot list %.2e' % tuple(errs))
    assert max(errs) <= tol

    # vector-Jacobian products of the stacked chain against the single calls'
    g = torch.as_tensor(_rand((len(wls), focal.size), 33, np.complex128), device='cuda').to(cdt)
    xs = x.clone().requires_grad_(True)
    (prop.forward_stacked(coro.forward_stacked(xs, wls), wls) * g.conj()).real.sum().backward()
    grads = []
    for i, wl in enumerate(wls):
        xi = x[i].clone().requires_grad_(True)
        out = prop.forward(coro.forward(ha.Wavefront(ha.Field(xi, pupil), wl))).electric_field.as_subclass(torch.Tensor)
        (out * g[i].conj()).real.sum().backward()
        grads.append(xi.grad)
    eg = peak_err(xs.grad, torch.stack(grads))
    say('stacked wavelengths', precision, 'vjp %.2e' % eg)
    assert eg <= tol


def test_row_scaled_phase_matches_per_row_calls():
    """multiply_phase with one scale per row (hcipy_pw_phase_rows) equals one call per row bit for
    bit, and its gradient with respect to the shared phase equals the sum of the per-row
    gradients (FP64, 1e-14 of the peak)."""
    n, rows = 4096, 5
    rng = np.random.default_rng(41)
    x = torch.as_tensor(_rand((rows, n), 42, np.complex128), device='cuda')
    phi = torch.as_tensor(rng.standard_normal(n) * 1e-9, device='cuda')
    scales_np = 4 * np.pi / (500e-9 * np.linspace(0.95, 1.05, rows))
    scales = torch.as_tensor(scales_np, dtype=torch.float64, device='cuda')
    y = ha.pointwise.multiply_phase(x, phi, scales)
    y1 = torch.stack([ha.pointwise.multiply_phase(x[i], phi, float(scales_np[i])) for i in range(rows)])
    assert torch.equal(y, y1)
    g = torch.as_tensor(_rand((rows, n), 43, np.complex128), device='cuda')
    p = phi.clone().requires_grad_(True)
    (ha.pointwise.multiply_phase(x, p, scales) * g.conj()).real.sum().backward()
    gsum = torch.zeros_like(phi)
    for i in range(rows):
        pi = phi.clone().requires_grad_(True)
        (ha.pointwise.multiply_phase(x[i], pi, float(scales_np[i])) * g[i].conj()).real.sum().backward()
        gsum = gsum + pi.grad
    e = (p.grad - gsum).abs().max().item() / gsum.abs().max().item()
    say('row-scaled phase gradient vs per-row sum %.2e' % e)
    assert e <= 1e-14
    with pytest.raises(ValueError):
        ha.pointwise.multiply_phase(x, phi, scales[:3])



def test_stacked_wavelengths_with_batch_dimension():
    """Stacked fields of shape (wavelengths, batch, n): each wavelength's batch runs as one grouped
    item and matches per-field calls (split FP16, 5e-7 of the peak)."""
    pupil = ha.make_pupil_grid(64, 1.0)
    focal = ha.make_focal_grid(4, 6, spatial_resolution=500e-9)
    wls = list(500e-9 * np.array([0.97, 1.03]))
    prop = ha.FraunhoferPropagator(pupil, focal, precision='f16x2')
    x = torch.as_tensor(_rand((2, 3, pupil.size), 61, np.complex64), device='cuda')
    F = prop.forward_stacked(x, wls)
    assert tuple(F.shape) == (2, 3, focal.size)
    worst = 0.0
    for i, wl in enumerate(wls):
        for k in range(3):
            ref = prop.forward(ha.Wavefront(ha.Field(x[i, k], pupil), wl)).electric_field.as_subclass(torch.Tensor)
            worst = max(worst, (F[i, k] - ref).abs().max().item() / ref.abs().max().item())
    say('stacked wavelengths with a batch dimension: %.2e' % worst)
    assert worst <= 5e-7
