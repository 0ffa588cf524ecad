"""Parity of the MFMA matrix Fourier transform with the FP64 oracle (GPU).

Tolerances (peak-relative, i.e. max error / max |reference|):
  complex128 (FP64 MFMA)                 field 1e-12
  complex64 split FP16x2 (FP64 flushes)  intensity 1e-7 at 1024^2 -> 512^2 (the bar of the
                                         throughput metric), field 2e-7
  complex64 plain FP32 MFMA              intensity 3e-6, field 3e-6
"""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from hcipy_amd import workloads
from oracle import hcipy_oracle as o

from helpers import dev, host, og, peak_rel_field, peak_rel_intensity, say

pytestmark = pytest.mark.gpu


def _rand(n, seed, lead=()):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(lead + (n,)) + 1j * rng.standard_normal(lead + (n,))


@pytest.mark.parametrize('dims_in,dims_out', [([7, 5], [6, 9]), ([16, 16], [8, 8]), ([33, 65], [70, 31])])
def test_fp64_forward_backward_small(dims_in, dims_out):
    gi = ha.make_uniform_grid(dims_in, [1.0, 0.8]).shifted([0.03, -0.02])
    go = ha.make_uniform_grid(dims_out, [20.0, 30.0]).shifted([0.7, -0.4])
    f = _rand(gi.size, 0)
    F = _rand(go.size, 1)
    mft = ha.MatrixFourierTransform(gi, go)
    say('fp64 small', dims_in, dims_out)
    a = host(mft.forward(ha.Field(dev(f), gi)))
    b = o.mft_forward(f, og(gi), og(go))
    assert peak_rel_field(a, b) < 1e-12
    a = host(mft.backward(ha.Field(dev(F), go)))
    b = o.mft_backward(F, og(gi), og(go))
    assert peak_rel_field(a, b) < 1e-12


def test_fp64_batched_tensor_field():
    gi = ha.make_pupil_grid(24)
    go = ha.make_focal_grid(3, 5)
    f = _rand(gi.size, 2, (2, 3))
    mft = ha.MatrixFourierTransform(gi, go)
    a = host(mft.forward(ha.Field(dev(f), gi)))
    assert a.shape == (2, 3, go.size)
    b = o.mft_forward(f, og(gi), og(go))
    assert peak_rel_field(a, b) < 1e-12
    a = host(mft.backward(ha.Field(dev(b), go)))
    assert peak_rel_field(a, o.mft_backward(b, og(gi), og(go))) < 1e-12


def test_fp64_one_dimensional():
    gi = ha.make_uniform_grid(65, 1.0)
    go = ha.make_uniform_grid(50, 40.0).shifted(0.3)
    f = _rand(gi.size, 3, (3,))
    mft = ha.MatrixFourierTransform(gi, go)
    a = host(mft.forward(ha.Field(dev(f), gi)))
    assert peak_rel_field(a, o.mft_forward(f, og(gi), og(go))) < 1e-12
    a = host(mft.backward(ha.Field(dev(f[:, :50]), go)))
    assert peak_rel_field(a, o.mft_backward(f[:, :50], og(gi), og(go))) < 1e-12


def test_fraunhofer_256_to_128_complex128():
    """256^2 pupil -> 128^2 focal (q=4, 16 lambda/D), complex128, parity 1e-12."""
    pupil, focal, E = workloads.plain_mft_256()
    prop = ha.FraunhoferPropagator(pupil, focal)
    say('fraunhofer 256->128 c128')
    a = host(prop(ha.Wavefront(ha.Field(dev(E), pupil), 1.0)).electric_field)
    b = o.fraunhofer_forward(E, og(pupil), og(focal))
    assert peak_rel_field(a, b) < 1e-12
    assert peak_rel_intensity(a, b) < 1e-12


def test_fraunhofer_1024_to_512_complex128():
    """1024^2 -> 512^2 (4 spider vanes, 0.05 rad phase), complex128 on the FP64 MFMA path against
    the oracle's FP64 transform: 1e-12 of the peak field and intensity."""
    pupil, focal, E = workloads.plain_mft_1024()
    prop = ha.FraunhoferPropagator(pupil, focal)
    say('fraunhofer 1024->512 c128')
    wf = ha.Wavefront(ha.Field(dev(E), pupil), 1.0)
    a = host(prop(wf).electric_field)
    assert a.dtype == np.complex128
    b = o.fraunhofer_forward(E, og(pupil), og(focal))
    ef, ei = peak_rel_field(a, b), peak_rel_intensity(a, b)
    say('  field %.3e  intensity %.3e' % (ef, ei))
    assert ef < 1e-12 and ei < 1e-12
    back = host(prop.backward(ha.Wavefront(ha.Field(dev(b), focal), 1.0)).electric_field)
    assert peak_rel_field(back, o.fraunhofer_backward(b, og(pupil), og(focal))) < 1e-12


@pytest.mark.parametrize('precision', ['f16x2', 'bf16x3', 'fp32'])
def test_fraunhofer_1024_to_512_complex64(precision):
    """1024^2 -> 512^2 complex64 against the FP64 result on the same (complex64) input.

    Bars (peak-relative): f16x2 and bf16x3 field 1e-7, fused intensity 1e-7, intensity of the
    complex64 field 2.5e-7 (rounding re/im to FP32 alone moves |E|^2 by up to 2^-23 = 1.2e-7 at
    the peak); fp32 field and intensities 3e-6.
    """
    pupil, focal, E = workloads.plain_mft_1024()
    E64 = E.astype(np.complex64)
    prop = ha.FraunhoferPropagator(pupil, focal, precision=precision)
    say('fraunhofer 1024->512 c64', precision)
    wf = ha.Wavefront(ha.Field(dev(E64), pupil), 1.0)
    a = host(prop(wf).electric_field)
    assert a.dtype == np.complex64
    I = host(prop.forward_intensity(wf))
    assert I.dtype == np.float32
    b = o.fraunhofer_forward(E64.astype(np.complex128), og(pupil), og(focal))
    Ib = np.abs(b) ** 2
    ef = peak_rel_field(a, b)
    ei_field = peak_rel_intensity(a, b)
    ei = np.abs(I.astype(np.float64) - Ib).max() / Ib.max()
    say('  field %.3e  intensity(fused) %.3e  intensity(from c64 field) %.3e' % (ef, ei, ei_field))
    if precision in ('f16x2', 'bf16x3'):
        assert ef < 1e-7 and ei < 1e-7 and ei_field < 2.5e-7
    else:
        assert ef < 3e-6 and ei < 3e-6 and ei_field < 3e-6


@pytest.mark.parametrize('precision,tol', [('f16x2', 5e-7), ('bf16x3', 5e-7), ('fp32', 3e-6)])
def test_backward_256_to_1024_complex64(precision, tol):
    """Lyot return leg: 256^2 focal (q=32, 4 lambda/D) -> 1024^2 pupil, adjoint MFT, on random
    complex data.  Random data has no coherent peak, so the error is set by the 22-bit hi/lo
    representation (2^-22 = 2.4e-7 per element, x sqrt(2) for re/im); bar 5e-7."""
    pupil = ha.make_pupil_grid(1024)
    fpm_grid = ha.make_focal_grid(32, 4)
    F = _rand(fpm_grid.size, 7).astype(np.complex64)
    prop = ha.FraunhoferPropagator(pupil, fpm_grid, precision=precision)
    a = host(prop.backward(ha.Wavefront(ha.Field(dev(F), fpm_grid), 1.0)).electric_field)
    b = o.fraunhofer_backward(F.astype(np.complex128), og(pupil), og(fpm_grid))
    e = peak_rel_field(a, b)
    say('backward 256->1024', precision, 'field peak-rel %.3e' % e)
    assert e < tol


def test_f16x2_handles_wide_dynamic_range():
    """Per-row power-of-two scaling keeps the split path accurate for tiny and huge fields."""
    gi = ha.make_pupil_grid(64)
    go = ha.make_focal_grid(4, 8)
    f = _rand(gi.size, 9)
    b = o.mft_forward(f, og(gi), og(go))
    for prec in ('f16x2', 'bf16x3'):
        mft = ha.MatrixFourierTransform(gi, go, precision=prec)
        for scale in [1e-12, 1.0, 1e9]:
            a = host(mft.forward(ha.Field(dev(f * scale).to(torch.complex64), gi))) / scale
            assert peak_rel_field(a, b) < 2e-6, (prec, scale)


def test_precompute_and_intermediate_flags():
    """M1 / intermediate_array presence follows the flags (tests/test_fourier.py:256-271)."""
    gi = ha.make_pupil_grid(32)
    go = ha.make_fft_grid(gi, 1, 0.25)
    f = _rand(gi.size, 4)
    ref = None
    for pm in [True, False]:
        for ai in [True, False]:
            mft = ha.MatrixFourierTransform(gi, go, precompute_matrices=pm, allocate_intermediate=ai)
            mft.forward(ha.Field(dev(np.zeros(gi.size)), gi))
            r = host(mft.forward(ha.Field(dev(f), gi)))
            assert (mft.M1 is not None) == pm
            assert (mft.intermediate_array is not None) == ai
            ref = r if ref is None else ref
            assert np.array_equal(r, ref)


@pytest.mark.parametrize('dtype', ['complex128', 'complex64'])
def test_energy_and_roundtrip_fov1(dtype):
    """forward x backward at q=1, fov=1 is the identity; energy conserved
    (tests/test_fourier.py:22-87 bars: c128 1e-12, c64 1e-6 / 1e-5)."""
    g = ha.make_uniform_grid([9, 18], 1, has_center=True).shifted([0.1, 0.1]).scaled(2)
    go = ha.make_fft_grid(g, 1, 1, [0.1, 0.1])
    rng = np.random.default_rng(0)
    f = rng.standard_normal(g.size).astype(dtype)
    mft = ha.MatrixFourierTransform(g, go)
    F = mft.forward(ha.Field(dev(f), g))
    back = host(mft.backward(F))
    assert back.dtype == np.dtype(dtype)
    e_in = np.sum(np.abs(f) ** 2) * g.weights
    e_out = np.sum(np.abs(back) ** 2) * g.weights
    tol = 1e-12 if dtype == 'complex128' else 1e-5
    assert abs(e_out / e_in - 1) < tol
    assert np.std(np.abs(back - f)) / np.abs(f).mean() < (1e-12 if dtype == 'complex128' else 1e-6)


@pytest.mark.parametrize('dtype', ['complex128', 'complex64'])
def test_even_odd_symmetry_and_linearity(dtype):
    tol = 1e-12 if dtype == 'complex128' else 1e-6
    g = ha.make_uniform_grid([9, 9], 1, has_center=False).scaled(2)
    go = ha.make_fft_grid(g, 3, 0.8, [0.1, 0.1])
    go = ha.make_fft_grid(g, 3, 0.8)
    rng = np.random.default_rng(0)
    f = rng.standard_normal(g.size)
    fe = ((f + f[::-1]) / 2).astype(dtype)
    fo = ((f - f[::-1]) / 2).astype(dtype)
    mft = ha.MatrixFourierTransform(g, go)
    Fe = host(mft.forward(ha.Field(dev(fe), g)))
    Fo = host(mft.forward(ha.Field(dev(fo), g)))
    assert np.abs(Fe.imag).max() < tol * max(1, np.abs(Fe).max())
    assert np.abs(Fo.real).max() < tol * max(1, np.abs(Fo).max())
    Fb = host(mft.forward(ha.Field(dev(f.astype(dtype)), g)))
    assert np.allclose(Fb, Fe + Fo, atol=tol * np.abs(Fb).max())


def test_fraunhofer_airy_closed_form():
    """tests/test_propagation.py:6-35 through this library (512^2 pupil, bar 3e-5)."""
    import scipy.special
    pupil = ha.make_pupil_grid(512)
    focal = ha.make_focal_grid(16, 8)
    ap = ha.evaluate_supersampled(ha.make_circular_aperture(0.7), pupil, 8)
    prop = ha.FraunhoferPropagator(pupil, focal, focal_length=0.8)
    img = host(prop(ha.Wavefront(ap, 0.5)).electric_field)
    img = img / img[np.argmax(np.abs(img))]
    r = np.hypot(focal.x, focal.y)
    x = r * np.pi / 0.5 * 0.7 / 0.8
    ind = np.argmin(r)
    x[ind] = 1
    ref = 2 * scipy.special.jv(1, x) / x
    ref[ind] = 1
    assert np.abs(img - ref).max() < 3e-5


@pytest.mark.parametrize('precision,cdt,tol', [('fp64', np.complex128, 1e-12), ('fp32', np.complex64, 3e-6),
                                               ('f16x2', np.complex64, 5e-7), ('bf16x3', np.complex64, 5e-7)])
def test_every_tile_shape_matches_oracle(precision, cdt, tol):
    """Each workgroup tile shape (pinned on the plan for both stages and directions) gives the
    oracle's result: ragged sizes exercise the k tails, the row clamps and the k-split sums."""
    from hcipy_amd import _lib
    gi = ha.make_uniform_grid([100, 90], [1.0, 0.9]).shifted([0.01, -0.02])
    go = ha.make_uniform_grid([70, 50], [30.0, 25.0]).shifted([0.4, -0.3])
    f = _rand(gi.size, 11, (2,)).astype(cdt)
    F = _rand(go.size, 12, (2,)).astype(cdt)
    ref_f = o.mft_forward(f.astype(np.complex128), og(gi), og(go))
    ref_b = o.mft_backward(F.astype(np.complex128), og(gi), og(go))
    mft = ha.MatrixFourierTransform(gi, go, precision=None if precision == 'fp64' else precision)
    tdt = torch.complex128 if cdt == np.complex128 else torch.complex64
    plan = mft.plan_for(tdt, torch.device('cuda', 0))
    for cfg in _lib.TILE_CFGS[precision]:
        for d in (_lib.FORWARD, _lib.BACKWARD):
            for stage in (0, 1):
                plan.set_tiles(d, stage, cfg)
                assert plan.get_tiles(d, stage) == cfg
        a = host(mft.forward(ha.Field(dev(f), gi)))
        b = host(mft.backward(ha.Field(dev(F), go)))
        say(precision, 'tile shape', cfg, 'errors %.3e %.3e' % (peak_rel_field(a, ref_f), peak_rel_field(b, ref_b)))
        assert peak_rel_field(a, ref_f) < tol, cfg
        assert peak_rel_field(b, ref_b) < tol, cfg


def test_autotune_pins_supported_shapes_and_keeps_results():
    """The autotuner picks one available shape per stage and direction (of the parity form on
    these centred grids); results after tuning still match the oracle."""
    from hcipy_amd import _lib
    gi = ha.make_pupil_grid(96)
    go = ha.make_focal_grid(4, 6)
    f = _rand(gi.size, 21, (2,)).astype(np.complex64)
    mft = ha.MatrixFourierTransform(gi, go, precision='f16x2')
    chosen = mft.autotune(ha.Field(dev(f), gi), repeats=2)
    say('autotune chose', chosen)
    plan = mft.plan_for(torch.complex64, torch.device('cuda', 0))
    shapes = _lib.PARITY_TILE_CFGS['f16x2'] if plan.parity else _lib.TILE_CFGS['f16x2']
    stage2 = _lib.PARITY_STAGE2_TILE_CFGS['f16x2'] if plan.parity else ()
    for d, key in ((_lib.FORWARD, 'forward'), (_lib.BACKWARD, 'backward')):
        for stage in (0, 1):
            assert chosen[key][stage] in shapes + (stage2 if stage == 1 else ())
            assert plan.get_tiles(d, stage) == chosen[key][stage]
    a = host(mft.forward(ha.Field(dev(f), gi)))
    assert peak_rel_field(a, o.mft_forward(f.astype(np.complex128), og(gi), og(go))) < 5e-7
    F = host(mft.forward(ha.Field(dev(f), gi)))
    b = host(mft.backward(ha.Field(dev(F.astype(np.complex64)), go)))
    assert peak_rel_field(b, o.mft_backward(F.astype(np.complex128), og(gi), og(go))) < 5e-7


@pytest.mark.parametrize('dims_in,dims_out', [([128, 64], [64, 64]), ([256, 128], [64, 128]), ([96, 64], [70, 50])])
def test_lds_dma_shapes_match_oracle(dims_in, dims_out):
    """The LDS-DMA split-FP16 kernels (contraction lengths in whole 128-element k-tiles; other
    lengths take the compact kernel) against the oracle, both directions, batch 2."""
    from hcipy_amd import _lib
    gi = ha.make_uniform_grid(dims_in, [1.0, 0.9]).shifted([0.01, -0.02])
    go = ha.make_uniform_grid(dims_out, [30.0, 25.0]).shifted([0.4, -0.3])
    f = _rand(gi.size, 13, (2,)).astype(np.complex64)
    F = _rand(go.size, 14, (2,)).astype(np.complex64)
    ref_f = o.mft_forward(f.astype(np.complex128), og(gi), og(go))
    ref_b = o.mft_backward(F.astype(np.complex128), og(gi), og(go))
    mft = ha.MatrixFourierTransform(gi, go, precision='f16x2')
    plan = mft.plan_for(torch.complex64, torch.device('cuda', 0))
    for cfg in (_lib.TILES_D64X64_W4, _lib.TILES_D64X64_W8, _lib.TILES_D64X32_W4, _lib.TILES_D64X32_W8):
        for d in (_lib.FORWARD, _lib.BACKWARD):
            for stage in (0, 1):
                plan.set_tiles(d, stage, cfg)
        say('lds dma', dims_in, dims_out, cfg)
        a = host(mft.forward(ha.Field(dev(f), gi)))
        b = host(mft.backward(ha.Field(dev(F), go)))
        assert peak_rel_field(a, ref_f) < 5e-7, cfg
        assert peak_rel_field(b, ref_b) < 5e-7, cfg


@pytest.mark.parametrize('precision,dtype,tol', [(None, np.complex128, 1e-12), ('f16x2', np.complex64, 5e-7)])
def test_mft_nonuniform_weights_vs_oracle(precision, dtype, tol):
    """A separated input grid with uneven x spacing (per-sample weights): the weights are applied
    by the pointwise kernel before the plain product, forward and backward match the oracle, and
    the gradient through that pass equals the adjoint transform times the weights."""
    rng = np.random.default_rng(51)
    x = np.sort(np.linspace(-0.5, 0.5, 48) + rng.uniform(-0.004, 0.004, 48))
    y = np.linspace(-0.5, 0.5, 40)
    gi = ha.CartesianGrid(ha.SeparatedCoords([x, y]))
    go = ha.make_focal_grid(4, 5)
    assert np.ptp(np.asarray(gi.weights)) > 0
    f = (rng.standard_normal(gi.size) + 1j * rng.standard_normal(gi.size)).astype(dtype)
    F = (rng.standard_normal(go.size) + 1j * rng.standard_normal(go.size)).astype(dtype)
    mft = ha.MatrixFourierTransform(gi, go, precision=precision)
    a = host(mft.forward(ha.Field(dev(f), gi)))
    ref = o.mft_forward(f.astype(np.complex128), og(gi), og(go))
    b = host(mft.backward(ha.Field(dev(F), go)))
    refb = o.mft_backward(F.astype(np.complex128), og(gi), og(go))
    ea, eb = peak_rel_field(a, ref), peak_rel_field(b, refb)
    say('non-uniform weights', precision, 'forward %.2e backward %.2e' % (ea, eb))
    assert ea <= tol and eb <= tol
    # d/df of Re<forward(f), g> = conj-linear adjoint: w * M^H g (M the plain product)
    t = dev(f).requires_grad_(True)
    (mft.forward(ha.Field(t, gi)).as_subclass(torch.Tensor) * dev(F).conj()).real.sum().backward()
    w = np.asarray(gi.weights, dtype=float)
    # M^H g from the oracle's backward (which carries the output-grid weight / (2 pi)^2)
    Mh = o.mft_backward(F.astype(np.complex128), og(gi), og(go)) / (np.asarray(go.weights).ravel()[0] / (2 * np.pi) ** 2)
    eg = peak_rel_field(host(t.grad), w * Mh)
    say('non-uniform weights', precision, 'gradient %.2e' % eg)
    assert eg <= tol
