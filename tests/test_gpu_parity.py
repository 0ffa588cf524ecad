"""Parity form of the split-precision MFT between centred grids (GPU): the field's four x/y
parity components times real half-size cosine and sine matrices, recombined into the four
output quadrants.  Checked against the complex form of the same plan and against the oracle's
FP64 MFT (hcipy/fourier/matrix_fourier_transform.py:171-172, 219-220), both directions, field
and intensity outputs, batches, every parity tile shape.

Bars (peak-relative): random fields 5e-7 (the 22-bit split of each operand, 2^-22 = 2.4e-7 per
element, as for the complex form); the 1024^2 -> 512^2 Fraunhofer workload 1e-7 in field and
fused intensity, the headline's bar."""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from hcipy_amd import _lib, workloads
from oracle import hcipy_oracle as o

from helpers import dev, host, og, peak_rel_field, say

pytestmark = pytest.mark.gpu


def _rand(n, seed, lead=()):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(lead + (n,)) + 1j * rng.standard_normal(lead + (n,))


def _plan(mft):
    return mft.plan_for(torch.complex64, torch.device('cuda', 0))


@pytest.mark.parametrize('precision', ['f16x2', 'bf16x3'])
def test_parity_headline_1024_to_512(precision):
    pupil, focal, E = workloads.plain_mft_1024()
    E64 = E.astype(np.complex64)
    prop = ha.FraunhoferPropagator(pupil, focal, precision=precision)
    wf = ha.Wavefront(ha.Field(dev(E64), pupil), 1.0)
    plan = _plan(prop.instance(1.0).fourier_transform)
    assert plan.parity_available and plan.parity
    b = o.fraunhofer_forward(E64.astype(np.complex128), og(pupil), og(focal))
    Ib = np.abs(b) ** 2
    res = {}
    for parity in (True, False):
        plan.parity = parity
        a = host(prop(wf).electric_field)
        I = host(prop.forward_intensity(wf)).astype(np.float64)
        res[parity] = a
        ef = peak_rel_field(a, b)
        ei = np.abs(I - Ib).max() / Ib.max()
        say('%s 1024->512 parity=%s: field %.3e intensity %.3e' % (precision, parity, ef, ei))
        assert ef < 1e-7 and ei < 1e-7
    plan.parity = True
    d = peak_rel_field(res[True], res[False])
    say('  parity vs complex form: %.3e' % d)
    assert d < 1e-7


@pytest.mark.parametrize('precision', ['f16x2', 'bf16x3'])
@pytest.mark.parametrize('dims_in,dims_out', [([64, 64], [32, 32]), ([96, 40], [48, 70]), ([63, 50], [33, 64]),
                                              ([256, 256], [128, 128]), ([1024, 1024], [256, 256])])
def test_parity_matches_oracle_random_fields(precision, dims_in, dims_out):
    """Centred rectangular grids (odd sizes hold the origin, which has no separate mirror),
    batch 3, forward and backward, field outputs, against the oracle and against the complex
    form of the same plan."""
    _check_against_oracle(ha.make_uniform_grid(dims_in, [1.0, 0.8]), ha.make_uniform_grid(dims_out, [40.0, 30.0]),
                          precision, '%s->%s' % (dims_in, dims_out))


@pytest.mark.parametrize('precision', ['f16x2', 'bf16x3'])
@pytest.mark.parametrize('dims_in,dims_out', [([128, 96], [64, 90]), ([2048, 64], [64, 64]),
                                              ([1024, 1024], [512, 512])])
def test_parity_fft_style_grids_with_nyquist_rows(precision, dims_in, dims_out):
    """FFT-style grids (a sample at the origin and an unpaired -n/2 sample) whose unpaired
    magnitude would take a stage-1 row tile of its own (x axes of 128, 2048 and 1024 samples, u
    axes of 64 and 512): the parity split forms that row of the stage-1 product in FP64 instead
    (over 1025 contraction values with 2048 x samples), in both directions.  The smaller cases run
    every parity tile shape."""
    gi = ha.make_uniform_grid(dims_in, [1.0, 0.8], has_center=True)
    go = ha.make_uniform_grid(dims_out, [40.0, 30.0], has_center=True)
    _check_against_oracle(gi, go, precision, 'FFT-style %s->%s' % (dims_in, dims_out),
                          all_tiles=dims_in[0] != 1024)


def _check_against_oracle(gi, go, precision, label, all_tiles=False):
    f = _rand(gi.size, 3, (3,)).astype(np.complex64)
    F = _rand(go.size, 4, (3,)).astype(np.complex64)
    ref_f = o.mft_forward(f.astype(np.complex128), og(gi), og(go))
    ref_b = o.mft_backward(F.astype(np.complex128), og(gi), og(go))
    mft = ha.MatrixFourierTransform(gi, go, precision=precision)
    plan = _plan(mft)
    assert plan.parity_available and plan.parity
    out = {}
    for parity in (True, False):
        plan.parity = parity
        a = host(mft.forward(ha.Field(dev(f), gi)))
        bb = host(mft.backward(ha.Field(dev(F), go)))
        out[parity] = (a, bb)
        ea, eb = peak_rel_field(a, ref_f), peak_rel_field(bb, ref_b)
        say('%s %s parity=%s: forward %.3e backward %.3e' % (precision, label, parity, ea, eb))
        assert ea < 5e-7 and eb < 5e-7
    assert peak_rel_field(out[True][0], out[False][0]) < 5e-7
    assert peak_rel_field(out[True][1], out[False][1]) < 5e-7
    if all_tiles:
        plan.parity = True
        for cfg1, cfg2 in _shape_pairs(precision):
            for d in (_lib.FORWARD, _lib.BACKWARD):
                plan.set_tiles(d, 0, cfg1)
                plan.set_tiles(d, 1, cfg2)
            ea = peak_rel_field(host(mft.forward(ha.Field(dev(f), gi))), ref_f)
            eb = peak_rel_field(host(mft.backward(ha.Field(dev(F), go))), ref_b)
            say('  tile shapes %d/%d: forward %.3e backward %.3e' % (cfg1, cfg2, ea, eb))
            assert ea < 5e-7 and eb < 5e-7, (cfg1, cfg2)


def _shape_pairs(precision):
    """(stage 1, stage 2) tile shapes: every shared parity shape on both stages, then the
    recombining stage-2 shapes after the default stage 1."""
    return ([(c, c) for c in _lib.PARITY_TILE_CFGS[precision]] +
            [(_lib.TILES_AUTO, c) for c in _lib.PARITY_STAGE2_TILE_CFGS[precision]])


@pytest.mark.parametrize('precision', ['f16x2', 'bf16x3'])
def test_every_parity_tile_shape(precision):
    gi = ha.make_pupil_grid(128)
    go = ha.make_focal_grid(4, 10)
    f = _rand(gi.size, 5, (2,)).astype(np.complex64)
    F = _rand(go.size, 6, (2,)).astype(np.complex64)
    ref_f = o.mft_forward(f.astype(np.complex128), og(gi), og(go))
    ref_b = o.mft_backward(F.astype(np.complex128), og(gi), og(go))
    mft = ha.MatrixFourierTransform(gi, go, precision=precision)
    plan = _plan(mft)
    assert plan.parity
    for cfg1, cfg2 in _shape_pairs(precision):
        for d in (_lib.FORWARD, _lib.BACKWARD):
            plan.set_tiles(d, 0, cfg1)
            plan.set_tiles(d, 1, cfg2)
        for output in ('field', 'intensity'):
            if output == 'field':
                a = host(mft.forward(ha.Field(dev(f), gi)))
                ea = peak_rel_field(a, ref_f)
            else:
                a = host(mft.forward_intensity(ha.Field(dev(f), gi))).astype(np.float64)
                ea = np.abs(a - np.abs(ref_f) ** 2).max() / (np.abs(ref_f) ** 2).max()
            assert ea < (5e-7 if output == 'field' else 1e-6), (cfg1, cfg2, output)
        b = host(mft.backward(ha.Field(dev(F), go)))
        say(precision, 'parity tile shapes', cfg1, cfg2, 'errors %.3e %.3e' % (ea, peak_rel_field(b, ref_b)))
        assert peak_rel_field(b, ref_b) < 5e-7, (cfg1, cfg2)
    # shapes of the complex form only are refused while the parity form is on, and the
    # recombining shapes on stage 1
    with pytest.raises(_lib.HcipyError):
        plan.set_tiles(_lib.FORWARD, 0, _lib.TILES_C64X64_W4)
    with pytest.raises(_lib.HcipyError):
        plan.set_tiles(_lib.FORWARD, 0, _lib.PARITY_STAGE2_TILE_CFGS[precision][0])


def test_parity_availability_rules():
    """Grids symmetric about 0 on every axis (pupil-style, FFT-style with its one unpaired
    sample, odd sizes) take the parity form; shifted grids, FP32 plans and 1-D plans keep the
    complex form."""
    p = ha.make_pupil_grid(64)
    f = ha.make_focal_grid(4, 8)
    cases = [(p, f, 'f16x2', True), (p, f, 'bf16x3', True), (p, f, 'fp32', False),
             (ha.make_pupil_grid(63), f, 'f16x2', True), (p, ha.make_focal_grid(4, 8, 1.0).shifted([0.1, 0]), 'f16x2', False),
             (ha.make_uniform_grid([64], [1.0]), ha.make_uniform_grid([32], [20.0]), 'f16x2', False)]
    for gi, go, prec, ok in cases:
        mft = ha.MatrixFourierTransform(gi, go, precision=prec)
        plan = _plan(mft)
        assert plan.parity_available == ok and plan.parity == ok, (gi.dims, go.dims, prec)
        if not ok:
            with pytest.raises(_lib.HcipyError):
                plan.parity = True
        x = _rand(gi.size, 8).astype(np.complex64)
        ref = o.mft_forward(x.astype(np.complex128), og(gi), og(go))
        tol = 5e-7 if prec != 'fp32' else 3e-6
        assert peak_rel_field(host(mft.forward(ha.Field(dev(x), gi))), ref) < tol
