"""The parity form as one persistent launch (GPU): the parity split, stage 1 and the recombining
stage 2 of a single execution run as tasks of one kernel that wait for each other through device
counters of the plan.  Checked against the oracle's FP64 MFT (hcipy/fourier/matrix_fourier_transform.py:171-172,
219-220), against the three-launch execution of the same plan, over repeated executions (the
counters reset themselves), graph replays, batches, FFT-style grids with their unpaired row and
odd grids with tabulated mirror indices.

Bars (peak-relative): random fields 5e-7 against the oracle, as the three-launch parity form;
1024^2 -> 512^2 Fraunhofer field and intensity 1e-7 (the headline's bar); fused against three
launches 1e-12 (the same arithmetic per tile; only the FP64 sum of the unpaired stage-1 row may
round differently, since the fused split gives each thread other samples)."""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from hcipy_amd import _lib, workloads
from oracle import hcipy_oracle as o

from helpers import dev, host, og, peak_rel_field, say

pytestmark = pytest.mark.gpu

R64X64, RC32X32 = 16, 19

# The persistent launch is disabled in the library (hcipy_mft_plan_set_fused refuses to enable it):
# its first launches on MI355X (1024^2 -> 512^2, then 256^2 -> 128^2 with every dependency wait
# bounded to 10 ms) did not finish within the test time limit.  The kernel's ISA, read on the CPU,
# showed the cause: a task loop whose exit the compiler took as divergent, so that all lanes but
# lane 0 repeated the previous task.  The loop now has a uniform skeleton, which has not run on a
# GPU; the tests that run it stay here, skipped, until it has.
HUNG = pytest.mark.skip(reason='fused parity launch disabled: its launches did not finish on MI355X; '
                               'the rewritten task loop has not run on a GPU')


def _rand(n, seed, lead=()):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(lead + (n,)) + 1j * rng.standard_normal(lead + (n,))


def _plan(mft):
    return mft.plan_for(torch.complex64, torch.device('cuda', 0))


def _pin_fused_shapes(plan):
    for d in (_lib.FORWARD, _lib.BACKWARD):
        plan.set_tiles(d, 0, R64X64)
        plan.set_tiles(d, 1, RC32X32)


@HUNG
def test_fused_first_execution_completes():
    """One fused execution at 256^2 -> 128^2 finishes with every wait satisfied, against the
    oracle."""
    gi = ha.make_pupil_grid(256)
    go = ha.make_focal_grid(4, 16)
    mft = ha.MatrixFourierTransform(gi, go, precision='f16x2')
    plan = _plan(mft)
    _pin_fused_shapes(plan)
    assert plan.fused
    x = _rand(gi.size, 1).astype(np.complex64)
    a = host(mft.forward(ha.Field(dev(x), gi)))
    n = plan.fused_timeouts()
    e = peak_rel_field(a, o.mft_forward(x.astype(np.complex128), og(gi), og(go)))
    say('first fused execution: %d timeouts, error %.3e' % (n, e))
    assert n == 0 and e < 5e-7


@HUNG
def test_fused_headline_1024_to_512():
    pupil, focal, E = workloads.plain_mft_1024()
    E64 = E.astype(np.complex64)
    prop = ha.FraunhoferPropagator(pupil, focal, precision='f16x2')
    wf = ha.Wavefront(ha.Field(dev(E64), pupil), 1.0)
    plan = _plan(prop.instance(1.0).fourier_transform)
    assert plan.parity and plan.fused_available and plan.fused
    b = o.fraunhofer_forward(E64.astype(np.complex128), og(pupil), og(focal))
    Ib = np.abs(b) ** 2
    res = {}
    for fused in (True, False):
        plan.fused = fused
        a = host(prop(wf).electric_field)
        I = host(prop.forward_intensity(wf)).astype(np.float64)
        res[fused] = (a, I)
        ef = peak_rel_field(a, b)
        ei = np.abs(I - Ib).max() / Ib.max()
        say('f16x2 1024->512 fused=%s: field %.3e intensity %.3e' % (fused, ef, ei))
        assert ef < 1e-7 and ei < 1e-7
    plan.fused = True
    assert peak_rel_field(res[True][0], res[False][0]) < 1e-12
    assert np.abs(res[True][1] - res[False][1]).max() / res[False][1].max() < 1e-12
    # the adjoint direction of the same plan (512^2 -> 1024^2: stage-1 and stage-2 shapes of its own)
    F = b.astype(np.complex64)
    back_ref = o.fraunhofer_backward(b, og(pupil), og(focal))
    wf_f = ha.Wavefront(ha.Field(dev(F), focal), 1.0)
    back = {}
    for fused in (True, False):
        plan.fused = fused
        back[fused] = host(prop.backward(wf_f).electric_field)
        e = peak_rel_field(back[fused], back_ref)
        say('  backward fused=%s: %.3e' % (fused, e))
        assert e < 1e-7
    plan.fused = True
    assert peak_rel_field(back[True], back[False]) < 1e-12
    assert plan.fused_timeouts() == 0


@HUNG
@pytest.mark.parametrize('dims_in,dims_out,center', [
    ([256, 256], [128, 128], False),
    ([96, 40], [48, 70], False),
    ([63, 50], [33, 64], False),
    ([128, 96], [64, 90], True),
    ([1024, 1024], [512, 512], True),
])
def test_fused_matches_oracle_and_three_launches(dims_in, dims_out, center):
    """Batch 3, both directions, field outputs, the fused shapes pinned so that the persistent
    launch runs at every size; centred grids (odd sizes: the origin has no separate mirror, and
    the mirror indices are tabulated), and FFT-style grids whose unpaired stage-1 row the split
    forms in FP64."""
    gi = ha.make_uniform_grid(dims_in, [1.0, 0.8], has_center=center)
    go = ha.make_uniform_grid(dims_out, [40.0, 30.0], has_center=center)
    f = _rand(gi.size, 3, (3,)).astype(np.complex64)
    F = _rand(go.size, 4, (3,)).astype(np.complex64)
    ref_f = o.mft_forward(f.astype(np.complex128), og(gi), og(go))
    ref_b = o.mft_backward(F.astype(np.complex128), og(gi), og(go))
    mft = ha.MatrixFourierTransform(gi, go, precision='f16x2')
    plan = _plan(mft)
    assert plan.parity and plan.fused_available
    _pin_fused_shapes(plan)
    out = {}
    for fused in (True, False):
        plan.fused = fused
        a = host(mft.forward(ha.Field(dev(f), gi)))
        bb = host(mft.backward(ha.Field(dev(F), go)))
        out[fused] = (a, bb)
        ea, eb = peak_rel_field(a, ref_f), peak_rel_field(bb, ref_b)
        say('fused=%s %s->%s center=%s: forward %.3e backward %.3e' % (fused, dims_in, dims_out, center, ea, eb))
        assert ea < 5e-7 and eb < 5e-7
    assert peak_rel_field(out[True][0], out[False][0]) < 1e-12
    assert peak_rel_field(out[True][1], out[False][1]) < 1e-12
    assert plan.fused_timeouts() == 0


@HUNG
def test_fused_repeated_executions_and_intensity():
    """The counters clear themselves: ten executions in a row, alternating inputs and output
    kinds, each equal to its first result bit for bit, and the intensity output against the
    oracle."""
    gi = ha.make_pupil_grid(256)
    go = ha.make_focal_grid(4, 16)
    mft = ha.MatrixFourierTransform(gi, go, precision='f16x2')
    plan = _plan(mft)
    _pin_fused_shapes(plan)
    assert plan.fused
    xs = [_rand(gi.size, s, (2,)).astype(np.complex64) for s in (11, 12)]
    refs = [o.mft_forward(x.astype(np.complex128), og(gi), og(go)) for x in xs]
    first = {}
    for it in range(10):
        k = it % 2
        x = ha.Field(dev(xs[k]), gi)
        a = host(mft.forward(x))
        I = host(mft.forward_intensity(x)).astype(np.float64)
        if k not in first:
            first[k] = (a, I)
            assert peak_rel_field(a, refs[k]) < 5e-7
            Ir = np.abs(refs[k]) ** 2
            assert np.abs(I - Ir).max() / Ir.max() < 1e-6
        else:
            assert np.array_equal(a, first[k][0]) and np.array_equal(I, first[k][1]), it
    assert plan.fused_timeouts() == 0


@HUNG
def test_fused_graph_replays():
    """Fused executions recorded in a HIP graph replay correctly, with eager work of the same plan
    between the replays, every replay equal to the eager result."""
    gi = ha.make_pupil_grid(256)
    go = ha.make_focal_grid(4, 16)
    mft = ha.MatrixFourierTransform(gi, go, precision='f16x2')
    plan = _plan(mft)
    _pin_fused_shapes(plan)
    x = ha.Field(dev(_rand(gi.size, 21).astype(np.complex64)), gi)
    y = ha.Field(dev(_rand(gi.size, 22).astype(np.complex64)), gi)
    eager_x = host(mft.forward(x))
    eager_y = host(mft.forward(y))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        mft.forward(x)  # warm: workspace and instance in place before capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = mft.forward(x)
    for r in range(4):
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(host(out), eager_x), r
        assert np.array_equal(host(mft.forward(y)), eager_y), r
        torch.cuda.synchronize()
    assert plan.fused_timeouts() == 0


def test_fused_switch_and_availability():
    """No plan offers the fused execution while it is disabled, enabling it is an error, and every
    parity tile-shape pair (the fused kernel's own among them) runs as three launches."""
    gi = ha.make_pupil_grid(128)
    go = ha.make_focal_grid(4, 8)
    for prec in ('fp32', 'bf16x3', 'f16x2'):
        plan = _plan(ha.MatrixFourierTransform(gi, go, precision=prec))
        assert not plan.fused_available and not plan.fused, prec
        with pytest.raises(_lib.HcipyError):
            plan.fused = True
        plan.fused = False
    mft = ha.MatrixFourierTransform(gi, go, precision='f16x2')
    plan = _plan(mft)
    x = _rand(gi.size, 31).astype(np.complex64)
    ref = o.mft_forward(x.astype(np.complex128), og(gi), og(go))
    for cfg1, cfg2 in ((R64X64, RC32X32), (_lib.TILES_AUTO, _lib.TILES_AUTO), (17, 20)):
        for d in (_lib.FORWARD, _lib.BACKWARD):
            plan.set_tiles(d, 0, cfg1)
            plan.set_tiles(d, 1, cfg2)
        assert peak_rel_field(host(mft.forward(ha.Field(dev(x), gi))), ref) < 5e-7, (cfg1, cfg2)
