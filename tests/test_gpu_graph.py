"""The propagation chain inside a captured HIP graph: every replay (with eager work and new
inputs in between) must equal the eager result of the same input bit for bit.  Plans and
workspaces are created by an eager warm-up call; no execute path allocates or records memsets."""
import numpy as np
import pytest
import torch

import hcipy_amd as ha

from helpers import host, say

pytestmark = pytest.mark.gpu


def test_graph_replays_equal_eager():
    pupil = ha.make_pupil_grid(128)
    focal = ha.make_focal_grid(4, 16)
    fpm_grid = ha.make_focal_grid(8, 4)
    aperture = torch.as_tensor(host(ha.make_circular_aperture(1.0)(pupil)), device='cuda')
    fpm = 1 - ha.make_circular_aperture(3.0)(fpm_grid)
    lyot = ha.make_circular_aperture(0.95)(pupil)
    prop32 = ha.FraunhoferPropagator(pupil, focal, precision='f16x2')
    coro64 = ha.LyotCoronagraph(pupil, fpm, lyot, focal_plane_mask_grid=fpm_grid)
    prop64 = ha.FraunhoferPropagator(pupil, focal)
    asp_grid = ha.make_pupil_grid(128, 1e-2)
    asp = ha.AngularSpectrumPropagator(asp_grid, 0.2)
    IF = ha.make_gaussian_influence_functions(pupil, 8, 1 / 8)
    dm = ha.DeformableMirror(IF)
    dm.actuators = np.random.default_rng(1).standard_normal(64) * 1e-3

    x = torch.zeros(pupil.size, dtype=torch.complex128, device='cuda')

    def chain():
        wf = ha.Wavefront(ha.Field(x * aperture, pupil), 1.0)
        wf = dm.forward(wf)
        i32 = prop32.forward(ha.Wavefront(ha.Field(wf.electric_field.as_subclass(torch.Tensor).to(torch.complex64), pupil), 1.0)).intensity
        i64 = prop64.forward(coro64.forward(wf)).intensity
        near = asp.forward(ha.Wavefront(ha.Field(x.to(torch.complex64), asp_grid), 5e-7)).electric_field
        return [t.as_subclass(torch.Tensor) for t in (i32, i64, near)]

    def new_input(seed):
        g = torch.Generator().manual_seed(seed)
        ph = torch.randn(pupil.size, generator=g, dtype=torch.float64) * 0.1
        return torch.exp(1j * ph).cuda()

    x.copy_(new_input(0))
    chain()  # warm-up: plans, matrices, workspaces
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        chain()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        outs = chain()
    torch.cuda.synchronize()
    for r in range(5):
        x.copy_(new_input(r + 1))
        graph.replay()
        torch.cuda.synchronize()
        got = [o.clone() for o in outs]
        ref = chain()  # eager, same input
        torch.cuda.synchronize()
        say('replay', r)
        for a, b in zip(got, ref):
            assert torch.equal(a, b)


def test_grouped_launches_replay_in_graphs():
    """Grouped MFT launches recorded under capture: the vortex coronagraph runs two grouped
    launches per forward, and the captured sequence holds two forwards, so each group records two
    launches with different tables.  Every replay must equal the eager result bit for bit."""
    pupil = ha.make_pupil_grid(128)
    vc = ha.VortexCoronagraph(pupil, 2, q=64)
    assert len(vc.focal_grids) == 4
    x1 = torch.zeros(pupil.size, dtype=torch.complex64, device='cuda')
    x2 = torch.zeros_like(x1)

    def chain():
        a = vc.forward(ha.Wavefront(ha.Field(x1, pupil))).electric_field
        b = vc.forward(ha.Wavefront(ha.Field(x2, pupil))).electric_field
        return [t.as_subclass(torch.Tensor) for t in (a, b)]

    def new_input(seed):
        g = torch.Generator().manual_seed(seed)
        return torch.complex(torch.randn(pupil.size, generator=g), torch.randn(pupil.size, generator=g)).cuda()

    x1.copy_(new_input(0))
    x2.copy_(new_input(1))
    chain()
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        chain()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        outs = chain()
    torch.cuda.synchronize()
    for r in range(4):
        x1.copy_(new_input(2 * r + 2))
        x2.copy_(new_input(2 * r + 3))
        graph.replay()
        torch.cuda.synchronize()
        got = [o.clone() for o in outs]
        ref = chain()
        torch.cuda.synchronize()
        say('grouped replay', r)
        assert not torch.equal(got[0], got[1])
        for a, b in zip(got, ref):
            assert torch.equal(a, b)


def test_zoom_filter_and_vector_vortex_replay_in_graphs():
    """The zoom FFT (both complex64 precisions and complex128), a 2048-point angular-spectrum
    filter and the vector vortex inside one captured graph: every replay equals the eager result
    bit for bit."""
    pupil = ha.make_pupil_grid(256)
    focal = ha.make_focal_grid(4, 16)
    uv = ha.FraunhoferPropagator(pupil, focal).instance(1.0).uv_grid
    z64 = ha.ZoomFastFourierTransform(pupil, uv, precision='fp64')
    z32 = ha.ZoomFastFourierTransform(pupil, uv, precision='fp32')
    ag = ha.make_pupil_grid(1024, 1.0)
    asp = ha.AngularSpectrumPropagator(ag, 1000.0)
    big = torch.zeros(ag.size, dtype=torch.complex64, device='cuda')
    vvc = ha.VectorVortexCoronagraph(2, q=64)
    x = torch.zeros(pupil.size, dtype=torch.complex64, device='cuda')

    def chain():
        xin = ha.Field(x, pupil)
        outs = [z64.forward(xin), z32.forward(xin), z64.forward(ha.Field(x.to(torch.complex128), pupil)),
                asp.forward(ha.Wavefront(ha.Field(big, ag), 5e-7)).electric_field,
                vvc.forward(ha.Wavefront(xin)).electric_field]
        return [t.as_subclass(torch.Tensor) for t in outs]

    def new_input(seed):
        g = torch.Generator().manual_seed(seed)
        return torch.complex(torch.randn(pupil.size, generator=g), torch.randn(pupil.size, generator=g)).cuda()

    def new_big(seed):
        g = torch.Generator().manual_seed(seed + 100)
        return torch.complex(torch.randn(big.numel(), generator=g), torch.randn(big.numel(), generator=g)).cuda()

    x.copy_(new_input(0))
    big.copy_(new_big(0))
    chain()
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        chain()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        outs = chain()
    torch.cuda.synchronize()
    for r in range(4):
        x.copy_(new_input(r + 1))
        big.copy_(new_big(r + 1))
        graph.replay()
        torch.cuda.synchronize()
        got = [o.clone() for o in outs]
        ref = chain()
        torch.cuda.synchronize()
        say('zoom/filter/vector-vortex replay', r)
        for a, b in zip(got, ref):
            assert torch.equal(a, b)


def test_plan_creation_under_capture_raises_cleanly():
    """Plan creation allocates device memory and waits on the stream, which a graph cannot record:
    the MFT, filter and zoom-FFT plan constructors refuse a capturing stream with a clear error
    before touching the device, and the same objects then build and run eagerly."""
    from hcipy_amd import _lib
    from hcipy_amd.fourier import _MftPlan, _ZoomPlan
    from hcipy_amd.fourier_filter import _FilterPlan

    dev = torch.device('cuda', 0)
    pupil = ha.make_pupil_grid(64)
    focal = ha.make_focal_grid(4, 8)
    tf = torch.exp(1j * torch.linspace(0, 1, 128 * 128, dtype=torch.float64)).to(torch.complex64).cuda()
    filt = ha.FourierFilter(pupil, tf, q=2)
    torch.cuda.synchronize()
    makers = [('mft', lambda: _MftPlan(pupil, focal, 'f16x2', dev)),
              ('zoom', lambda: _ZoomPlan(pupil, focal, torch.complex64, 'fp64', dev)),
              ('filter', lambda: _FilterPlan(filt, torch.complex64, dev, tf=tf))]
    for name, make in makers:
        say('capture refusal', name)
        g = torch.cuda.CUDAGraph()
        with pytest.raises(_lib.HcipyError, match='before the capture starts'):
            with torch.cuda.graph(g):
                make()
        del g
        torch.cuda.synchronize()
        plan = make()  # eager creation still works
        torch.cuda.synchronize()
        del plan
    torch.cuda.synchronize()
    x = torch.ones(pupil.size, dtype=torch.complex64, device='cuda')
    y = ha.MatrixFourierTransform(pupil, focal, precision='f16x2').forward(ha.Field(x, pupil))
    torch.cuda.synchronize()
    assert torch.isfinite(torch.view_as_real(y.as_subclass(torch.Tensor))).all()


def test_group_capture_snapshots_are_released_on_request():
    """A group serves 64 captured calls, refuses the next one with a clear error, and serves again
    after release_captures(); every graph replays equal to the eager result."""
    from hcipy_amd import _lib
    from hcipy_amd.fourier import MftGroup

    pupil = ha.make_pupil_grid(64)
    focal = ha.make_focal_grid(4, 8)
    mft = ha.MatrixFourierTransform(pupil, focal, precision='f16x2')
    g0 = torch.Generator().manual_seed(3)
    x = torch.complex(torch.randn(2, pupil.size, generator=g0), torch.randn(2, pupil.size, generator=g0)).cuda()
    group = MftGroup(capacity=2)
    calls = [(mft, _lib.FORWARD, x[0], 1.0, _lib.OUTPUT_FIELD), (mft, _lib.FORWARD, x[1], 0.5j, _lib.OUTPUT_INTENSITY)]
    ref = [o.clone() for o in group.execute(calls)]
    torch.cuda.synchronize()

    def capture_and_check(i):
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            outs = group.execute(calls)
        for _ in range(2):
            g.replay()
            torch.cuda.synchronize()
            for a, b in zip(outs, ref):
                assert torch.equal(a, b), i
        del g, outs

    for i in range(64):
        capture_and_check(i)
    say('64 captured group calls replayed')
    g = torch.cuda.CUDAGraph()
    with pytest.raises(_lib.HcipyError, match='capture snapshots'):
        with torch.cuda.graph(g):
            group.execute(calls)
    del g
    torch.cuda.synchronize()
    assert group.release_captures() == 64
    for i in range(6):
        capture_and_check(64 + i)
    group.close()
