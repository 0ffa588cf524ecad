"""hcipy_mft_execute_noalloc (static device scratch, no allocation) against hcipy_mft_execute_ex
with a caller workspace: two calls in a row through the C ABI, the second on a smaller input,
must give bit-identical outputs; a batch beyond the scratch is refused without a launch."""
import ctypes

import numpy as np
import pytest
import torch

from hcipy_amd import _lib, make_focal_grid, make_pupil_grid
from hcipy_amd.fourier import _MftPlan
from tests.helpers import say

pytestmark = pytest.mark.gpu


def _field(grid, batch, dtype, seed):
    rng = np.random.default_rng(seed)
    r2 = grid.x ** 2 + grid.y ** 2
    ap = (r2 <= 0.25).astype(float)
    e = ap * np.exp(1j * 0.3 * rng.standard_normal((batch, grid.size)))
    return torch.as_tensor(e.astype(dtype), device='cuda').contiguous()


def _run_pair(pupil, focal, precision, batch, output, seed):
    lib = _lib.load()
    cdt = np.complex128 if precision == 'fp64' else np.complex64
    plan = _MftPlan(pupil, focal, precision, torch.device('cuda', 0))
    x = _field(pupil, batch, cdt, seed)
    if output == _lib.OUTPUT_FIELD:
        odt = torch.complex128 if precision == 'fp64' else torch.complex64
    else:
        odt = torch.float64 if precision == 'fp64' else torch.float32
    out_ws = torch.empty((batch, focal.size), dtype=odt, device='cuda')
    out_na = torch.full((batch, focal.size), float('nan'), dtype=odt, device='cuda')
    nbytes = plan.workspace_bytes(batch)  # either direction; the noalloc call checks its own need
    # caller workspace full of NaN bytes: any read before write would show in the output
    ws = torch.full((nbytes,), 0xFF, dtype=torch.uint8, device='cuda')
    alpha = complex(0.7, -0.2)
    stream = _lib.stream_ptr(x.device)
    say('noalloc case', pupil.shape, '->', focal.shape, precision, 'batch', batch, 'output', output)
    st = lib.hcipy_mft_execute_ex(plan.handle, _lib.FORWARD, _lib.dptr(x), _lib.dptr(out_ws), batch, alpha.real,
                                  alpha.imag, output, _lib.dptr(ws), nbytes, stream)
    _lib.check(st, 'hcipy_mft_execute_ex')
    st = lib.hcipy_mft_execute_noalloc(plan.handle, _lib.FORWARD, _lib.dptr(x), _lib.dptr(out_na), batch, alpha.real,
                                       alpha.imag, output, stream)
    _lib.check(st, 'hcipy_mft_execute_noalloc')
    torch.cuda.synchronize()
    assert torch.isfinite(torch.view_as_real(out_ws) if out_ws.is_complex() else out_ws).all()
    assert torch.equal(out_ws, out_na)
    return plan, x


def test_noalloc_matches_workspace_entry_point_on_two_calls():
    # first call: the bench's transform, split-FP16, two fields (24 MiB of the 32 MiB scratch)
    plan, x = _run_pair(make_pupil_grid(1024), make_focal_grid(4, 64), 'f16x2', 2, _lib.OUTPUT_INTENSITY, 1)
    # second call: a smaller transform in FP64 over the same scratch, whose earlier contents stay
    _run_pair(make_pupil_grid(192), make_focal_grid(4, 12), 'fp64', 3, _lib.OUTPUT_FIELD, 2)
    _run_pair(make_pupil_grid(96), make_focal_grid(2, 10), 'f16x2', 1, _lib.OUTPUT_FIELD, 3)


def test_noalloc_refuses_batches_beyond_the_scratch():
    lib = _lib.load()
    pupil, focal = make_pupil_grid(1024), make_focal_grid(4, 64)
    plan = _MftPlan(pupil, focal, 'f16x2', torch.device('cuda', 0))
    assert plan.workspace_bytes(3) > 32 * 1024 * 1024
    x = torch.zeros((3, pupil.size), dtype=torch.complex64, device='cuda')
    out = torch.full((3, focal.size), 5.0, dtype=torch.float32, device='cuda')
    st = lib.hcipy_mft_execute_noalloc(plan.handle, _lib.FORWARD, _lib.dptr(x), _lib.dptr(out), 3, 1.0, 0.0,
                                       _lib.OUTPUT_INTENSITY, _lib.stream_ptr(x.device))
    assert st == 3
    assert b'static scratch' in lib.hcipy_get_last_error()
    torch.cuda.synchronize()
    assert (out == 5.0).all()
    st = lib.hcipy_mft_execute_noalloc(None, _lib.FORWARD, None, None, 1, 1.0, 0.0, 0, None)
    assert st == 1
