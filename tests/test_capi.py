"""The C-ABI library loads, exports every symbol include/hcipy_amd.h declares, and rejects
invalid arguments with a status code and message (no GPU needed)."""
import ctypes
import os
import re

import pytest

from hcipy_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'hcipy_amd.h')


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(hcipy_\w+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _lib.SIGNATURES, 'ctypes signature missing for %s' % s


def test_abi_version():
    assert _lib.load().hcipy_abi_version() == 1


def test_plan_create_rejects_bad_ndim():
    lib = _lib.load()
    h = ctypes.c_void_p()
    x, px = _lib.darray([0.0, 1.0])
    st = lib.hcipy_mft_plan_create(ctypes.byref(h), 3, 0, px, 2, px, 2, px, 2, px, 2, None)
    assert st == 1
    assert b'one- or two-dimensional' in lib.hcipy_get_last_error()
    st = lib.hcipy_mft_plan_create(ctypes.byref(h), 2, 7, px, 2, px, 2, px, 2, px, 2, None)
    assert st == 1 and b'precision' in lib.hcipy_get_last_error()
    st = lib.hcipy_mft_plan_create(ctypes.byref(h), 1, 0, px, 2, px, 2, px, 2, px, 2, None)
    assert st == 1


def test_null_arguments_are_errors():
    lib = _lib.load()
    n = ctypes.c_size_t()
    assert lib.hcipy_mft_workspace_size(None, 1, ctypes.byref(n)) == 1
    assert lib.hcipy_mft_execute(None, 0, None, None, 1, 1.0, 0.0, None, 0, None) == 1
    assert lib.hcipy_mft_plan_destroy(None) == 0
    assert lib.hcipy_pw_sub_mul(None, None, None, 0, 0, 0, None, 5, 1, 1, None) == 1
    assert lib.hcipy_pw_power_workspace(4, ctypes.byref(n)) == 0 and n.value > 0


def test_kernels_refuse_cpu_tensors():
    import torch
    from hcipy_amd import make_pupil_grid, make_focal_grid, MatrixFourierTransform
    g = make_pupil_grid(8)
    mft = MatrixFourierTransform(g, make_focal_grid(2, 2))
    with pytest.raises(_lib.HcipyError):
        mft.forward(torch.zeros(g.size, dtype=torch.complex128))


def test_zoom_entry_points_reject_bad_arguments():
    """Argument checks of the zoom-FFT entry points run before any device work."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    ci = ctypes.c_int * 2
    cd = ctypes.c_double * 2
    n, d, z = ci(8, 8), cd(0.1, 0.1), cd(0.0, 0.0)
    assert lib.hcipy_zoom_plan_create(ctypes.byref(h), 3, n, d, z, n, d, z, 1, 0, None) == 1
    assert b'one- and two-dimensional' in lib.hcipy_get_last_error()
    assert lib.hcipy_zoom_plan_create(ctypes.byref(h), 2, n, d, z, n, d, z, 2, 0, None) == 1
    assert lib.hcipy_zoom_plan_create(ctypes.byref(h), 2, n, d, z, n, d, z, 1, 1, None) == 1  # FP32 for complex128
    assert b'precision' in lib.hcipy_get_last_error()
    big = ci(6000, 8)
    assert lib.hcipy_zoom_plan_create(ctypes.byref(h), 2, big, d, z, big, d, z, 1, 0, None) == 1
    assert b'8192' in lib.hcipy_get_last_error()
    size = ctypes.c_size_t()
    assert lib.hcipy_zoom_workspace_size(None, 1, ctypes.byref(size)) == 1
    assert lib.hcipy_zoom_execute(None, 0, None, None, 1, 1.0, 0.0, None, 0, None) == 1
    assert lib.hcipy_zoom_plan_destroy(None) == 0


def test_matrix_filter_entry_points_reject_bad_arguments():
    """Argument checks of the matrix filter entry points run before any device work."""
    lib = _lib.load()
    size = ctypes.c_size_t()
    hs = (ctypes.c_void_p * 25)()
    assert lib.hcipy_filter_workspace_size_matrix(None, 2, 2, 1, ctypes.byref(size)) == 1
    assert lib.hcipy_filter_execute_matrix(None, 2, 2, 0, None, None, 1, None, 0, None) == 1
    assert lib.hcipy_filter_execute_matrix(hs, 5, 2, 0, None, None, 1, None, 0, None) == 1
    assert b'at most 4 x 4' in lib.hcipy_get_last_error()
    assert lib.hcipy_filter_execute_matrix(hs, 1, 1, 0, None, None, 1, None, 0, None) == 1
    assert b'NULL plan 0' in lib.hcipy_get_last_error()


def test_zoom_fft_pair_entry_point_checks_the_grids():
    """hcipy_zoom_plan_create_fft: fft_n is required, positive, and must satisfy
    delta_in * delta_out = 2 pi / fft_n; the checks run before any device work."""
    import math
    lib = _lib.load()
    h = ctypes.c_void_p()
    ci = ctypes.c_int * 1
    cd = ctypes.c_double * 1
    n, z = ci(8), cd(0.0)
    d_in, d_out = cd(0.25), cd(2 * math.pi / (16 * 0.25))
    assert lib.hcipy_zoom_plan_create_fft(ctypes.byref(h), 1, n, d_in, z, n, d_out, z, None, 1, 0, None) == 1
    assert b'fft_n' in lib.hcipy_get_last_error()
    assert lib.hcipy_zoom_plan_create_fft(ctypes.byref(h), 1, n, d_in, z, n, d_out, z, ci(0), 1, 0, None) == 1
    assert lib.hcipy_zoom_plan_create_fft(ctypes.byref(h), 1, n, d_in, z, n, d_out, z, ci(17), 1, 0, None) == 1
    assert b'is not 2 pi / 17' in lib.hcipy_get_last_error()


@pytest.mark.parametrize('dims,q,fov,shift,method', [([64, 64], 2, 1, 0, 'fft'), ([64, 32], 2, 0.5, 0.1, 'fft'),
                                                     ([4096], 2, 1, 0.3, 'fft'), ([24, 16], 1.234, 0.6, 0.1, 'zft'),
                                                     ([65], 3, 0.8, 0, 'zft')])
def test_fft_class_hands_its_plan_the_fft_lengths(dims, q, fov, shift, method):
    """FastFourierTransform picks the direct FFT when every internal length is a power of two
    (else the zoom FFT), passes the internal grid's dims as fft_n, and those grids pass the
    plan's delta_in * delta_out = 2 pi / fft_n check (here the call then stops at device
    selection, with no GPU, not at the grid check)."""
    import hcipy_amd as ha
    g = ha.make_uniform_grid(dims, 1.0)
    ft = ha.FastFourierTransform(g, q, fov, shift)
    assert ft.method == method
    assert ft._impl._fft_n == [int(n) for n in ft.internal_grid.dims]
    lib = _lib.load()
    nd = g.ndim
    ci, cd = ctypes.c_int * nd, ctypes.c_double * nd
    out = ft.output_grid
    h = ctypes.c_void_p()
    create = lib.hcipy_fft_plan_create if method == 'fft' else lib.hcipy_zoom_plan_create_fft
    st = create(ctypes.byref(h), nd, ci(*[int(v) for v in g.dims]), cd(*g.delta), cd(*g.zero),
                ci(*[int(v) for v in out.dims]), cd(*out.delta), cd(*out.zero), ci(*ft._impl._fft_n), 1, 0, None)
    if st != 0:
        assert b'is not 2 pi' not in lib.hcipy_get_last_error()
        assert b'power of two' not in lib.hcipy_get_last_error()
    else:
        (lib.hcipy_fft_plan_destroy if method == 'fft' else lib.hcipy_zoom_plan_destroy)(h)


def test_direct_fft_entry_points_check_their_arguments():
    """hcipy_fft_plan_create: power-of-two fft_n in [4, 8192], sample counts within fft_n,
    delta_in * delta_out = 2 pi / fft_n, complex dtypes, FP32 only for complex64; the checks run
    before any device work."""
    import math
    lib = _lib.load()
    h = ctypes.c_void_p()
    ci = ctypes.c_int * 1
    cd = ctypes.c_double * 1
    n, z = ci(8), cd(0.0)
    d_in, d_out = cd(0.25), cd(2 * math.pi / (16 * 0.25))
    create = lib.hcipy_fft_plan_create
    assert create(ctypes.byref(h), 1, n, d_in, z, n, d_out, z, None, 1, 0, None) == 1
    assert create(ctypes.byref(h), 3, n, d_in, z, n, d_out, z, ci(16), 1, 0, None) == 1
    assert b'one- and two-dimensional' in lib.hcipy_get_last_error()
    assert create(ctypes.byref(h), 1, n, d_in, z, n, d_out, z, ci(16), 3, 0, None) == 1
    assert b'complex' in lib.hcipy_get_last_error()
    assert create(ctypes.byref(h), 1, n, d_in, z, n, d_out, z, ci(16), 1, 1, None) == 1
    assert b'precision' in lib.hcipy_get_last_error()
    assert create(ctypes.byref(h), 1, n, d_in, z, n, d_out, z, ci(24), 1, 0, None) == 1
    assert b'power of two' in lib.hcipy_get_last_error()
    assert create(ctypes.byref(h), 1, n, d_in, z, n, d_out, z, ci(16384), 1, 0, None) == 1
    assert b'power of two' in lib.hcipy_get_last_error()
    assert create(ctypes.byref(h), 1, ci(17), d_in, z, n, d_out, z, ci(16), 1, 0, None) == 1
    assert b'17 input' in lib.hcipy_get_last_error()
    assert create(ctypes.byref(h), 1, n, d_in, z, n, cd(1.0), z, ci(16), 1, 0, None) == 1
    assert b'is not 2 pi / 16' in lib.hcipy_get_last_error()
    size = ctypes.c_size_t()
    assert lib.hcipy_fft_workspace_size(None, 1, ctypes.byref(size)) == 1
    assert lib.hcipy_fft_execute(None, 0, None, None, 1, 1.0, 0.0, None, 0, None) == 1
    assert lib.hcipy_fft_plan_destroy(None) == 0
