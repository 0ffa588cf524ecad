"""ctypes binding of the C-ABI library ``libhcipy_amd.so`` (declarations in include/hcipy_amd.h).

There is no fallback: if the library is missing or no ROCm device is present the operations
raise.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libhcipy_amd.so')

PREC_FP64 = 0
PREC_FP32 = 1
PREC_F16X2 = 2
PREC_BF16X3 = 3
PRECISIONS = {'fp64': PREC_FP64, 'fp32': PREC_FP32, 'f16x2': PREC_F16X2, 'bf16x3': PREC_BF16X3}
FORWARD = 0
BACKWARD = 1
OUTPUT_FIELD = 0
OUTPUT_FIELD_C128 = 1
OUTPUT_INTENSITY = 2
OUTPUT_INTENSITY_F64 = 3
TILES_AUTO = -1


TILES_64X64_W2X2 = 0
TILES_64X32_W2X1 = 1
TILES_64X64_K4 = 2
TILES_64X32_K4 = 3
TILES_128X64 = 4
TILES_64X128 = 5
TILES_64X64_W8 = 6
TILES_64X32_W8 = 7
TILES_C64X64_W4 = 8
TILES_C64X64_W8 = 9
TILES_C64X32_W4 = 10
TILES_C128X32_W8 = 11
TILES_D64X64_W4 = 12
TILES_D64X64_W8 = 13
TILES_D64X32_W4 = 14
TILES_D64X32_W8 = 15
TILES_R64X64_W4 = 16
TILES_R64X32_W4 = 17
TILES_R32X64_W4 = 18
TILES_RC32X32_W4 = 19
TILES_RC32X32_H = 20
TILES_RC32X32_W8 = 21
TILES_R64X64_H = 22
TILES_128X64_H = 23
# tile shapes of the poke sets' grouped launches (hcipy_poke_set_autotune / hcipy_poke_set_set_tiles)
POKE_TILE_CFGS = {'fp64': (0, 1, 6), 'fp32': (0, 1, 4, 6, 7), 'f16x2': (0, 1, 4, 6, 7, 23), 'bf16x3': (1, 4, 7)}
TILE_CFGS = {'fp64': (0, 1, 6), 'fp32': (0, 1, 2, 3, 4, 5, 6, 7), 'f16x2': tuple(range(16)),
             'bf16x3': (0, 1, 4, 7)}
# tile shapes of the parity form (hcipy_mft_plan_set_parity)
PARITY_TILE_CFGS = {'f16x2': (0, 1, 4, 5, 6, 7, 16, 17, 18, 22), 'bf16x3': (0, 1, 4, 6, 7, 22)}
# stage 2 of the parity form only: the recombination runs in the stage's epilogue
PARITY_STAGE2_TILE_CFGS = {'f16x2': (19, 20, 21), 'bf16x3': (20, 21)}


class MftGroupItem(ctypes.Structure):
    """hcipy_mft_group_item of include/hcipy_amd.h."""
    _fields_ = [('plan', ctypes.c_void_p), ('direction', ctypes.c_int), ('in_', ctypes.c_void_p),
                ('out', ctypes.c_void_p), ('batch', ctypes.c_int64), ('alpha_re', ctypes.c_double),
                ('alpha_im', ctypes.c_double), ('output', ctypes.c_int), ('workspace', ctypes.c_void_p),
                ('workspace_bytes', ctypes.c_size_t)]


class DmGaussianDesc(ctypes.Structure):
    """hcipy_dm_gaussian_desc of include/hcipy_amd.h."""
    _fields_ = [('nx', ctypes.c_int64), ('ny', ctypes.c_int64), ('x0', ctypes.c_double), ('y0', ctypes.c_double),
                ('dx', ctypes.c_double), ('dy', ctypes.c_double), ('n_act', ctypes.c_int64),
                ('act_x0', ctypes.c_double), ('act_y0', ctypes.c_double), ('pitch', ctypes.c_double),
                ('sigma', ctypes.c_double), ('radius', ctypes.c_double), ('x_tilt', ctypes.c_double),
                ('y_tilt', ctypes.c_double), ('z_tilt', ctypes.c_double), ('oversampling', ctypes.c_int64)]


class DmTableDesc(ctypes.Structure):
    """hcipy_dm_table_desc of include/hcipy_amd.h."""
    _fields_ = [('nx', ctypes.c_int64), ('ny', ctypes.c_int64), ('x0', ctypes.c_double), ('y0', ctypes.c_double),
                ('dx', ctypes.c_double), ('dy', ctypes.c_double), ('n_act', ctypes.c_int64),
                ('act_x0', ctypes.c_double), ('act_y0', ctypes.c_double), ('pitch', ctypes.c_double),
                ('x_tilt', ctypes.c_double), ('y_tilt', ctypes.c_double), ('z_tilt', ctypes.c_double),
                ('table_nx', ctypes.c_int64), ('table_ny', ctypes.c_int64), ('table_x0', ctypes.c_double),
                ('table_y0', ctypes.c_double), ('table_dx', ctypes.c_double), ('table_dy', ctypes.c_double),
                ('table', ctypes.c_void_p)]


class PokeItem(ctypes.Structure):
    """hcipy_poke_item of include/hcipy_amd.h."""
    _fields_ = [('plan', ctypes.c_void_p), ('field', ctypes.c_void_p), ('alpha_re', ctypes.c_double),
                ('alpha_im', ctypes.c_double), ('dm', DmGaussianDesc), ('phase_scale', ctypes.c_double),
                ('base_out', ctypes.c_void_p), ('poke_out', ctypes.c_void_p), ('output', ctypes.c_int)]


_c_int64 = ctypes.c_int64
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); every symbol include/hcipy_amd.h declares.
SIGNATURES = {
    'hcipy_abi_version': (ctypes.c_int, []),
    'hcipy_get_last_error': (ctypes.c_char_p, []),
    'hcipy_mft_plan_create': (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _dp, _c_int64, _dp,
                                             _c_int64, _dp, _c_int64, _dp, _c_int64, _vp]),
    'hcipy_mft_plan_destroy': (ctypes.c_int, [_vp]),
    'hcipy_mft_workspace_size': (ctypes.c_int, [_vp, _c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    'hcipy_mft_plan_bytes': (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    'hcipy_mft_plan_set_tiles': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    'hcipy_mft_plan_get_tiles': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    'hcipy_mft_plan_set_parity': (ctypes.c_int, [_vp, ctypes.c_int]),
    'hcipy_mft_plan_get_parity': (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    'hcipy_mft_plan_set_fused': (ctypes.c_int, [_vp, ctypes.c_int]),
    'hcipy_mft_plan_get_fused': (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                                ctypes.POINTER(ctypes.c_uint)]),
    'hcipy_mft_group_create': (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int]),
    'hcipy_mft_group_destroy': (ctypes.c_int, [_vp]),
    'hcipy_mft_execute_grouped': (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    'hcipy_mft_group_release_captures': (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    'hcipy_mft_autotune': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _c_int64, ctypes.c_int, _vp, ctypes.c_size_t,
                                          ctypes.c_int, ctypes.POINTER(ctypes.c_int), _vp]),
    'hcipy_mft_execute': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _c_int64, ctypes.c_double, ctypes.c_double,
                                         _vp, ctypes.c_size_t, _vp]),
    'hcipy_mft_execute_ex': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _c_int64, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_int, _vp, ctypes.c_size_t, _vp]),
    'hcipy_mft_execute_noalloc': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _c_int64, ctypes.c_double,
                                                 ctypes.c_double, ctypes.c_int, _vp]),
    'hcipy_mft_profile': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _c_int64, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, _vp, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                         _vp]),
    'hcipy_mft_profile_pipeline': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _c_int64, ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_int, _vp, ctypes.c_size_t, ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_float), _vp]),
    'hcipy_pw_sub_mul': (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _c_int64, ctypes.c_int, _vp, ctypes.c_int,
                                        _c_int64, _c_int64, _vp]),
    'hcipy_pw_phase': (ctypes.c_int, [_vp, _vp, ctypes.c_int, _c_int64, ctypes.c_double, _vp, ctypes.c_int, _c_int64,
                                      _c_int64, _vp]),
    'hcipy_pw_phase_rows': (ctypes.c_int, [_vp, _vp, ctypes.c_int, _c_int64, _vp, ctypes.c_double, _vp, ctypes.c_int,
                                           _c_int64, _c_int64, _vp]),
    'hcipy_pw_phase_rows_adjoint': (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_double, _vp, _c_int64,
                                                   _c_int64, _c_int64, _vp]),
    'hcipy_pw_intensity': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _c_int64, _vp]),
    'hcipy_pw_power_workspace': (ctypes.c_int, [_c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    'hcipy_pw_power': (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, _c_int64, ctypes.c_double, _c_int64,
                                      _c_int64, _vp, _vp, ctypes.c_size_t, _vp]),
    'hcipy_pw_masked_mean_workspace': (ctypes.c_int, [_c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    'hcipy_pw_masked_mean': (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, _c_int64, _c_int64, _c_int64, _vp,
                                            _vp, _vp, ctypes.c_size_t, _vp]),
    'hcipy_pw_masked_mean_adjoint': (ctypes.c_int, [_vp, ctypes.c_int, _c_int64, _vp, _vp, _vp, ctypes.c_int, _c_int64,
                                                    _c_int64, _vp]),
    'hcipy_pw_phase_adjoint': (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_double, _vp, _c_int64, _c_int64, _c_int64,
                                              _vp]),
    'hcipy_pw_intensity_adjoint': (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, _c_int64, _vp, _c_int64, _c_int64,
                                                  _vp]),
    'hcipy_dm_gaussian_surface': (ctypes.c_int, [_vp, _vp, _vp, _c_int64, _vp]),
    'hcipy_dm_gaussian_surface_adjoint': (ctypes.c_int, [_vp, _vp, _vp, _c_int64, _vp]),
    'hcipy_poke_set_create': (ctypes.c_int, [ctypes.POINTER(_vp), _vp, ctypes.c_int, _vp]),
    'hcipy_poke_set_bytes': (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    'hcipy_poke_set_execute': (ctypes.c_int, [_vp, _vp]),
    'hcipy_poke_set_autotune': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _vp]),
    'hcipy_poke_set_set_tiles': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    'hcipy_poke_set_info': (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double)]),
    'hcipy_poke_set_destroy': (ctypes.c_int, [_vp]),
    'hcipy_inverse_tikhonov_workspace': (ctypes.c_int, [_c_int64, _c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    'hcipy_inverse_tikhonov': (ctypes.c_int, [_vp, _c_int64, _c_int64, ctypes.c_double, _vp,
                                              ctypes.POINTER(ctypes.c_double), _vp, ctypes.c_size_t, _vp]),
    'hcipy_filter_plan_create': (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int, _vp]),
    'hcipy_filter_plan_destroy': (ctypes.c_int, [_vp]),
    'hcipy_filter_set_transfer_function': (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    'hcipy_filter_nearfield_transfer_function': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                                ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                                ctypes.c_double, ctypes.c_int, _vp]),
    'hcipy_nearfield_impulse_response': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                        ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                        ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                        ctypes.c_int, _vp, _vp]),
    'hcipy_pw_field_dot': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                                          ctypes.c_int, _c_int64, _vp, _c_int64, _vp]),
    'hcipy_pw_sub_many': (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_int, _vp, ctypes.c_int, _c_int64, _vp]),
    'hcipy_pw_phase_ramp': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _vp,
                                           _c_int64, _vp]),
    'hcipy_pw_stokes': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                       _vp, _c_int64, _vp]),
    'hcipy_dm_table_surface': (ctypes.c_int, [ctypes.POINTER(DmTableDesc), _vp, _vp, _c_int64, _vp]),
    'hcipy_dm_table_surface_adjoint': (ctypes.c_int, [ctypes.POINTER(DmTableDesc), _vp, _vp, _c_int64, _vp]),
    'hcipy_filter_workspace_size': (ctypes.c_int, [_vp, _c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    'hcipy_filter_execute': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _c_int64, _vp, ctypes.c_size_t, _vp]),
    'hcipy_filter_workspace_size_multi': (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _c_int64,
                                                          ctypes.POINTER(ctypes.c_size_t)]),
    'hcipy_filter_execute_multi': (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _vp,
                                                   ctypes.POINTER(_vp), _c_int64, _vp, ctypes.c_size_t, _vp]),
    'hcipy_filter_workspace_size_matrix': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _c_int64,
                                                           ctypes.POINTER(ctypes.c_size_t)]),
    'hcipy_filter_execute_matrix': (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    _vp, _vp, _c_int64, _vp, ctypes.c_size_t, _vp]),
    'hcipy_zoom_plan_create': (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_int, _vp]),
    'hcipy_nearfield_transfer_function': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                         ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                         ctypes.c_double, ctypes.c_double, ctypes.c_int, _vp, _vp]),
    'hcipy_zoom_plan_create_fft': (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                                  ctypes.c_int, ctypes.c_int, _vp]),
    'hcipy_zoom_plan_destroy': (ctypes.c_int, [_vp]),
    'hcipy_zoom_workspace_size': (ctypes.c_int, [_vp, _c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    'hcipy_zoom_execute': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _c_int64, ctypes.c_double, ctypes.c_double, _vp,
                                          ctypes.c_size_t, _vp]),
    'hcipy_fft_plan_create': (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                             ctypes.c_int, ctypes.c_int, _vp]),
    'hcipy_fft_plan_destroy': (ctypes.c_int, [_vp]),
    'hcipy_fft_workspace_size': (ctypes.c_int, [_vp, _c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    'hcipy_fft_execute': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _c_int64, ctypes.c_double, ctypes.c_double, _vp,
                                         ctypes.c_size_t, _vp]),
}

DTYPE_C64 = 0
DTYPE_C128 = 1
DTYPE_F32 = 2
DTYPE_F64 = 3
_DTYPE_CODES = {torch.complex64: DTYPE_C64, torch.complex128: DTYPE_C128, torch.float32: DTYPE_F32,
                torch.float64: DTYPE_F64}


def dtype_code(t):
    try:
        return _DTYPE_CODES[t.dtype]
    except KeyError:
        raise HcipyError('unsupported dtype %s' % t.dtype)

_lib = None


class HcipyError(RuntimeError):
    pass


def load():
    """Loads the library (once) and declares the signatures."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HcipyError('%s is not built; run `python -m hcipy_amd.build`' % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status, what=''):
    if status != 0:
        msg = load().hcipy_get_last_error().decode(errors='replace')
        raise HcipyError('%s failed (status %d): %s' % (what, status, msg))


def require_device(t):
    """Raises unless `t` lives on a ROCm device (the kernels have no CPU path)."""
    if not isinstance(t, torch.Tensor) or t.device.type != 'cuda':
        raise HcipyError('hcipy_amd kernels need tensors on a ROCm GPU device; got %s'
                         % (t.device if isinstance(t, torch.Tensor) else type(t)))


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def darray(a):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)
