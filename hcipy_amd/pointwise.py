"""Pointwise operations on Fields through the library's HBM-bound kernels
(C-ABI ``hcipy_pw_*``, include/hcipy_amd.h)."""
import ctypes

import numpy as np
import torch

from . import _lib
from .field import Field

__all__ = ['multiply', 'sub_mul', 'multiply_phase', 'intensity', 'power_sum', 'total_power']


def _plain(t):
    return t.as_subclass(torch.Tensor) if isinstance(t, Field) else t


def _operand(a, ref):
    """Converts a modulating array (Field / tensor / ndarray / scalar) to a device tensor and
    its per-field stride (0 = broadcast over the batch)."""
    if a is None:
        return None, 0
    if not isinstance(a, torch.Tensor):
        a = torch.as_tensor(np.asarray(a), device=ref.device)
    a = _plain(a)
    if a.dtype not in (torch.complex64, torch.complex128, torch.float32, torch.float64):
        a = a.to(torch.float64)
    a = a.to(ref.device).contiguous()
    n = ref.shape[-1]
    if a.numel() == 1:
        a = a.reshape(1).expand(n).contiguous()
    if a.shape[-1] != n:
        raise ValueError('operand has %d samples, field has %d' % (a.shape[-1], n))
    stride = 0 if a.numel() == n else n
    if stride and a.numel() != ref.numel():
        raise ValueError('operand batch shape does not match the field')
    return a, stride


def sub_mul(x, y=None, a=None, conjugate=False, out=None):
    """(x - y) * a (conj(a) if `conjugate`) in one pass; returns a Field on x's grid."""
    grid = getattr(x, 'grid', None)
    xt = _plain(x).contiguous()
    _lib.require_device(xt)
    yt = None
    if y is not None:
        yt = _plain(y).to(xt.dtype).contiguous()
    at, stride = _operand(a, xt)
    o = torch.empty_like(xt) if out is None else _plain(out)
    n = xt.shape[-1]
    batch = xt.numel() // n if n else 0
    st = _lib.load().hcipy_pw_sub_mul(_lib.dptr(xt), _lib.dptr(yt) if yt is not None else None,
                                      _lib.dptr(at) if at is not None else None,
                                      _lib.dtype_code(at) if at is not None else 0, stride, int(bool(conjugate)),
                                      _lib.dptr(o), _lib.dtype_code(xt), n, batch, _lib.stream_ptr(xt.device))
    _lib.check(st, 'hcipy_pw_sub_mul')
    return Field(o, grid) if grid is not None else o


def multiply(x, a, conjugate=False):
    """x * a (or x * conj(a))."""
    return sub_mul(x, None, a, conjugate)


def multiply_phase(x, phi, scale=1.0):
    """x * exp(i * scale * phi), phase in FP64."""
    grid = getattr(x, 'grid', None)
    xt = _plain(x).contiguous()
    _lib.require_device(xt)
    pt, stride = _operand(phi, xt)
    if pt.is_complex():
        raise ValueError('phase must be real')
    o = torch.empty_like(xt)
    n = xt.shape[-1]
    batch = xt.numel() // n if n else 0
    st = _lib.load().hcipy_pw_phase(_lib.dptr(xt), _lib.dptr(pt), _lib.dtype_code(pt), stride, float(scale),
                                    _lib.dptr(o), _lib.dtype_code(xt), n, batch, _lib.stream_ptr(xt.device))
    _lib.check(st, 'hcipy_pw_phase')
    return Field(o, grid) if grid is not None else o


def intensity(x):
    """|x|^2 elementwise (float32 for complex64, float64 for complex128)."""
    grid = getattr(x, 'grid', None)
    xt = _plain(x).contiguous()
    _lib.require_device(xt)
    rdt = torch.float32 if xt.dtype == torch.complex64 else torch.float64
    if not xt.is_complex():
        xt = xt.to(torch.complex64 if xt.dtype == torch.float32 else torch.complex128)
    o = torch.empty(xt.shape, dtype=rdt, device=xt.device)
    st = _lib.load().hcipy_pw_intensity(_lib.dptr(xt), _lib.dtype_code(xt), _lib.dptr(o), xt.numel(),
                                        _lib.stream_ptr(xt.device))
    _lib.check(st, 'hcipy_pw_intensity')
    return Field(o, grid) if grid is not None else o


def power_sum(x, weight=1.0, mask=None):
    """FP64 tensor of shape lead: weight * sum(mask * |x|^2) over the last axis."""
    xt = _plain(x).contiguous()
    _lib.require_device(xt)
    if not xt.is_complex():
        xt = xt.to(torch.complex64 if xt.dtype == torch.float32 else torch.complex128)
    mt, stride = _operand(mask, xt)
    n = xt.shape[-1]
    lead = tuple(xt.shape[:-1])
    batch = int(np.prod(lead)) if lead else 1
    lib = _lib.load()
    nb = ctypes.c_size_t()
    _lib.check(lib.hcipy_pw_power_workspace(batch, ctypes.byref(nb)), 'hcipy_pw_power_workspace')
    ws = torch.empty(max(nb.value, 8), dtype=torch.uint8, device=xt.device)
    out = torch.empty(batch, dtype=torch.float64, device=xt.device)
    st = lib.hcipy_pw_power(_lib.dptr(xt), _lib.dtype_code(xt), _lib.dptr(mt) if mt is not None else None,
                            _lib.dtype_code(mt) if mt is not None else 0, stride, float(weight), n, batch,
                            _lib.dptr(out), _lib.dptr(ws), ws.numel(), _lib.stream_ptr(xt.device))
    _lib.check(st, 'hcipy_pw_power')
    return out.reshape(lead)


def total_power(x, weights):
    """sum |x|^2 * weights as a Python float (scalar weights use the fused FP64 reduction)."""
    w = np.asarray(weights)
    if w.size == 1:
        return float(power_sum(x, float(w.ravel()[0])).sum().item())
    return float(power_sum(x, 1.0, mask=torch.as_tensor(w, device=x.device)).sum().item())
