"""Regularised matrix inverses on the device (hcipy/util/matrix_inversion.py).

``inverse_tikhonov`` keeps the reference's signature and result,
V diag(s / (s^2 + (rcond max s)^2)) U^T, computed as (M^T M + lam^2 I)^-1 M^T by the
``hcipy_inverse_tikhonov`` kernels (Gram matrix on the FP64 MFMA stage kernel, power iteration for
the largest singular value, blocked Cholesky, block triangular solves).  Its relative accuracy
is about 1e-16 / rcond^2, the conditioning of the regularised normal equations.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .field import default_device

__all__ = ['inverse_tikhonov']


def inverse_tikhonov(M, rcond=1e-15, svd=None):
    """Tikhonov-regularised inverse of the real matrix ``M`` (m x n) with strength ``rcond`` times
    its largest singular value (matrix_inversion.py:71-100).  Returns an (n x m) float64 device
    tensor.  ``svd`` (a precomputed SVD in the reference) is not used: the kernels need none."""
    if svd is not None:
        raise ValueError('inverse_tikhonov computes no SVD here; pass svd=None')
    t = M.as_subclass(torch.Tensor) if isinstance(M, torch.Tensor) else torch.as_tensor(np.asarray(M))
    if t.is_complex():
        raise ValueError('M must be real')
    if t.dim() != 2:
        raise ValueError('M must be two-dimensional')
    dev = t.device if t.device.type == 'cuda' else default_device()
    t = t.to(device=dev, dtype=torch.float64).contiguous()
    _lib.require_device(t)
    m, n = t.shape
    lib = _lib.load()
    nb = ctypes.c_size_t()
    _lib.check(lib.hcipy_inverse_tikhonov_workspace(m, n, ctypes.byref(nb)), 'hcipy_inverse_tikhonov_workspace')
    ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
    out = torch.empty((n, m), dtype=torch.float64, device=dev)
    smax = ctypes.c_double()
    with torch.cuda.device(dev):
        st = lib.hcipy_inverse_tikhonov(_lib.dptr(t), m, n, float(rcond), _lib.dptr(out), ctypes.byref(smax),
                                        _lib.dptr(ws), ws.numel(), _lib.stream_ptr(dev))
    _lib.check(st, 'hcipy_inverse_tikhonov')
    out.s_max = smax.value
    return out
