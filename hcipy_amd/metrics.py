"""Focal-plane metrics of the dark-hole loop (hcipy/metrics/contrast.py:1-76).

The region-of-interest mean is an FP64 reduction on the device (``hcipy_pw_masked_mean``);
the Strehl helpers index the peak of the reference image.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .field import Field, default_device

__all__ = ['get_mean_intensity_in_roi', 'get_mean_raw_contrast', 'get_strehl_from_focal', 'get_strehl_from_pupil']


def _device_tensor(x):
    t = x.as_subclass(torch.Tensor) if isinstance(x, Field) else x
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(np.asarray(x), device=default_device())
    return t


def get_mean_intensity_in_roi(img, mask):
    """Mean of ``img`` over the samples where ``mask`` is non-zero (contrast.py:37-52), summed in
    FP64 on the device.  Leading dimensions of ``img`` are a batch (one mean each); ``mask`` is
    broadcast when it has one image's size.  Returns a float for a single image, a tensor for a
    batch or when ``img`` requires grad (the backward spreads g / count over the region)."""
    t = _device_tensor(img)
    _lib.require_device(t)
    if t.is_complex():
        raise ValueError('img must be real (an intensity image)')
    if t.dtype not in (torch.float32, torch.float64):
        t = t.to(torch.float64)
    t = t.contiguous()
    m = _device_tensor(mask).to(t.device)
    if m.dtype == torch.bool:
        m = m.to(torch.float32)
    elif m.is_complex() or m.dtype not in (torch.float32, torch.float64):
        m = m.to(torch.float64)
    m = m.contiguous()
    n = t.shape[-1]
    lead = tuple(t.shape[:-1])
    batch = int(np.prod(lead)) if lead else 1
    if m.numel() not in (n, n * batch):
        raise ValueError('mask must have the size of one image or of the whole batch')
    stride = 0 if m.numel() == n else n
    if t.requires_grad and torch.is_grad_enabled():
        out = _RoiMeanFn.apply(t, m, stride)
        return out.reshape(lead)
    out, _ = _roi_mean_plain(t, m, stride, with_count=False)
    if not lead:
        return float(out[0].item())
    return out.reshape(lead)


def _roi_mean_plain(t, m, stride, with_count):
    n = t.shape[-1]
    batch = t.numel() // n if n else 0
    ws_n = _workspace_bytes(batch)
    ws = torch.empty(max(ws_n, 8), dtype=torch.uint8, device=t.device)
    out = torch.empty(max(batch, 1), dtype=torch.float64, device=t.device)
    cnt = torch.empty(max(batch, 1), dtype=torch.float64, device=t.device) if with_count else None
    st = _lib.load().hcipy_pw_masked_mean(_lib.dptr(t), _lib.dtype_code(t), _lib.dptr(m), _lib.dtype_code(m), stride,
                                          n, batch, _lib.dptr(out), _lib.dptr(cnt) if cnt is not None else None,
                                          _lib.dptr(ws), ws_n, _lib.stream_ptr(t.device))
    _lib.check(st, 'hcipy_pw_masked_mean')
    return out, cnt


class _RoiMeanFn(torch.autograd.Function):
    """Region-of-interest mean with the adjoint g / count on the selected samples."""

    @staticmethod
    def forward(ctx, t, m, stride):
        out, cnt = _roi_mean_plain(t, m, stride, with_count=True)
        ctx.save_for_backward(m, cnt)
        ctx.stride, ctx.shape, ctx.dtype = stride, t.shape, t.dtype
        return out

    @staticmethod
    def backward(ctx, g):
        m, cnt = ctx.saved_tensors
        g = g.to(torch.float64).contiguous()
        n = ctx.shape[-1]
        batch = int(np.prod(ctx.shape[:-1])) if len(ctx.shape) > 1 else 1
        gt = torch.empty(ctx.shape, dtype=ctx.dtype, device=g.device)
        st = _lib.load().hcipy_pw_masked_mean_adjoint(_lib.dptr(m), _lib.dtype_code(m), ctx.stride, _lib.dptr(g),
                                                      _lib.dptr(cnt), _lib.dptr(gt), _lib.dtype_code(gt), n, batch,
                                                      _lib.stream_ptr(g.device))
        _lib.check(st, 'hcipy_pw_masked_mean_adjoint')
        return gt, None, None


def _workspace_bytes(batch):
    n = ctypes.c_size_t()
    _lib.check(_lib.load().hcipy_pw_masked_mean_workspace(batch, ctypes.byref(n)), 'hcipy_pw_masked_mean_workspace')
    return n.value


def get_strehl_from_focal(img, ref_img):
    """img at the peak of ref_img, over that peak (contrast.py:3-17)."""
    i = _device_tensor(img)
    r = _device_tensor(ref_img)
    k = int(torch.argmax(r).item())
    return float((i.reshape(-1)[k] / r.max()).item())


def get_strehl_from_pupil(aperture, ref_aperture):
    """|sum(aperture) / sum(ref_aperture)|^2 (contrast.py:19-35)."""
    a = _device_tensor(aperture).to(torch.complex128)
    r = _device_tensor(ref_aperture).to(torch.complex128)
    return float((torch.abs(a.sum() / r.sum()) ** 2).item())


def get_mean_raw_contrast(img, mask, ref_img):
    """Mean intensity in the region over the Strehl ratio (contrast.py:54-76)."""
    return get_mean_intensity_in_roi(img, mask) / get_strehl_from_focal(img, ref_img)
