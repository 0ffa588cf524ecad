"""Pointwise operations on Fields through the library's HBM-bound kernels
(C-ABI ``hcipy_pw_*``, include/hcipy_amd.h).

Every operation is differentiable: when an input requires grad (and grad mode is on) it runs
inside a torch.autograd.Function whose backward is the adjoint kernel (conjugate multiply,
``hcipy_pw_phase_adjoint``, ``hcipy_pw_intensity_adjoint``), with PyTorch's convention for complex
gradients.  The modulating array of ``sub_mul``/``multiply`` is treated as a constant."""
import ctypes

import numpy as np
import torch

from . import _lib
from .field import Field

__all__ = ['multiply', 'sub_mul', 'multiply_phase', 'intensity', 'power_sum', 'total_power', 'needs_grad']


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


def needs_grad(*ts):
    """True when autograd must record an operation on these inputs."""
    return torch.is_grad_enabled() and any(isinstance(t, torch.Tensor) and t.requires_grad for t in ts)


def _sub_mul_raw(xt, yt, at, stride, conjugate, o=None):
    """(x - y) a with x or y None (x None: zero)."""
    ref = xt if xt is not None else yt
    o = torch.empty_like(ref) if o is None else o
    n = ref.shape[-1]
    batch = ref.numel() // n if n else 0
    st = _lib.load().hcipy_pw_sub_mul(_lib.dptr(xt) if xt is not None else None,
                                      _lib.dptr(yt) if yt is not None else None,
                                      _lib.dptr(at) if at is not None else None,
                                      _lib.dtype_code(at) if at is not None else 0, stride, int(bool(conjugate)),
                                      _lib.dptr(o), _lib.dtype_code(ref), n, batch, _lib.stream_ptr(ref.device))
    _lib.check(st, 'hcipy_pw_sub_mul')
    return o


class _SubMulFn(torch.autograd.Function):
    """(x - y) a; adjoint g conj(a) for x and -g conj(a) for y."""

    @staticmethod
    def forward(ctx, xt, yt, at, stride, conjugate):
        ctx.save_for_backward(at)
        ctx.stride, ctx.conjugate, ctx.has_y = stride, conjugate, yt is not None
        return _sub_mul_raw(xt, yt, at, stride, conjugate)

    @staticmethod
    def backward(ctx, g):
        (at,) = ctx.saved_tensors
        g = g.contiguous()
        gx = _sub_mul_raw(g, None, at, ctx.stride, not ctx.conjugate) if at is not None else g
        # -g conj(a) from the kernel with x absent (exactly the negation of gx)
        gy = _sub_mul_raw(None, g, at, ctx.stride, not ctx.conjugate) if ctx.has_y and ctx.needs_input_grad[1] else None
        return (gx if ctx.needs_input_grad[0] else None), gy, None, None, None


def sub_mul(x, y=None, a=None, conjugate=False, out=None):
    """(x - y) * a (conj(a) if `conjugate`) in one pass; returns a Field on x's grid."""
    grid = getattr(x, 'grid', None)
    xt = _plain(x).contiguous()
    _lib.require_device(xt)
    yt = None
    if y is not None:
        yt = _plain(y).to(xt.dtype).contiguous()
    at, stride = _operand(a, xt)
    if needs_grad(xt, yt):
        if out is not None:
            raise ValueError('out= is not supported while recording gradients')
        o = _SubMulFn.apply(xt, yt, at, stride, bool(conjugate))
    else:
        o = _sub_mul_raw(xt, yt, at, stride, conjugate, None if out is None else _plain(out))
    return Field(o, grid) if grid is not None else o


def multiply(x, a, conjugate=False):
    """x * a (or x * conj(a))."""
    return sub_mul(x, None, a, conjugate)


def _phase_raw(xt, pt, stride, scale, rows=None):
    o = torch.empty_like(xt)
    n = xt.shape[-1]
    batch = xt.numel() // n if n else 0
    st = _lib.load().hcipy_pw_phase_rows(_lib.dptr(xt), _lib.dptr(pt), _lib.dtype_code(pt), stride,
                                         _lib.dptr(rows) if rows is not None else None, float(scale), _lib.dptr(o),
                                         _lib.dtype_code(xt), n, batch, _lib.stream_ptr(xt.device))
    _lib.check(st, 'hcipy_pw_phase_rows')
    return o


class _PhaseFn(torch.autograd.Function):
    """y = x exp(i s phi); adjoints g exp(-i s phi) for x and -s Im(conj(g) y) for phi."""

    @staticmethod
    def forward(ctx, xt, pt, stride, scale, rows):
        y = _phase_raw(xt, pt, stride, scale, rows)
        ctx.save_for_backward(pt, y)
        ctx.stride, ctx.scale, ctx.rows = stride, scale, rows
        return y

    @staticmethod
    def backward(ctx, g):
        pt, y = ctx.saved_tensors
        g = g.contiguous()
        gx = _phase_raw(g, pt, ctx.stride, -ctx.scale, ctx.rows) if ctx.needs_input_grad[0] else None
        gp = None
        if ctx.needs_input_grad[1]:
            n = y.shape[-1]
            batch = y.numel() // n if n else 0
            gp = torch.empty(pt.shape, dtype=torch.float64, device=y.device)
            st = _lib.load().hcipy_pw_phase_rows_adjoint(_lib.dptr(y), _lib.dptr(g), _lib.dtype_code(y),
                                                         _lib.dptr(ctx.rows) if ctx.rows is not None else None,
                                                         float(ctx.scale), _lib.dptr(gp), ctx.stride, n, batch,
                                                         _lib.stream_ptr(y.device))
            _lib.check(st, 'hcipy_pw_phase_rows_adjoint')
            gp = gp.to(pt.dtype)
        return gx, gp, None, None, None


def multiply_phase(x, phi, scale=1.0):
    """x * exp(i * scale * phi), phase in FP64.  ``scale`` may also be a float64 device tensor
    with one scale per field of the batch (the first dimension of ``x``; for example 4 pi / lambda
    of each wavelength of a stacked broadband field): one launch for all of them."""
    grid = getattr(x, 'grid', None)
    xt = _plain(x).contiguous()
    _lib.require_device(xt)
    pt, stride = _operand(phi, xt)
    if pt.is_complex():
        raise ValueError('phase must be real')
    rows = None
    if isinstance(scale, torch.Tensor):
        n = xt.shape[-1]
        batch = xt.numel() // n if n else 0
        rows = scale
        if rows.dtype != torch.float64 or rows.device != xt.device or rows.numel() != batch or not rows.is_contiguous():
            raise ValueError('per-field scales must be a contiguous float64 tensor of %d values on %s' %
                             (batch, xt.device))
        if rows.requires_grad:
            raise ValueError('per-field scales are constants')
        scale = 1.0
    if needs_grad(xt, pt):
        o = _PhaseFn.apply(xt, pt, stride, float(scale), rows)
    else:
        o = _phase_raw(xt, pt, stride, scale, rows)
    return Field(o, grid) if grid is not None else o


def intensity(x):
    """|x|^2 elementwise (float32 for complex64, float64 for complex128)."""
    grid = getattr(x, 'grid', None)
    xt = _plain(x).contiguous()
    _lib.require_device(xt)
    if not xt.is_complex():
        xt = xt.to(torch.complex64 if xt.dtype == torch.float32 else torch.complex128)
    o = _IntensityFn.apply(xt) if needs_grad(xt) else _intensity_raw(xt)
    return Field(o, grid) if grid is not None else o


def _intensity_raw(xt):
    rdt = torch.float32 if xt.dtype == torch.complex64 else torch.float64
    o = torch.empty(xt.shape, dtype=rdt, device=xt.device)
    st = _lib.load().hcipy_pw_intensity(_lib.dptr(xt), _lib.dtype_code(xt), _lib.dptr(o), xt.numel(),
                                        _lib.stream_ptr(xt.device))
    _lib.check(st, 'hcipy_pw_intensity')
    return o


class _IntensityFn(torch.autograd.Function):
    """I = |x|^2; adjoint 2 g x."""

    @staticmethod
    def forward(ctx, xt):
        ctx.save_for_backward(xt)
        return _intensity_raw(xt)

    @staticmethod
    def backward(ctx, g):
        (xt,) = ctx.saved_tensors
        g = g.contiguous()
        n = xt.shape[-1]
        batch = xt.numel() // n if n else 0
        gx = torch.empty_like(xt)
        st = _lib.load().hcipy_pw_intensity_adjoint(_lib.dptr(xt), _lib.dtype_code(xt), _lib.dptr(g), _lib.dtype_code(g),
                                                    n, _lib.dptr(gx), n, batch, _lib.stream_ptr(xt.device))
        _lib.check(st, 'hcipy_pw_intensity_adjoint')
        return gx


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


def real_sum(x):
    """FP64 sum of a real array (the deterministic two-pass reduction of ``hcipy_pw_masked_mean``
    without a mask: mean times count), as a Python float."""
    t = _plain(x).contiguous()
    _lib.require_device(t)
    if t.dtype not in (torch.float32, torch.float64):
        raise ValueError('real_sum needs a real array')
    t = t.reshape(1, -1)
    n = t.shape[-1]
    if n == 0:
        return 0.0
    lib = _lib.load()
    nb = ctypes.c_size_t()
    _lib.check(lib.hcipy_pw_masked_mean_workspace(1, ctypes.byref(nb)), 'hcipy_pw_masked_mean_workspace')
    ws = torch.empty(max(nb.value, 8), dtype=torch.uint8, device=t.device)
    out = torch.empty(1, dtype=torch.float64, device=t.device)
    cnt = torch.empty(1, dtype=torch.float64, device=t.device)
    st = lib.hcipy_pw_masked_mean(_lib.dptr(t), _lib.dtype_code(t), None, 0, 0, n, 1, _lib.dptr(out), _lib.dptr(cnt),
                                  _lib.dptr(ws), ws.numel(), _lib.stream_ptr(t.device))
    _lib.check(st, 'hcipy_pw_masked_mean')
    return float(out.item()) * float(cnt.item())


def stokes(E, input_stokes_vector=None, component=-1):
    """Stokes parameters of a polarized field (wavefront.py:112-235): E is a Jones-vector field
    (2, n), or the (2, 2, n) Jones-matrix field of a partially polarized wavefront acting on
    ``input_stokes_vector``.  ``component`` 0..3 returns I, Q, U or V (n,); -1 all four (4, n).
    FP64 arithmetic, float32 results for complex64 fields."""
    grid = getattr(E, 'grid', None)
    et = _plain(E).contiguous()
    _lib.require_device(et)
    vector = input_stokes_vector is None
    if tuple(et.shape[:-1]) != ((2,) if vector else (2, 2)):
        raise ValueError('expected a %s field, got shape %s' % ('(2, n) Jones-vector' if vector else '(2, 2, n) Jones-matrix',
                                                                 tuple(et.shape)))
    if not et.is_complex():
        et = et.to(torch.complex128)
    n = et.shape[-1]
    rdt = torch.float32 if et.dtype == torch.complex64 else torch.float64
    out = torch.empty((n,) if component >= 0 else (4, n), dtype=rdt, device=et.device)
    sv = None
    if not vector:
        sv = (ctypes.c_double * 4)(*[float(v) for v in np.asarray(input_stokes_vector, dtype=float).reshape(4)])
    st = _lib.load().hcipy_pw_stokes(_lib.dptr(et), _lib.dtype_code(et), int(vector), sv, int(component), _lib.dptr(out),
                                     n, _lib.stream_ptr(et.device))
    _lib.check(st, 'hcipy_pw_stokes')
    return Field(out, grid) if grid is not None else out


def sub_many(x, ys):
    """x - ys[0] - ys[1] - ... in one pass (the rounding of the chained subtractions); the
    operands share x's shape and dtype."""
    grid = getattr(x, 'grid', None)
    xt = _plain(x).contiguous()
    _lib.require_device(xt)
    out = torch.empty_like(xt)
    done = 0
    src = xt
    ys = [_plain(y).contiguous() for y in ys]
    for y in ys:
        if y.shape != xt.shape or y.dtype != xt.dtype:
            raise ValueError('sub_many operands must match the field in shape and dtype')
    while True:
        chunk = ys[done:done + 16]
        arr = (ctypes.c_void_p * max(1, len(chunk)))(*[_lib.dptr(y) for y in chunk])
        st = _lib.load().hcipy_pw_sub_many(_lib.dptr(src), arr, len(chunk), _lib.dptr(out), _lib.dtype_code(xt),
                                           xt.numel(), _lib.stream_ptr(xt.device))
        _lib.check(st, 'hcipy_pw_sub_many')
        done += len(chunk)
        src = out
        if done >= len(ys):
            break
    return Field(out, grid) if grid is not None else out


def phase_ramp(E, grid, ax, ay=0.0, out=None):
    """E * exp(i (ax x + ay y)) on the regular grid of E (any leading dimensions), the phase in
    FP64 on the device; ``out`` may be E itself (in place)."""
    et = _plain(E).contiguous()
    _lib.require_device(et)
    if not et.is_complex():
        et = et.to(torch.complex128)
    if not grid.is_regular:
        raise ValueError('phase_ramp needs a regular grid')
    n = et.shape[-1]
    if n != grid.size:
        raise ValueError('field size %d does not match the grid size %d' % (n, grid.size))
    batch = int(np.prod(et.shape[:-1])) if et.dim() > 1 else 1
    dims = [int(d) for d in grid.dims]
    delta = [float(d) for d in grid.delta]
    zero = [float(z) for z in grid.zero]
    nx, ny = dims[0], (dims[1] if grid.ndim == 2 else 1)
    zy, dy = (zero[1], delta[1]) if grid.ndim == 2 else (0.0, 0.0)
    o = et if out is not None and _plain(out).data_ptr() == et.data_ptr() else torch.empty_like(et)
    st = _lib.load().hcipy_pw_phase_ramp(_lib.dptr(et), _lib.dtype_code(et), nx, ny, zero[0], delta[0], zy, dy,
                                         float(ax), float(ay), _lib.dptr(o), batch, _lib.stream_ptr(et.device))
    _lib.check(st, 'hcipy_pw_phase_ramp')
    return o


def _field_dot_raw(jt, et, adjoint):
    p, q = int(jt.shape[0]), int(jt.shape[1])
    n = et.shape[-1]
    rest = tuple(et.shape[1:-1])
    K = int(np.prod(rest)) if rest else 1
    o = torch.empty(((q if adjoint else p),) + rest + (n,), dtype=et.dtype, device=et.device)
    st = _lib.load().hcipy_pw_field_dot(_lib.dptr(jt), _lib.dtype_code(jt), p, q, int(bool(adjoint)), _lib.dptr(et),
                                        _lib.dtype_code(et), K, _lib.dptr(o), n, _lib.stream_ptr(et.device))
    _lib.check(st, 'hcipy_pw_field_dot')
    return o


class _FieldDotFn(torch.autograd.Function):
    """out = J E per pixel (J constant); the gradient with respect to E is J^H g."""

    @staticmethod
    def forward(ctx, jt, et, adjoint):
        ctx.save_for_backward(jt)
        ctx.adjoint = adjoint
        return _field_dot_raw(jt, et, adjoint)

    @staticmethod
    def backward(ctx, g):
        (jt,) = ctx.saved_tensors
        return None, _field_dot_raw(jt, g.contiguous(), not ctx.adjoint), None


def field_dot(J, E, adjoint=False):
    """Per-pixel matrix product of a (p, q, n) matrix field J with a field E of shape (q, ..., n)
    (field_dot of hcipy/field/field.py; ``adjoint`` applies J^H), returning (p, ..., n)."""
    grid = getattr(E, 'grid', None)
    et = _plain(E).contiguous()
    _lib.require_device(et)
    jt = _plain(J)
    if not isinstance(jt, torch.Tensor):
        jt = torch.as_tensor(np.asarray(J), device=et.device)
    jt = jt.to(et.device)
    if not jt.is_complex():
        jt = jt.to(torch.complex128)
    jt = jt.contiguous()
    if jt.dim() != 3 or jt.shape[-1] != et.shape[-1]:
        raise ValueError('J must be a (p, q, n) matrix field on the same grid as E')
    if et.shape[0] != (jt.shape[0] if adjoint else jt.shape[1]):
        raise ValueError('E has %d components along its first dimension, J needs %d'
                         % (et.shape[0], jt.shape[0] if adjoint else jt.shape[1]))
    if needs_grad(et):
        o = _FieldDotFn.apply(jt, et, bool(adjoint))
    else:
        o = _field_dot_raw(jt, et, adjoint)
    return Field(o, grid) if grid is not None else o


__all__ += ['real_sum', 'stokes', 'field_dot', 'phase_ramp', 'sub_many']
