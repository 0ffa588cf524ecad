"""Fourier transforms: the FourierTransform protocol, FFT-grid geometry and the MFMA matrix
Fourier transform.

Protocol (hcipy/fourier/fourier_transform.py:17-183): ``forward(field) -> Field`` on
``output_grid``; ``backward(field) -> Field`` on ``input_grid``; classmethods
``check_if_supported`` (raises ValueError), ``is_supported`` and ``compute_complexity``.
"""
import ctypes
import math
import time
from collections.abc import Iterable
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .config import Configuration
from .field import CartesianGrid, Field, RegularCoords, field_dtype

__all__ = [
    'ComputationalComplexity', 'FourierTransform', 'MatrixFourierTransform', 'MftGroup', 'make_fft_grid',
    'get_fft_parameters', 'is_fft_grid', 'complex_dtype_for',
]


@dataclass(frozen=True)
class ComputationalComplexity:
    """hcipy/fourier/fourier_transform.py:7-15."""
    num_multiplications: int = 0
    num_additions: int = 0
    expected_execution_time: float = 0

    @property
    def num_operations(self):
        return self.num_multiplications + self.num_additions


def complex_dtype_for(dtype):
    """float32/complex64 -> complex64, everything else -> complex128
    (hcipy/fourier/fourier_transform.py:323-343, hcipy/optics/wavefront.py:77-88)."""
    return torch.complex64 if dtype in (torch.float32, torch.complex64) else torch.complex128


def _as_field_tensor(field, grid):
    """Returns (tensor on device, complex dtype) for a Field / tensor / array input."""
    if not isinstance(field, torch.Tensor):
        field = Field(field, grid)
    t = field.as_subclass(torch.Tensor) if isinstance(field, Field) else field
    cdt = complex_dtype_for(t.dtype)
    _lib.require_device(t)
    return t.to(cdt).contiguous(), cdt


class FourierTransform:
    """Base class (hcipy/fourier/fourier_transform.py:17-183)."""

    def forward(self, field):
        raise NotImplementedError()

    def backward(self, field):
        raise NotImplementedError()

    @classmethod
    def is_supported(cls, input_grid, output_grid):
        try:
            cls.check_if_supported(input_grid, output_grid)
        except ValueError:
            return False
        return True

    @classmethod
    def check_if_supported(cls, input_grid, output_grid):
        raise NotImplementedError()

    @classmethod
    def compute_complexity(cls, input_grid, output_grid):
        raise NotImplementedError()

    @classmethod
    def _predict_execution_time(cls, num_operations, prediction_coefficients):
        """t_ms = e^a (Gop)^b + e^c, returned in seconds (fourier_transform.py:159-183)."""
        a = prediction_coefficients['a']
        b = prediction_coefficients['b']
        c = prediction_coefficients['c']
        return (math.exp(a) * (num_operations / 1e9) ** b + math.exp(c)) * 1e-3


# ----------------------------------------------------------------------------------------------
# FFT grid geometry (hcipy/fourier/fast_fourier_transform.py:23-130)


def make_fft_grid(input_grid, q=1, fov=1, shift=0):
    """Grid of an FFT of `input_grid` with zero padding q, crop fov and shift
    (fast_fourier_transform.py:23-62)."""
    nd = input_grid.ndim
    if not isinstance(q, Iterable):
        q = (q,) * nd
    if not isinstance(fov, Iterable):
        fov = (fov,) * nd
    if not isinstance(shift, Iterable):
        shift = (shift,) * nd
    if not input_grid.is_regular:
        raise ValueError('The input_grid must be regular.')
    if not input_grid.is_('cartesian'):
        raise ValueError('The input_grid must be cartesian.')
    dims_in = [int(d) for d in input_grid.dims]
    q = tuple(round(q_i * d_in) / d_in for q_i, d_in in zip(q, dims_in))
    delta = tuple(float((2 * math.pi / (d_in * s_in)) / q_i) for d_in, s_in, q_i in zip(input_grid.delta, dims_in, q))
    dims = tuple(int(d_in * f_i * q_i) for d_in, f_i, q_i in zip(dims_in, fov, q))
    zero = tuple(float(d_i * (-dim / 2 + (dim % 2) * 0.5) + s_i) for d_i, dim, s_i in zip(delta, dims, shift))
    return CartesianGrid(RegularCoords(delta, dims, zero))


def get_fft_parameters(fft_grid, input_grid):
    """Recovers (q, fov, shift) of an FFT grid or raises ValueError
    (fast_fourier_transform.py:64-130)."""
    if not input_grid.is_regular:
        raise ValueError('The input grid must be regular to reconstruct an fft grid.')
    if not fft_grid.is_regular:
        raise ValueError('The fft grid is not regular and therefore cannot be an fft grid.')
    if input_grid.ndim != fft_grid.ndim:
        raise ValueError('The fft grid does not have the same number of dimensions as input_grid.')
    dims_in = [int(d) for d in input_grid.dims]
    fdims = [int(d) for d in fft_grid.dims]
    q = tuple((2 * math.pi / (d_in * s_in)) / f_delta for d_in, s_in, f_delta in zip(input_grid.delta, dims_in, fft_grid.delta))
    if any(q_i < 1 for q_i in q):
        raise ValueError(f'fft_grid is not an FFT grid of input_grid: q of {q} would be < 1.')
    zp = tuple(q_i * d_in for q_i, d_in in zip(q, dims_in))
    if any(abs(zn - round(zn)) > 1e-10 for zn in zp):
        raise ValueError(f'fft_grid is not an FFT grid of input_grid: q of {q} does not correspond to an integer zeropadding.')
    fov = tuple(f_dim / zp_dim for f_dim, zp_dim in zip(fdims, zp))
    if any(n > int(zn + 0.5) for n, zn in zip(fdims, zp)):
        raise ValueError(f'fft_grid is not an FFT grid of input_grid: fov of {fov} would be > 1 .')
    dummy = make_fft_grid(input_grid, q, fov)
    wrong = tuple(int(n) != m for n, m in zip(dummy.dims, fdims))
    fov = tuple((f_dim + 0.5) / (d_in * q_i) if w else f for w, f, f_dim, d_in, q_i in zip(wrong, fov, fdims, dims_in, q))
    shift = tuple(f_zero - f_delta * (-f_dim / 2 + (f_dim % 2) * 0.5) for f_zero, f_delta, f_dim in zip(fft_grid.zero, fft_grid.delta, fdims))
    return q, fov, shift


def is_fft_grid(grid, input_grid):
    try:
        get_fft_parameters(grid, input_grid)
    except ValueError:
        return False
    return True


# ----------------------------------------------------------------------------------------------
# Matrix Fourier transform


class _MftPlan:
    """Owns one C-ABI plan (device Fourier matrices in one operand format)."""

    def __init__(self, input_grid, output_grid, precision, device):
        lib = _lib.load()
        self.device = device
        self.precision = precision
        with torch.cuda.device(device):
            sep_in = input_grid.separated_coords
            sep_out = output_grid.separated_coords
            x, px = _lib.darray(sep_in[0])
            u, pu = _lib.darray(sep_out[0])
            if input_grid.ndim == 2:
                y, py = _lib.darray(sep_in[1])
                v, pv = _lib.darray(sep_out[1])
                ny, nv = len(y), len(v)
            else:
                py = pv = None
                ny = nv = 1
            h = ctypes.c_void_p()
            st = lib.hcipy_mft_plan_create(ctypes.byref(h), input_grid.ndim, _lib.PRECISIONS[precision], px, len(x),
                                           py, ny, pu, len(u), pv, nv, _lib.stream_ptr(device))
            _lib.check(st, 'hcipy_mft_plan_create')
            self.handle = h
            n = ctypes.c_size_t()
            _lib.check(lib.hcipy_mft_plan_bytes(h, ctypes.byref(n)), 'hcipy_mft_plan_bytes')
            self.nbytes = n.value

    def workspace_bytes(self, batch):
        n = ctypes.c_size_t()
        _lib.check(_lib.load().hcipy_mft_workspace_size(self.handle, batch, ctypes.byref(n)), 'hcipy_mft_workspace_size')
        return n.value

    def set_tiles(self, direction, stage, tile_cfg):
        """Pins the workgroup tile shape of one stage (``_lib.TILES_*``; ``TILES_AUTO`` resets)."""
        _lib.check(_lib.load().hcipy_mft_plan_set_tiles(self.handle, direction, stage, tile_cfg),
                   'hcipy_mft_plan_set_tiles')

    def get_tiles(self, direction, stage):
        v = ctypes.c_int()
        _lib.check(_lib.load().hcipy_mft_plan_get_tiles(self.handle, direction, stage, ctypes.byref(v)),
                   'hcipy_mft_plan_get_tiles')
        return v.value

    @property
    def parity_available(self):
        """True for 2-D split-precision plans between centred grids (even sizes, axes symmetric
        about 0): their transforms can run in the parity form (real half-size cosine and sine
        products of the field's four parity components)."""
        a = ctypes.c_int()
        _lib.check(_lib.load().hcipy_mft_plan_get_parity(self.handle, ctypes.byref(a), None), 'hcipy_mft_plan_get_parity')
        return bool(a.value)

    @property
    def parity(self):
        """Whether single transforms run in the parity form (on by default where available)."""
        e = ctypes.c_int()
        _lib.check(_lib.load().hcipy_mft_plan_get_parity(self.handle, None, ctypes.byref(e)), 'hcipy_mft_plan_get_parity')
        return bool(e.value)

    @parity.setter
    def parity(self, enable):
        _lib.check(_lib.load().hcipy_mft_plan_set_parity(self.handle, 1 if enable else 0), 'hcipy_mft_plan_set_parity')

    @property
    def fused_available(self):
        """True for split-FP16 plans with the parity form available: a single transform can then run
        as one persistent launch (parity split, stage 1 and stage 2 as tasks of one kernel)."""
        a = ctypes.c_int()
        _lib.check(_lib.load().hcipy_mft_plan_get_fused(self.handle, ctypes.byref(a), None, None),
                   'hcipy_mft_plan_get_fused')
        return bool(a.value)

    @property
    def fused(self):
        """Whether single transforms in the parity form run as one persistent launch (on by default
        where available, for the default tile shapes)."""
        e = ctypes.c_int()
        _lib.check(_lib.load().hcipy_mft_plan_get_fused(self.handle, None, ctypes.byref(e), None),
                   'hcipy_mft_plan_get_fused')
        return bool(e.value)

    @fused.setter
    def fused(self, enable):
        _lib.check(_lib.load().hcipy_mft_plan_set_fused(self.handle, 1 if enable else 0), 'hcipy_mft_plan_set_fused')

    def fused_timeouts(self):
        """Waits of the persistent launch that gave up since the last call (synchronises the device;
        0 unless a launch could not make progress)."""
        n = ctypes.c_uint()
        _lib.check(_lib.load().hcipy_mft_plan_get_fused(self.handle, None, None, ctypes.byref(n)),
                   'hcipy_mft_plan_get_fused')
        return n.value

    def autotune(self, direction, x, out, batch, workspace, output=_lib.OUTPUT_FIELD, repeats=20):
        """Times every tile shape of each stage on these buffers and pins the fastest.

        Returns the chosen shape per stage.  ``out`` and ``workspace`` are overwritten.
        """
        chosen = (ctypes.c_int * 2)()
        st = _lib.load().hcipy_mft_autotune(self.handle, direction, _lib.dptr(x), _lib.dptr(out), batch, output,
                                            _lib.dptr(workspace) if workspace is not None else None,
                                            0 if workspace is None else workspace.numel(), repeats, chosen,
                                            _lib.stream_ptr(x.device))
        _lib.check(st, 'hcipy_mft_autotune')
        return chosen[0], chosen[1]

    def execute(self, direction, x, out, batch, alpha, workspace, output=_lib.OUTPUT_FIELD):
        st = _lib.load().hcipy_mft_execute_ex(self.handle, direction, _lib.dptr(x), _lib.dptr(out), batch,
                                              float(alpha.real), float(alpha.imag), output,
                                              _lib.dptr(workspace) if workspace is not None else None,
                                              0 if workspace is None else workspace.numel(),
                                              _lib.stream_ptr(x.device))
        _lib.check(st, 'hcipy_mft_execute_ex')

    def close(self):
        if getattr(self, 'handle', None) is not None and _lib._lib is not None:
            _lib.load().hcipy_mft_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MftGroup:
    """Grouped execution of many matrix Fourier transforms as one launch per contraction stage
    (``hcipy_mft_execute_grouped``).  The calls may use different transforms (for example one
    per wavelength, or pupils of different sizes), directions and batch sizes, but one precision.
    Each item gets its own workspace, held by the group.  The transforms must keep their
    matrices (``precompute_matrices=True``, the default)."""

    def __init__(self, capacity=64, device=None):
        self.device = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
        self.capacity = int(capacity)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().hcipy_mft_group_create(ctypes.byref(h), self.capacity), 'hcipy_mft_group_create')
        self.handle = h
        self._ws = []

    def execute(self, calls, outs=None):
        """``calls``: sequence of ``(mft, direction, field, alpha, output)`` with direction
        ``_lib.FORWARD``/``_lib.BACKWARD``, a complex field tensor or Field (leading dimensions
        = batch), a complex scale and an ``_lib.OUTPUT_*`` kind.  Returns the output tensors,
        written into ``outs`` (contiguous tensors of the outputs' shapes and dtypes, for example
        rows of one stacked tensor) when given."""
        given = outs
        n = len(calls)
        if n == 0:
            return []
        if n > self.capacity:
            raise ValueError('%d calls for a group of capacity %d' % (n, self.capacity))
        items = (_lib.MftGroupItem * n)()
        outs, keep = [], []
        precs = set()
        while len(self._ws) < n:
            self._ws.append(None)
        for i, (mft, direction, field, alpha, output) in enumerate(calls):
            in_grid = mft.input_grid if direction == _lib.FORWARD else mft.output_grid
            out_grid = mft.output_grid if direction == _lib.FORWARD else mft.input_grid
            t, cdt = _as_field_tensor(field, in_grid)
            if t.device != self.device:
                raise ValueError('call %d: field on %s, group on %s' % (i, t.device, self.device))
            lead = tuple(t.shape[:-1])
            batch = int(np.prod(lead)) if lead else 1
            plan = mft.plan_for(cdt, t.device)
            precs.add(plan.precision)
            odt = {_lib.OUTPUT_FIELD: cdt, _lib.OUTPUT_FIELD_C128: torch.complex128,
                   _lib.OUTPUT_INTENSITY: torch.float64 if cdt == torch.complex128 else torch.float32,
                   _lib.OUTPUT_INTENSITY_F64: torch.float64}[output]
            if given is None:
                out = torch.empty(lead + (out_grid.size,), dtype=odt, device=t.device)
            else:
                out = given[i]
                if (out.dtype != odt or tuple(out.shape) != lead + (out_grid.size,) or not out.is_contiguous()
                        or out.device != t.device):
                    raise ValueError('call %d: output tensor of shape %s, %s; expected %s, %s, contiguous' %
                                     (i, tuple(out.shape), out.dtype, lead + (out_grid.size,), odt))
            need = max(plan.workspace_bytes(batch), 256)
            if self._ws[i] is None or self._ws[i].numel() < need:
                self._ws[i] = torch.empty(need, dtype=torch.uint8, device=t.device)
            ws = self._ws[i]
            a = complex(alpha)
            items[i] = _lib.MftGroupItem(plan.handle.value, direction, _lib.dptr(t), _lib.dptr(out), batch, a.real,
                                         a.imag, output, _lib.dptr(ws), ws.numel())
            outs.append(out)
            keep.append(t)
        if len(precs) != 1:
            raise ValueError('all calls of a grouped launch share one precision, got %s' % sorted(precs))
        st = _lib.load().hcipy_mft_execute_grouped(self.handle, ctypes.cast(items, ctypes.c_void_p), n,
                                                   _lib.stream_ptr(self.device))
        _lib.check(st, 'hcipy_mft_execute_grouped')
        return outs

    def release_captures(self):
        """Returns the group's capture snapshots (``hcipy_mft_group_release_captures``); graphs that
        recorded this group earlier must not be replayed afterwards.  Returns how many were held."""
        held = ctypes.c_int()
        _lib.check(_lib.load().hcipy_mft_group_release_captures(self.handle, ctypes.byref(held)),
                   'hcipy_mft_group_release_captures')
        return held.value

    def close(self):
        if getattr(self, 'handle', None) is not None and _lib._lib is not None:
            _lib.load().hcipy_mft_group_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _MftFn(torch.autograd.Function):
    """One MFT direction as an autograd node.  The vector-Jacobian product of
    out = alpha M1 . in . M2 is conj(alpha) M1^H . g . M2^H: the opposite direction's kernels with
    the conjugate scale (and vice versa), in the same precision."""

    @staticmethod
    def forward(ctx, t, mft, direction, alpha):
        ctx.mft, ctx.direction, ctx.alpha = mft, direction, alpha
        return mft._execute_plain(t, direction, alpha)

    @staticmethod
    def backward(ctx, g):
        other = _lib.BACKWARD if ctx.direction == _lib.FORWARD else _lib.FORWARD
        return ctx.mft._execute_plain(g.contiguous(), other, ctx.alpha.conjugate()), None, None, None


class _GroupedMftFn(torch.autograd.Function):
    """Several MFTs (one per call: transform, direction, scalar scale) as one grouped execution
    and one autograd node.  Each call's vector-Jacobian product is its transform's opposite
    direction with the conjugate scale (as for ``_MftFn``); the products of all calls run as one
    grouped execution too."""

    @staticmethod
    def forward(ctx, group, specs, *tensors):
        ctx.group, ctx.specs = group, specs
        outs = group.execute([(m, d, t, a, _lib.OUTPUT_FIELD) for (m, d, a), t in zip(specs, tensors)])
        return tuple(outs)

    @staticmethod
    def backward(ctx, *grads):
        calls = []
        for (m, d, a), g in zip(ctx.specs, grads):
            other = _lib.BACKWARD if d == _lib.FORWARD else _lib.FORWARD
            calls.append((m, other, g.contiguous(), a.conjugate(), _lib.OUTPUT_FIELD))
        return (None, None) + tuple(ctx.group.execute(calls))


class _GroupedStackFn(torch.autograd.Function):
    """As ``_GroupedMftFn`` for transforms of one input and one output grid (for example one
    propagator at several wavelengths) on a stacked tensor: row i of the input goes through call
    i, and its output is row i of the stacked output, so that pointwise operations on all rows
    run as one launch."""

    @staticmethod
    def forward(ctx, group, specs, out_size, x):
        ctx.group, ctx.specs, ctx.in_shape = group, specs, tuple(x.shape)
        out = torch.empty(tuple(x.shape[:-1]) + (out_size,), dtype=x.dtype, device=x.device)
        group.execute([(m, d, x[i], a, _lib.OUTPUT_FIELD) for i, (m, d, a) in enumerate(specs)],
                      outs=[out[i] for i in range(len(specs))])
        return out

    @staticmethod
    def backward(ctx, g):
        g = g.contiguous()
        gx = torch.empty(ctx.in_shape, dtype=g.dtype, device=g.device)
        calls = []
        for i, (m, d, a) in enumerate(ctx.specs):
            other = _lib.BACKWARD if d == _lib.FORWARD else _lib.FORWARD
            calls.append((m, other, g[i], a.conjugate(), _lib.OUTPUT_FIELD))
        ctx.group.execute(calls, outs=[gx[i] for i in range(len(ctx.specs))])
        return None, None, None, gx


def grouped_transforms_stacked(mfts, direction, x, extra_alphas, group):
    """Row i of the stacked complex tensor ``x`` (shape ``(n, ..., size)``) through
    ``mfts[i]._apply(., direction, extra_alphas[i])`` as one grouped execution; returns the
    stacked output ``(n, ..., out_size)``.  Differentiable; the transforms share one input grid
    and one output grid (sizes), with scalar weights."""
    n = len(mfts)
    in_grid = mfts[0].input_grid if direction == _lib.FORWARD else mfts[0].output_grid
    out_grid = mfts[0].output_grid if direction == _lib.FORWARD else mfts[0].input_grid
    if x.shape[0] != n or x.shape[-1] != in_grid.size:
        raise ValueError('stacked field of shape %s for %d transforms of %d-sample inputs' %
                         (tuple(x.shape), n, in_grid.size))
    specs = []
    for i, (m, ea) in enumerate(zip(mfts, extra_alphas)):
        g_in = m.input_grid if direction == _lib.FORWARD else m.output_grid
        g_out = m.output_grid if direction == _lib.FORWARD else m.input_grid
        if g_in.size != in_grid.size or g_out.size != out_grid.size:
            raise ValueError('transform %d: grid sizes differ from the first transform' % i)
        a = m.scalar_scale(direction, ea)
        if a is None:
            raise ValueError('transform %d: grouped transforms need grids with scalar weights' % i)
        specs.append((m, direction, a))
    x = x.to(complex_dtype_for(x.dtype)).contiguous()
    if torch.is_grad_enabled() and x.requires_grad:
        return _GroupedStackFn.apply(group, tuple(specs), out_grid.size, x)
    out = torch.empty(tuple(x.shape[:-1]) + (out_grid.size,), dtype=x.dtype, device=x.device)
    group.execute([(m, d, x[i], a, _lib.OUTPUT_FIELD) for i, (m, d, a) in enumerate(specs)],
                  outs=[out[i] for i in range(n)])
    return out


def grouped_transforms(calls, group):
    """``calls``: sequence of ``(mft, direction, field, extra_alpha)``, each meaning
    ``mft._apply(field, direction, extra_alpha)`` (the transform's grid weights included).  Runs
    them as one grouped execution of ``group`` (an ``MftGroup``) and returns the output tensors;
    with gradients recorded, the backward pass is one grouped execution as well.  Every call's
    input grid must have scalar weights."""
    specs, tensors = [], []
    for i, (mft, direction, field, extra_alpha) in enumerate(calls):
        in_grid = mft.input_grid if direction == _lib.FORWARD else mft.output_grid
        t, _ = _as_field_tensor(field, in_grid)
        if t.shape[-1] != in_grid.size:
            raise ValueError('call %d: field size %d does not match the grid size %d' % (i, t.shape[-1], in_grid.size))
        alpha = mft.scalar_scale(direction, extra_alpha)
        if alpha is None:
            raise ValueError('call %d: grouped transforms need grids with scalar weights' % i)
        specs.append((mft, direction, alpha))
        tensors.append(t)
    if torch.is_grad_enabled() and any(t.requires_grad for t in tensors):
        return list(_GroupedMftFn.apply(group, tuple(specs), *tensors))
    return group.execute([(m, d, t, a, _lib.OUTPUT_FIELD) for (m, d, a), t in zip(specs, tensors)])


class MatrixFourierTransform(FourierTransform):
    """Matrix Fourier transform between separated Cartesian grids, on MFMA.

    Same protocol and flags as hcipy/fourier/matrix_fourier_transform.py:8-309.  ``forward``
    computes ``w_in * M1 . E . M2`` with ``M1 = exp(-i v y)``, ``M2 = exp(-i x u)``;
    ``backward`` computes ``w_out/(2 pi)^ndim * M1^H . F . M2^H``.  Leading tensor dimensions
    of the field are one batched launch (the reference loops over them,
    fourier_transform.py:302-321).

    ``precision`` picks the complex64 arithmetic ('f16x2', 'bf16x3' or 'fp32'; default from
    ``Configuration().fourier.mft.precision_complex64``).  complex128 always runs in FP64.
    """

    def __init__(self, input_grid, output_grid, precompute_matrices=None, allocate_intermediate=None, precision=None):
        self.check_if_supported(input_grid, output_grid)
        self.input_grid = input_grid
        self.output_grid = output_grid
        self.shape_input = input_grid.shape
        self.shape_output = output_grid.shape
        self.ndim = input_grid.ndim
        cfg = Configuration().fourier.mft
        self.precompute_matrices = cfg.precompute_matrices if precompute_matrices is None else precompute_matrices
        self.allocate_intermediate = cfg.allocate_intermediate if allocate_intermediate is None else allocate_intermediate
        if precision is not None and precision not in ('f16x2', 'bf16x3', 'fp32'):
            raise ValueError("precision must be 'f16x2', 'bf16x3' or 'fp32' (complex64 fields)")
        self.precision = precision
        self._plans = {}
        self._workspace = {}
        self.matrices_dtype = None
        self.intermediate_dtype = None

    # The reference exposes M1/M2 (or M) only while matrices are kept (test_fourier.py:256-271).
    @property
    def M1(self):
        return next(iter(self._plans.values())) if (self._plans and self.ndim == 2) else None

    @property
    def M2(self):
        return self.M1

    @property
    def M(self):
        return next(iter(self._plans.values())) if (self._plans and self.ndim == 1) else None

    @property
    def intermediate_array(self):
        return next(iter(self._workspace.values())) if self._workspace else None

    def precision_for(self, cdt):
        if cdt == torch.complex128:
            return 'fp64'
        return self.precision or Configuration().fourier.mft.precision_complex64

    def plan_for(self, cdt, device):
        prec = self.precision_for(cdt)
        key = (prec, device.index)
        plan = self._plans.get(key)
        if plan is None:
            plan = _MftPlan(self.input_grid, self.output_grid, prec, device)
            self._plans[key] = plan
        self.matrices_dtype = cdt
        return plan

    def _workspace_for(self, plan, batch, device):
        need = plan.workspace_bytes(batch)
        if need == 0:
            return None
        ws = self._workspace.get(device.index)
        if ws is None or ws.numel() < need:
            ws = torch.empty(need, dtype=torch.uint8, device=device)
            self._workspace[device.index] = ws
        return ws

    def _release(self):
        if not self.precompute_matrices:
            for p in self._plans.values():
                p.close()
            self._plans = {}
            self.matrices_dtype = None
        if not self.allocate_intermediate:
            self._workspace = {}

    def scalar_scale(self, direction, extra_alpha=1.0):
        """The complex scale of one direction when the weights of its input grid are a single
        value (the transform is then alpha times the plain matrix product), else None."""
        w = self._direction_weights(direction)
        if np.isscalar(w) or np.size(w) == 1 or np.all(np.asarray(w) == np.asarray(w).ravel()[0]):
            return complex(float(np.asarray(w).ravel()[0])) * complex(extra_alpha)
        return None

    def _direction_weights(self, direction):
        if direction == _lib.FORWARD:
            return self.input_grid.weights
        # weights of the transform's output grid, which is this call's input grid
        return self.output_grid.weights / (2 * np.pi) ** self.ndim

    def _apply(self, field, direction, extra_alpha=1.0, out=None, output=_lib.OUTPUT_FIELD):
        in_grid, out_grid = (self.input_grid, self.output_grid) if direction == _lib.FORWARD else (self.output_grid, self.input_grid)
        t, cdt = _as_field_tensor(field, in_grid)
        if t.shape[-1] != in_grid.size:
            raise ValueError('Field size %d does not match the grid size %d' % (t.shape[-1], in_grid.size))
        alpha = self.scalar_scale(direction, extra_alpha)
        if alpha is None:
            # per-sample weights: one pass of the pointwise kernel (differentiable), then the plain product
            from .pointwise import sub_mul
            w = np.asarray(self._direction_weights(direction), dtype=np.float64).ravel()
            t = sub_mul(t, None, torch.as_tensor(w, device=t.device)).as_subclass(torch.Tensor).to(cdt).contiguous()
            alpha = complex(extra_alpha)
        if t.requires_grad and torch.is_grad_enabled():
            if out is not None:
                raise ValueError('out= is not supported while recording gradients')
            o = _MftFn.apply(t, self, direction, alpha)
            if output == _lib.OUTPUT_FIELD:
                return Field(o, out_grid)
            if output == _lib.OUTPUT_FIELD_C128:
                return Field(o.to(torch.complex128), out_grid)
            from .pointwise import intensity
            i = intensity(o)
            return Field(i.to(torch.float64) if output == _lib.OUTPUT_INTENSITY_F64 else i, out_grid)
        return Field(self._execute_plain(t, direction, alpha, output, out), out_grid)

    def _execute_plain(self, t, direction, alpha, output=_lib.OUTPUT_FIELD, out=None):
        """Runs the kernels on a contiguous complex tensor; returns the output tensor."""
        out_grid = self.output_grid if direction == _lib.FORWARD else self.input_grid
        cdt = t.dtype
        lead = tuple(t.shape[:-1])
        batch = int(np.prod(lead)) if lead else 1
        plan = self.plan_for(cdt, t.device)
        if out is None:
            odt = {_lib.OUTPUT_FIELD: cdt, _lib.OUTPUT_FIELD_C128: torch.complex128,
                   _lib.OUTPUT_INTENSITY: torch.float64 if cdt == torch.complex128 else torch.float32,
                   _lib.OUTPUT_INTENSITY_F64: torch.float64}[output]
            out = torch.empty(lead + (out_grid.size,), dtype=odt, device=t.device)
        ws = self._workspace_for(plan, batch, t.device)
        plan.execute(direction, t, out, batch, alpha, ws, output)
        self._release()
        return out

    def autotune(self, field, repeats=20):
        """Times every kernel tile shape of both directions on fields like ``field`` (same dtype
        and leading shape) and pins the fastest on the plan.  Returns the chosen shapes
        ``{'forward': (stage0, stage1), 'backward': (...)}``.  The pinned shapes live as long as
        the plan (for the whole object with ``precompute_matrices=True``)."""
        t, cdt = _as_field_tensor(field, self.input_grid)
        lead = tuple(t.shape[:-1])
        batch = int(np.prod(lead)) if lead else 1
        plan = self.plan_for(cdt, t.device)
        ws = self._workspace_for(plan, batch, t.device)
        out = torch.empty(lead + (self.output_grid.size,), dtype=cdt, device=t.device)
        fwd = plan.autotune(_lib.FORWARD, t, out, batch, ws, repeats=repeats)
        bwd = plan.autotune(_lib.BACKWARD, out, torch.empty_like(t), batch, ws, repeats=repeats)
        return {'forward': fwd, 'backward': bwd}

    def forward(self, field):
        """Forward transform of a Field on ``input_grid`` (matrix_fourier_transform.py:132-176)."""
        return self._apply(field, _lib.FORWARD)

    def backward(self, field):
        """Adjoint-weighted inverse transform (matrix_fourier_transform.py:178-224)."""
        return self._apply(field, _lib.BACKWARD)

    def forward_intensity(self, field, float64=False):
        """|forward(field)|^2 evaluated in the GEMM epilogue from the FP64 result (float32 for
        complex64 fields unless ``float64``), before any rounding of the field."""
        return self._apply(field, _lib.FORWARD, output=_lib.OUTPUT_INTENSITY_F64 if float64 else _lib.OUTPUT_INTENSITY)

    def forward_c128(self, field):
        """forward(field) returned as complex128 whatever the arithmetic."""
        return self._apply(field, _lib.FORWARD, output=_lib.OUTPUT_FIELD_C128)

    @classmethod
    def check_if_supported(cls, input_grid, output_grid):
        """matrix_fourier_transform.py:227-254."""
        if not input_grid.is_separated or not input_grid.is_('cartesian'):
            raise ValueError('The input_grid must be separable in cartesian coordinates.')
        if not output_grid.is_separated or not output_grid.is_('cartesian'):
            raise ValueError('The output_grid must be separable in cartesian coordinates.')
        if input_grid.ndim not in [1, 2]:
            raise ValueError('The input_grid must be one- or two-dimensional.')
        if input_grid.ndim != output_grid.ndim:
            raise ValueError('The input_grid must have the same dimensions as the output_grid.')

    @classmethod
    def compute_complexity(cls, input_grid, output_grid):
        """Real-operation count of the two GEMMs (matrix_fourier_transform.py:256-309)."""
        cls.check_if_supported(input_grid, output_grid)
        if input_grid.ndim == 1:
            cm = input_grid.size * output_grid.size
            ca = (input_grid.size - 1) * output_grid.size
        else:
            nox, noy = output_grid.shape
            nix, niy = input_grid.shape
            cm = niy * nix * noy
            ca = noy * nix * (niy - 1)
            cm += nix * nox * noy
            ca += nox * noy * (nix - 1)
            cm += input_grid.size
        num_mult = 4 * cm
        num_add = 2 * cm + 2 * ca
        coeffs = Configuration().fourier.mft.execution_time_prediction_coefficients
        t = FourierTransform._predict_execution_time(num_mult + num_add, coeffs)
        return ComputationalComplexity(num_multiplications=num_mult, num_additions=num_add, expected_execution_time=t)


# ----------------------------------------------------------------------------------------------
# Zoom FFT (Bluestein chirp-z transform per axis)


class _ZoomPlan:
    """Owns one C-ABI zoom-FFT plan (per-axis chirps and filter spectra in one dtype), or with
    ``direct`` a direct power-of-two FFT plan (per-axis twiddles and phase ramps), which takes
    the same execute arguments."""

    def __init__(self, input_grid, output_grid, cdt, precision, device, fft_n=None, direct=False):
        lib = _lib.load()
        self._api = 'fft' if direct else 'zoom'

        nd = input_grid.ndim
        ci = ctypes.c_int * nd
        cd = ctypes.c_double * nd
        grids = (ci(*[int(v) for v in input_grid.dims]), cd(*[float(v) for v in input_grid.delta]),
                 cd(*[float(v) for v in input_grid.zero]), ci(*[int(v) for v in output_grid.dims]),
                 cd(*[float(v) for v in output_grid.delta]), cd(*[float(v) for v in output_grid.zero]))
        with torch.cuda.device(device):
            h = ctypes.c_void_p()
            if direct:
                st = lib.hcipy_fft_plan_create(ctypes.byref(h), nd, *grids, ci(*[int(v) for v in fft_n]),
                                               _lib._DTYPE_CODES[cdt], _lib.PRECISIONS[precision],
                                               _lib.stream_ptr(device))
                _lib.check(st, 'hcipy_fft_plan_create')
            elif fft_n is None:
                st = lib.hcipy_zoom_plan_create(ctypes.byref(h), nd, *grids, _lib._DTYPE_CODES[cdt],
                                                _lib.PRECISIONS[precision], _lib.stream_ptr(device))
                _lib.check(st, 'hcipy_zoom_plan_create')
            else:
                st = lib.hcipy_zoom_plan_create_fft(ctypes.byref(h), nd, *grids, ci(*[int(v) for v in fft_n]),
                                                    _lib._DTYPE_CODES[cdt], _lib.PRECISIONS[precision],
                                                    _lib.stream_ptr(device))
                _lib.check(st, 'hcipy_zoom_plan_create_fft')
            self.handle = h

    def workspace_bytes(self, batch):
        n = ctypes.c_size_t()
        name = 'hcipy_%s_workspace_size' % self._api
        _lib.check(getattr(_lib.load(), name)(self.handle, batch, ctypes.byref(n)), name)
        return n.value

    def execute(self, direction, x, out, batch, alpha, workspace):
        name = 'hcipy_%s_execute' % self._api
        st = getattr(_lib.load(), name)(self.handle, direction, _lib.dptr(x), _lib.dptr(out), batch,
                                        float(alpha.real), float(alpha.imag),
                                        _lib.dptr(workspace) if workspace is not None else None,
                                        0 if workspace is None else workspace.numel(), _lib.stream_ptr(x.device))
        _lib.check(st, name)

    def close(self):
        if getattr(self, 'handle', None) is not None and _lib._lib is not None:
            getattr(_lib.load(), 'hcipy_%s_plan_destroy' % self._api)(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _ZoomFn(torch.autograd.Function):
    """One zoom-FFT direction as an autograd node: the vector-Jacobian product of
    alpha sum_j x_j exp(s i p_j q_k) is conj(alpha) sum_k g_k exp(-s i p_j q_k), the opposite
    direction with the conjugate scale."""

    @staticmethod
    def forward(ctx, t, zfft, direction, alpha):
        ctx.zfft, ctx.direction, ctx.alpha = zfft, direction, alpha
        return zfft._execute_plain(t, direction, alpha)

    @staticmethod
    def backward(ctx, g):
        other = _lib.BACKWARD if ctx.direction == _lib.FORWARD else _lib.FORWARD
        return ctx.zfft._execute_plain(g.contiguous(), other, ctx.alpha.conjugate()), None, None, None


class ZoomFastFourierTransform(FourierTransform):
    """Fourier transform between regular Cartesian grids of any spacing by a Bluestein chirp-z
    transform along each axis (hcipy/fourier/zoom_fast_fourier_transform.py:10-219).

    ``forward`` computes ``w_in * sum_j f_j exp(-i x_j . u_k)``; ``backward`` computes
    ``w_out / (2 pi)^ndim * sum_k F_k exp(+i u_k . x_j)``, the same sums as the matrix Fourier
    transform.  complex128 fields run in FP64.  complex64 fields run with ``precision`` 'fp64'
    (FP32 loads and stores, FP64 chirps, line FFTs and intermediate; the default from
    ``Configuration().fourier.zfft.precision_complex64``) or 'fp32' (all FP32, chirps and filter
    spectra rounded from FP64).  One- and two-dimensional grids; each axis needs
    ``n_in + n_out - 1 <= 8192`` (one line FFT in LDS per row).
    """

    def __init__(self, input_grid, output_grid, precision=None):
        self.check_if_supported(input_grid, output_grid)
        if precision is not None and precision not in ('fp64', 'fp32'):
            raise ValueError("precision must be 'fp64' or 'fp32' (complex64 fields)")
        self.input_grid = input_grid
        self.output_grid = output_grid
        self.ndim = input_grid.ndim
        self.precision = precision
        # FFT length per axis when the grids are an FFT pair (set by FastFourierTransform): the
        # chirps then use the exact 2 pi / N
        self._fft_n = None
        self._direct = False
        self._plans = {}
        self._workspace = {}

    def precision_for(self, cdt):
        if cdt == torch.complex128:
            return 'fp64'
        return self.precision or Configuration().fourier.zfft.precision_complex64

    def plan_for(self, cdt, device):
        prec = self.precision_for(cdt)
        key = (cdt, prec, device.index)
        plan = self._plans.get(key)
        if plan is None:
            plan = _ZoomPlan(self.input_grid, self.output_grid, cdt, prec, device, fft_n=self._fft_n,
                             direct=self._direct)
            self._plans[key] = plan
        return plan

    def _workspace_for(self, plan, batch, device):
        need = plan.workspace_bytes(batch)
        if need == 0:
            return None
        ws = self._workspace.get(device.index)
        if ws is None or ws.numel() < need:
            ws = torch.empty(need, dtype=torch.uint8, device=device)
            self._workspace[device.index] = ws
        return ws

    def _apply(self, field, direction):
        in_grid, out_grid = (self.input_grid, self.output_grid) if direction == _lib.FORWARD else (self.output_grid, self.input_grid)
        t, cdt = _as_field_tensor(field, in_grid)
        if t.shape[-1] != in_grid.size:
            raise ValueError('Field size %d does not match the grid size %d' % (t.shape[-1], in_grid.size))
        w = in_grid.weights if direction == _lib.FORWARD else in_grid.weights / (2 * np.pi) ** self.ndim
        alpha = complex(float(np.asarray(w).ravel()[0]))
        if t.requires_grad and torch.is_grad_enabled():
            return Field(_ZoomFn.apply(t, self, direction, alpha), out_grid)
        return Field(self._execute_plain(t, direction, alpha), out_grid)

    def _execute_plain(self, t, direction, alpha):
        out_grid = self.output_grid if direction == _lib.FORWARD else self.input_grid
        lead = tuple(t.shape[:-1])
        batch = int(np.prod(lead)) if lead else 1
        plan = self.plan_for(t.dtype, t.device)
        out = torch.empty(lead + (out_grid.size,), dtype=t.dtype, device=t.device)
        plan.execute(direction, t, out, batch, alpha, self._workspace_for(plan, batch, t.device))
        return out

    def forward(self, field):
        """Forward transform (zoom_fast_fourier_transform.py:70-92)."""
        return self._apply(field, _lib.FORWARD)

    def backward(self, field):
        """Inverse transform with the output weights (zoom_fast_fourier_transform.py:94-116)."""
        return self._apply(field, _lib.BACKWARD)

    @classmethod
    def check_if_supported(cls, input_grid, output_grid):
        """zoom_fast_fourier_transform.py:118-141, and the size limits of the device transform."""
        if not input_grid.is_regular or not input_grid.is_('cartesian'):
            raise ValueError('The input grid should be regularly spaced in Cartesian coordinates.')
        if not output_grid.is_regular or not output_grid.is_('cartesian'):
            raise ValueError('The output grid should be regularly spaced in Cartesian coordinates.')
        if input_grid.ndim != output_grid.ndim:
            raise ValueError('The input_grid must have the same dimensions as the output_grid.')
        if input_grid.ndim not in (1, 2):
            raise ValueError('The zoom FFT of this library supports one- and two-dimensional grids.')
        for n, m in zip(input_grid.dims, output_grid.dims):
            if int(n) + int(m) - 1 > 8192:
                raise ValueError('n_in + n_out - 1 = %d exceeds the 8192-point line FFT of the zoom FFT.' % (int(n) + int(m) - 1))

    @classmethod
    def compute_complexity(cls, input_grid, output_grid):
        """Operation count of the reference's chirp-z passes (zoom_fast_fourier_transform.py:143-219):
        the same terms, with the reference's next_fast_len for the convolution length."""
        import scipy.fft
        cls.check_if_supported(input_grid, output_grid)
        cm = input_grid.size
        ca = 0
        for i, (n_in, n_out) in enumerate(zip(input_grid.shape, output_grid.shape)):
            czt_len = scipy.fft.next_fast_len(n_in + n_out - 1)
            lines = math.prod(output_grid.shape[:i]) * math.prod(input_grid.shape[i + 1:])
            before = math.prod(output_grid.shape[:i]) * math.prod(input_grid.shape[i:])
            after = math.prod(output_grid.shape[:i + 1]) * math.prod(input_grid.shape[i + 1:])
            fft_mults = 0.5 * czt_len * lines * math.log2(czt_len)
            fft_adds = czt_len * lines * math.log2(czt_len)
            cm += before + 2 * fft_mults + czt_len * lines + 2 * after
            ca += 2 * fft_adds
        num_mult = 4 * cm
        num_add = 2 * cm + 2 * ca
        coeffs = Configuration().fourier.zfft.execution_time_prediction_coefficients
        t = FourierTransform._predict_execution_time(num_mult + num_add, coeffs)
        return ComputationalComplexity(num_multiplications=num_mult, num_additions=num_add, expected_execution_time=t)


__all__ += ['ZoomFastFourierTransform']


# ----------------------------------------------------------------------------------------------
# FastFourierTransform and the planner


class _DirectFft(ZoomFastFourierTransform):
    """FastFourierTransform's power-of-two path: one line FFT of the internal length per row and
    per column of the unpadded field, with the FFT grid's phase ramps on load and store
    (``hcipy_fft_execute``).  Same sums, weights and autograd as the zoom FFT; no zoom-FFT size
    limit (each internal length is a power of two in [4, 8192])."""

    def __init__(self, input_grid, output_grid, fft_n, precision=None):
        if precision is not None and precision not in ('fp64', 'fp32'):
            raise ValueError("precision must be 'fp64' or 'fp32' (complex64 fields)")
        self.input_grid = input_grid
        self.output_grid = output_grid
        self.ndim = input_grid.ndim
        self.precision = precision
        self._fft_n = [int(n) for n in fft_n]
        self._direct = True
        self._plans = {}
        self._workspace = {}

    @staticmethod
    def supports(fft_n):
        return all(4 <= int(n) <= 8192 and int(n) & (int(n) - 1) == 0 for n in fft_n)


class FastFourierTransform(FourierTransform):
    """Zero-padded / cropped / shifted Fourier transform onto an FFT grid
    (hcipy/fourier/fast_fourier_transform.py:188-482).

    The reference's FFT output (with its input/output phase-ramp corrections) equals the
    Fourier sum onto ``make_fft_grid(input_grid, q, fov, shift)``.  This class evaluates that sum
    with a direct FFT of the internal length (``method`` 'fft') when every internal length is a
    power of two in [4, 8192]; otherwise with the Bluestein zoom FFT ('zft'; measured faster than
    the MFT at every size from 128^2 up) when each axis fits its 8192-point line FFT, and with
    the matrix Fourier transform ('mft') beyond that, or when ``precision`` 'f16x2' / 'bf16x3'
    asks for a split-precision MFT.  ``precision`` None / 'fp32' selects the complex64 arithmetic
    of the FFTs (FP64 internal / FP32) or of the MFT.  ``emulate_fftshifts`` is accepted for
    interface compatibility (it only changes the reference's rounding).
    """

    def __init__(self, input_grid, q=1, fov=1, shift=0, emulate_fftshifts=None, precision=None):
        if not input_grid.is_regular:
            raise ValueError('The input_grid must be regular.')
        if not input_grid.is_('cartesian'):
            raise ValueError('The input_grid must be Cartesian.')
        q_check = (q,) if np.isscalar(q) else q
        fov_check = (fov,) if np.isscalar(fov) else fov
        if any(q_i < 1 for q_i in q_check):
            raise ValueError('The amount of zeropadding (q) must be larger than 1.')
        if any(fov_i < 0 for fov_i in fov_check):
            raise ValueError('The amount of cropping (fov) must be positive.')
        self.input_grid = input_grid
        self.emulate_fftshifts = Configuration().fourier.fft.emulate_fftshifts if emulate_fftshifts is None else emulate_fftshifts
        self.output_grid = make_fft_grid(input_grid, q, fov, shift)
        self.internal_grid = make_fft_grid(input_grid, q, 1)
        if np.any(np.asarray(self.output_grid.dims) > np.asarray(self.internal_grid.dims)):
            raise ValueError('The amount of cropping (fov) must be smaller than 1.')
        self.shape_in = input_grid.shape
        self.shape_out = self.output_grid.shape
        self.internal_shape = self.internal_grid.shape
        self._impl = None
        fft_n = [int(n) for n in self.internal_grid.dims]
        if precision not in ('f16x2', 'bf16x3') and input_grid.ndim in (1, 2) and _DirectFft.supports(fft_n):
            self._impl = _DirectFft(input_grid, self.output_grid, fft_n, precision=precision)
        elif precision not in ('f16x2', 'bf16x3'):
            try:
                self._impl = ZoomFastFourierTransform(input_grid, self.output_grid, precision=precision)
                self._impl._fft_n = [int(n) for n in self.internal_grid.dims]
            except ValueError:
                self._impl = None
        if self._impl is None:
            self._impl = MatrixFourierTransform(input_grid, self.output_grid, precision=precision)

    @property
    def method(self):
        """'fft', 'zft' or 'mft': the transform that evaluates this FFT."""
        if isinstance(self._impl, _DirectFft):
            return 'fft'
        return 'zft' if isinstance(self._impl, ZoomFastFourierTransform) else 'mft'

    def forward(self, field):
        return self._impl.forward(field)

    def backward(self, field):
        return self._impl.backward(field)

    @classmethod
    def check_if_supported(cls, input_grid, output_grid):
        get_fft_parameters(output_grid, input_grid)

    @classmethod
    def compute_complexity(cls, input_grid, output_grid):
        """Operation count of the reference FFT (fast_fourier_transform.py:430-482)."""
        q, _, shift = get_fft_parameters(output_grid, input_grid)
        shape = tuple(n * qq for n, qq in zip(input_grid.shape, q))
        n_int = math.prod(shape)
        cm = 0.5 * n_int * math.log2(n_int)
        ca = n_int * math.log2(n_int)
        if not all(abs(s) <= 1e-8 for s in shift):
            cm += input_grid.size
        cm += output_grid.size
        num_mult = 4 * cm
        num_add = 2 * cm + 2 * ca
        coeffs = Configuration().fourier.fft.execution_time_prediction_coefficients
        t = FourierTransform._predict_execution_time(num_mult + num_add, coeffs)
        return ComputationalComplexity(num_multiplications=num_mult, num_additions=num_add, expected_execution_time=t)


def _time_forward(ft, input_grid, repeats=10):
    """Seconds per forward transform of a zero complex128 field on the current device."""
    a = torch.zeros(input_grid.size, dtype=torch.complex128, device=torch.device('cuda', torch.cuda.current_device()))
    f = Field(a, input_grid)
    for _ in range(2):
        ft.forward(f)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        ft.forward(f)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / repeats


def make_fourier_transform(input_grid, output_grid=None, q=1, fov=1, shift=0, planner='estimate', precision=None):
    """Fourier transform between two grids, the fastest of the candidate methods
    (hcipy/fourier/fourier_transform.py:217-300).

    Candidates: 'fft' (FastFourierTransform, when the output grid is an FFT grid of the input),
    'mft' (MatrixFourierTransform) and 'zft' (ZoomFastFourierTransform), each where it supports
    the grids.  ``planner='estimate'`` ranks them by ``compute_complexity`` with the time-model
    coefficients of the configuration (the reference's, so the choice is the reference's);
    ``'measure'`` times one forward transform of each on the device.  The reference's
    NaiveFourierTransform has no device implementation here, so unstructured grids raise
    ValueError.  ``precision`` is passed to the method that takes it ('f16x2'/'fp32' for the
    MFT-based methods, 'fp64'/'fp32' for the zoom FFT).
    """
    if planner not in ('estimate', 'measure'):
        raise ValueError("planner must be 'estimate' or 'measure'")
    options = {'fft': FastFourierTransform, 'mft': MatrixFourierTransform, 'zft': ZoomFastFourierTransform}
    if output_grid is not None:
        try:
            q, fov, shift = get_fft_parameters(output_grid, input_grid)
        except ValueError:
            del options['fft']
    else:
        output_grid = make_fft_grid(input_grid, q, fov, shift)

    def build(method):
        if method == 'fft':
            return FastFourierTransform(input_grid, q, fov, shift,
                                        precision=precision if precision in ('f16x2', 'bf16x3', 'fp32') else None)
        if method == 'mft':
            return MatrixFourierTransform(input_grid, output_grid,
                                          precision=precision if precision in ('f16x2', 'bf16x3', 'fp32') else None)
        return ZoomFastFourierTransform(input_grid, output_grid,
                                        precision=precision if precision in ('fp64', 'fp32') else None)

    expected = {}
    for method, ft in options.items():
        if not ft.is_supported(input_grid, output_grid):
            continue
        if planner == 'estimate':
            expected[method] = ft.compute_complexity(input_grid, output_grid).expected_execution_time
        else:
            expected[method] = _time_forward(build(method), input_grid)
    if not expected:
        raise ValueError('No Fourier transform of this library supports these grids (separated Cartesian grids are required).')
    return build(min(expected, key=expected.get))


__all__ += ['FastFourierTransform', 'make_fourier_transform']
