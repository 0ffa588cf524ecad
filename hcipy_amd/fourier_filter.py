"""FourierFilter: zero-pad -> FFT -> transfer function -> inverse FFT -> crop
(hcipy/fourier/fourier_operations.py:8-151).

Power-of-two internal grids run on the library's three-pass Stockham FFT kernels
(``hcipy_filter_*``).  Other sizes run the same operator as FastFourierTransform forward onto
the internal frequency grid, the transfer function, and its backward onto the input grid (the
zoom FFT with exact 2 pi / N chirps, or the MFT beyond its line length); on the FFT grid this
is the padded FFT filter, since the weights multiply to 1/(NX NY) and the kernel
exp(i u (x' - x)) is N-periodic.
"""
import ctypes

import numpy as np
import torch

from . import _lib, pointwise
from .field import Field
from .fourier import FastFourierTransform, complex_dtype_for, make_fft_grid

__all__ = ['FourierFilter', 'NearFieldTransferFunction']


class NearFieldTransferFunction:
    """Transfer function generated on the device in FP64 (kind 'angular_spectrum' or 'fresnel'),
    averaged over oversampling^2 dithered sub-grids of the internal grid."""

    KINDS = {'angular_spectrum': 0, 'fresnel': 1}

    def __init__(self, kind, wavenumber, distance, oversampling=2):
        self.kind = kind
        self.wavenumber = float(wavenumber)
        self.distance = float(distance)
        self.oversampling = int(round(oversampling))


def _pow2(n):
    return 4 <= n <= 8192 and (n & (n - 1)) == 0


class _FilterPlan:
    """One device plan: the filter geometry and one scalar transfer function (``tf``, if given, is
    a centred device tensor used instead of the filter's own transfer function)."""

    def __init__(self, filt, cdt, device, tf=None):
        lib = _lib.load()
        self.device = device
        self.dtype = cdt
        ny, nx = filt.input_grid.shape
        NY, NX = filt.internal_grid.shape
        dcode = _lib.DTYPE_C64 if cdt == torch.complex64 else _lib.DTYPE_C128
        h = ctypes.c_void_p()
        with torch.cuda.device(device):
            stream = _lib.stream_ptr(device)
            _lib.check(lib.hcipy_filter_plan_create(ctypes.byref(h), nx, ny, NX, NY, dcode, stream), 'hcipy_filter_plan_create')
            self.handle = h
            if tf is None:
                tf = filt.transfer_function
            if isinstance(tf, NearFieldTransferFunction):
                g = filt.internal_grid
                st = lib.hcipy_filter_nearfield_transfer_function(
                    h, NearFieldTransferFunction.KINDS[tf.kind], float(g.delta[0]), float(g.zero[0]), float(g.delta[1]),
                    float(g.zero[1]), tf.wavenumber, tf.distance, tf.oversampling, stream)
                _lib.check(st, 'hcipy_filter_nearfield_transfer_function')
            else:
                t = tf if isinstance(tf, torch.Tensor) and not isinstance(tf, Field) and tf.dim() == 1 \
                    else filt.evaluate_transfer_function(device)
                t = t.to(device=device, dtype=cdt).contiguous()
                _lib.check(lib.hcipy_filter_set_transfer_function(h, _lib.dptr(t), dcode, stream),
                           'hcipy_filter_set_transfer_function')
                torch.cuda.current_stream(device).synchronize()
        self._ws = None

    def workspace(self, batch):
        n = ctypes.c_size_t()
        _lib.check(_lib.load().hcipy_filter_workspace_size(self.handle, batch, ctypes.byref(n)), 'hcipy_filter_workspace_size')
        if self._ws is None or self._ws.numel() < n.value:
            self._ws = torch.empty(max(n.value, 256), dtype=torch.uint8, device=self.device)
        return self._ws

    def execute(self, adjoint, x, out, batch):
        ws = self.workspace(batch)
        st = _lib.load().hcipy_filter_execute(self.handle, int(adjoint), _lib.dptr(x), _lib.dptr(out), batch,
                                              _lib.dptr(ws), ws.numel(), _lib.stream_ptr(x.device))
        _lib.check(st, 'hcipy_filter_execute')

    @staticmethod
    def execute_multi(plans, adjoint, x, outs, batch, ws_holder):
        """The filters of several plans (same grids) applied to one input: one row pass over the
        input and one forward column FFT per column shared by all of them."""
        k = len(plans)
        lib = _lib.load()
        hs = (ctypes.c_void_p * k)(*[pl.handle.value for pl in plans])
        n = ctypes.c_size_t()
        _lib.check(lib.hcipy_filter_workspace_size_multi(hs, k, batch, ctypes.byref(n)), 'hcipy_filter_workspace_size_multi')
        ws = ws_holder.get('ws')
        if ws is None or ws.numel() < n.value or ws.device != x.device:
            ws = torch.empty(max(n.value, 256), dtype=torch.uint8, device=x.device)
            ws_holder['ws'] = ws
        op = (ctypes.c_void_p * k)(*[_lib.dptr(o) for o in outs])
        st = lib.hcipy_filter_execute_multi(hs, k, int(adjoint), _lib.dptr(x), op, batch, _lib.dptr(ws), ws.numel(),
                                            _lib.stream_ptr(x.device))
        _lib.check(st, 'hcipy_filter_execute_multi')

    @staticmethod
    def execute_matrix(plans, n_out, n_in, conj, x, out, batch, ws_holder):
        """Matrix filter: ``x`` is [n_in][batch] fields, ``out`` [n_out][batch];
        plans[o * n_in + c] filters input c into output o, with conj(tf) if ``conj``."""
        lib = _lib.load()
        hs = (ctypes.c_void_p * len(plans))(*[pl.handle.value for pl in plans])
        n = ctypes.c_size_t()
        _lib.check(lib.hcipy_filter_workspace_size_matrix(hs[0], n_out, n_in, batch, ctypes.byref(n)),
                   'hcipy_filter_workspace_size_matrix')
        ws = ws_holder.get('ws')
        if ws is None or ws.numel() < n.value or ws.device != x.device:
            ws = torch.empty(max(n.value, 256), dtype=torch.uint8, device=x.device)
            ws_holder['ws'] = ws
        st = lib.hcipy_filter_execute_matrix(hs, n_out, n_in, int(conj), _lib.dptr(x), _lib.dptr(out), batch,
                                             _lib.dptr(ws), ws.numel(), _lib.stream_ptr(x.device))
        _lib.check(st, 'hcipy_filter_execute_matrix')

    def __del__(self):
        try:
            if getattr(self, 'handle', None) is not None and _lib._lib is not None:
                _lib.load().hcipy_filter_plan_destroy(self.handle)
        except Exception:
            pass


class _FilterFn(torch.autograd.Function):
    """FourierFilter as an autograd node: crop . IFFT . diag(tf) . FFT . pad has the adjoint
    crop . IFFT . diag(conj tf) . FFT . pad, i.e. the filter with ``adjoint`` flipped."""

    @staticmethod
    def forward(ctx, t, flt, adjoint):
        ctx.flt, ctx.adjoint = flt, adjoint
        return flt._run_plain(t, adjoint)

    @staticmethod
    def backward(ctx, g):
        return ctx.flt._run_plain(g.contiguous(), not ctx.adjoint), None, None


class FourierFilter:
    """Filter in the Fourier domain with zero padding q (fourier_operations.py:8-151).

    ``transfer_function`` is a Field on the internal grid (centred order), a callable that
    maps the internal grid to such a Field, or a NearFieldTransferFunction.  A transfer function
    of shape (p, q, internal size) is a matrix field (fourier_operations.py:120-131): it
    contracts with the field's first dimension (q components, as field_dot does; any further
    dimensions are carried along) and the output has p there, ``out_i = sum_j filter(tf_ij)(in_j)``
    (the adjoint uses the conjugate transpose).  By linearity this is
    ``IFFT(sum_j tf_ij FFT(in_j))``: each input is transformed once and each output
    inverse-transformed once (``hcipy_filter_execute_matrix``).
    """

    def __init__(self, input_grid, transfer_function, q=1):
        if not input_grid.is_regular or input_grid.ndim != 2:
            raise ValueError('FourierFilter needs a regular two-dimensional input grid.')
        self.input_grid = input_grid
        self.internal_grid = make_fft_grid(input_grid, q, 1)
        self.q = q
        self.shape_in = input_grid.shape
        self.transfer_function = transfer_function
        self._plans = {}
        self._multi_ws = {}
        self._fft = None
        self._tf_cache = {}
        self._matrix_tf = None

    def evaluate_transfer_function(self, device):
        tf = self.transfer_function
        if isinstance(tf, NearFieldTransferFunction):
            # generated on the device in FP64, centred order, on the internal grid
            g = self.internal_grid
            out = torch.empty(g.size, dtype=torch.complex128, device=device)
            with torch.cuda.device(device):
                st = _lib.load().hcipy_nearfield_transfer_function(
                    NearFieldTransferFunction.KINDS[tf.kind], int(g.dims[0]), int(g.dims[1]), float(g.delta[0]),
                    float(g.zero[0]), float(g.delta[1]), float(g.zero[1]), tf.wavenumber, tf.distance, tf.oversampling,
                    _lib.dptr(out), _lib.stream_ptr(device))
                _lib.check(st, 'hcipy_nearfield_transfer_function')
            return out
        if callable(tf) and not isinstance(tf, torch.Tensor):
            tf = tf(self.internal_grid)
        if not isinstance(tf, torch.Tensor):
            tf = torch.as_tensor(np.asarray(tf), device=device)
        tf = tf.as_subclass(torch.Tensor) if isinstance(tf, Field) else tf
        if tf.dim() not in (1, 3) or tf.shape[-1] != self.internal_grid.size:
            raise ValueError('the transfer function must be a scalar or matrix field on the internal grid')
        return tf.to(device)

    def _is_matrix(self, device):
        tf = self.transfer_function
        if isinstance(tf, NearFieldTransferFunction):
            return False
        if self._matrix_tf is None:
            t = self.evaluate_transfer_function(device)
            self._matrix_tf = t if t.dim() == 3 else False
        return self._matrix_tf is not False

    def _element_plan(self, cdt, device, i, j):
        key = (cdt, device.index, i, j, False)
        plan = self._plans.get(key)
        if plan is None:
            plan = _FilterPlan(self, cdt, device, tf=self._matrix_tf[i, j])
            self._plans[key] = plan
        return plan

    def _run_matrix(self, t, adjoint):
        """Matrix transfer function: out_i = sum_j F_ij(in_j), adjoint in_j = sum_i F_ij^H(out_i).

        Up to 4 x 4 runs as one matrix filter (each input transformed once, each output
        inverse-transformed once); larger matrices run in 4 x 4 blocks whose partial sums are
        added."""
        tf = self._matrix_tf
        p, q = int(tf.shape[0]), int(tf.shape[1])
        n_in, n_out = (p, q) if adjoint else (q, p)
        if t.dim() < 2 or t.shape[0] != n_in:
            raise ValueError('a matrix transfer function of shape (%d, %d) needs %d field components, got shape %s'
                             % (p, q, n_in, tuple(t.shape)))
        cdt, dev_ = t.dtype, t.device
        lead = tuple(t.shape[1:-1])
        batch = int(np.prod(lead)) if lead else 1
        x = t.reshape(n_in, batch, t.shape[-1])
        out = torch.empty((n_out, batch, t.shape[-1]), dtype=cdt, device=dev_)
        B = 4
        for o0 in range(0, n_out, B):
            no = min(B, n_out - o0)
            acc = None
            for c0 in range(0, n_in, B):
                nc = min(B, n_in - c0)
                # plan of (output o, input c): element (o, c) of the matrix, or (c, o) for the adjoint
                plans = [self._element_plan(cdt, dev_, *((c, o) if adjoint else (o, c)))
                         for o in range(o0, o0 + no) for c in range(c0, c0 + nc)]
                dst = out[o0:o0 + no] if acc is None else torch.empty((no, batch, t.shape[-1]), dtype=cdt, device=dev_)
                for b0 in range(0, batch, 65535):
                    nb = min(65535, batch - b0)
                    xs = x[c0:c0 + nc, b0:b0 + nb]
                    xs = xs if (nb == batch and xs.is_contiguous()) else xs.contiguous()
                    whole = nb == batch
                    ds = dst[:, b0:b0 + nb] if whole else torch.empty((no, nb, t.shape[-1]), dtype=cdt, device=dev_)
                    _FilterPlan.execute_matrix(plans, no, nc, adjoint, xs, ds, nb, self._multi_ws)
                    if not whole:
                        dst[:, b0:b0 + nb] = ds
                if acc is None:
                    acc = dst
                else:
                    acc.add_(dst)
        return out.reshape((n_out,) + lead + (t.shape[-1],))

    def apply_elements(self, field, adjoint=False):
        """For a matrix transfer function (p, q): every element filter applied to the same scalar
        field, T[i][j] = F_ij(E) of shape (p, q, n), or with ``adjoint`` the adjoint element
        filters F_ij^H(E).  This is the matrix filter of E times the identity (E on the diagonal
        of a q x q tensor field) without the filters of its zero components."""
        t = field.as_subclass(torch.Tensor) if isinstance(field, Field) else field
        _lib.require_device(t)
        if not self._is_matrix(t.device):
            raise ValueError('apply_elements needs a matrix transfer function')
        cdt = complex_dtype_for(t.dtype)
        t = t.to(cdt).contiguous()
        tf = self._matrix_tf
        p, q = int(tf.shape[0]), int(tf.shape[1])
        lead = tuple(t.shape[:-1])
        batch = int(np.prod(lead)) if lead else 1
        NY, NX = self.internal_grid.shape
        if not (_pow2(NX) and _pow2(NY)):
            # one forward transform of E, every element's product in one pass (conj(tf) for the
            # adjoint), one batched backward transform
            if self._fft is None:
                self._fft = FastFourierTransform(self.input_grid, self.q)
            F = self._fft.forward(Field(t, self.input_grid)).as_subclass(torch.Tensor)
            n_int = F.shape[-1]
            Fb = F.reshape((1, 1) + tuple(F.shape)).expand((p, q) + tuple(F.shape)).contiguous()
            key = (t.device, cdt, 'elements')
            J = self._tf_cache.get(key)
            if J is None:
                J = tf.to(device=t.device, dtype=cdt).contiguous()
                self._tf_cache[key] = J
            Jb = J.reshape((p, q) + (1,) * len(lead) + (n_int,)).expand(Fb.shape).contiguous()
            G = pointwise.multiply(Fb, Jb, conjugate=adjoint)
            return Field(self._fft.backward(Field(G, self.internal_grid)).as_subclass(torch.Tensor), self.input_grid)
        out = torch.empty((p, q) + tuple(t.shape), dtype=cdt, device=t.device)
        plans = [self._element_plan(cdt, t.device, i, j) for i in range(p) for j in range(q)]
        outs = [out[i, j] for i in range(p) for j in range(q)]
        # at most 8 transfer functions per shared pass
        for k0 in range(0, len(plans), 8):
            _FilterPlan.execute_multi(plans[k0:k0 + 8], adjoint, t, outs[k0:k0 + 8], batch, self._multi_ws)
        return Field(out, self.input_grid)

    def _operation(self, field, adjoint):
        grid = self.input_grid
        if not isinstance(field, torch.Tensor):
            field = Field(field, grid)
        t = field.as_subclass(torch.Tensor) if isinstance(field, Field) else field
        _lib.require_device(t)
        cdt = complex_dtype_for(t.dtype)
        t = t.to(cdt).contiguous()
        if t.requires_grad and torch.is_grad_enabled():
            return Field(_FilterFn.apply(t, self, adjoint), grid)
        return Field(self._run_plain(t, adjoint), grid)

    def _run_plain(self, t, adjoint):
        """Filter (or its adjoint) of a contiguous complex tensor; returns a plain tensor."""
        cdt = t.dtype
        lead = tuple(t.shape[:-1])
        batch = int(np.prod(lead)) if lead else 1
        NY, NX = self.internal_grid.shape
        if self._is_matrix(t.device):
            if not (_pow2(NX) and _pow2(NY)):
                return self._fft_matrix_operation(t, adjoint, cdt).as_subclass(torch.Tensor)
            return self._run_matrix(t, adjoint)
        if _pow2(NX) and _pow2(NY):
            key = (cdt, t.device.index)
            plan = self._plans.get(key)
            if plan is None:
                plan = _FilterPlan(self, cdt, t.device)
                self._plans[key] = plan
            out = torch.empty_like(t)
            plan.execute(adjoint, t, out, batch)
            return out
        return self._fft_operation(t, adjoint, cdt).as_subclass(torch.Tensor)

    def _fft_operation(self, t, adjoint, cdt):
        """Non-power-of-two internal grids: FastFourierTransform onto the internal grid, the
        transfer function (conjugated for the adjoint; evaluated once per device and dtype),
        and the backward transform."""
        if self._fft is None:
            self._fft = FastFourierTransform(self.input_grid, self.q)
        key = (t.device, cdt)
        tf = self._tf_cache.get(key)
        if tf is None:
            tf = self.evaluate_transfer_function(t.device).to(cdt).contiguous()
            self._tf_cache[key] = tf
        F = self._fft.forward(Field(t, self.input_grid))
        F = Field(pointwise.multiply(F, tf, conjugate=adjoint), self.internal_grid)
        return self._fft.backward(F)

    def _fft_matrix_operation(self, t, adjoint, cdt):
        """Matrix transfer function on a non-power-of-two internal grid: each input component
        transformed onto the internal grid, the per-frequency matrix product (field_dot; J^H for
        the adjoint), each output component transformed back."""
        tf = self._matrix_tf
        p, q = int(tf.shape[0]), int(tf.shape[1])
        n_in = p if adjoint else q
        if t.dim() < 2 or t.shape[0] != n_in:
            raise ValueError('a matrix transfer function of shape (%d, %d) needs %d field components, got shape %s'
                             % (p, q, n_in, tuple(t.shape)))
        if self._fft is None:
            self._fft = FastFourierTransform(self.input_grid, self.q)
        key = (t.device, cdt, 'matrix')
        J = self._tf_cache.get(key)
        if J is None:
            J = tf.to(device=t.device, dtype=cdt).contiguous()
            self._tf_cache[key] = J
        F = self._fft.forward(Field(t, self.input_grid))
        G = pointwise.field_dot(J, F.as_subclass(torch.Tensor).contiguous(), adjoint=adjoint)
        return self._fft.backward(Field(G.as_subclass(torch.Tensor).contiguous(), self.internal_grid))

    def forward(self, field):
        return self._operation(field, adjoint=False)

    def backward(self, field):
        return self._operation(field, adjoint=True)
