"""Grids and Fields.

Grids carry FP64 coordinate metadata on the host (as in hcipy/field/coordinates.py and
hcipy/field/cartesian_grid.py); Fields are torch tensors on the GPU that carry their grid.

Layout (hcipy/field/field.py:52-65, hcipy/field/coordinates.py:671-675): a Field is a flat
tensor ``(..., Ny*Nx)`` with x fastest; ``grid.shape == (Ny, Nx)``; leading dimensions are
tensor (polarisation / batch) dimensions.
"""
import math
import numbers

import numpy as np
import torch

__all__ = [
    'Field', 'CartesianGrid', 'RegularCoords', 'SeparatedCoords', 'UnstructuredCoords',
    'make_uniform_grid', 'make_pupil_grid', 'make_focal_grid', 'evaluate_supersampled',
    'subsample_field', 'default_device', 'set_default_device', 'field_dtype',
]

_DEFAULT_DEVICE = None


def default_device():
    """Device on which new Fields are created: the current ROCm GPU when present."""
    global _DEFAULT_DEVICE
    if _DEFAULT_DEVICE is None:
        _DEFAULT_DEVICE = torch.device('cuda', torch.cuda.current_device()) if torch.cuda.is_available() else torch.device('cpu')
    return _DEFAULT_DEVICE


def set_default_device(device):
    global _DEFAULT_DEVICE
    _DEFAULT_DEVICE = torch.device(device)


_NP_TO_TORCH = {
    'float32': torch.float32, 'float64': torch.float64, 'complex64': torch.complex64,
    'complex128': torch.complex128, 'float': torch.float64, 'complex': torch.complex128,
    'bool': torch.bool, 'int64': torch.int64, 'int32': torch.int32,
}


def field_dtype(dtype):
    """Maps numpy-style dtype names / numpy dtypes / torch dtypes to torch dtypes."""
    if dtype is None or isinstance(dtype, torch.dtype):
        return dtype
    if dtype is float:
        return torch.float64
    if dtype is complex:
        return torch.complex128
    return _NP_TO_TORCH[np.dtype(dtype).name]


# ----------------------------------------------------------------------------------------------
# Coordinates


class Coords:
    is_regular = False
    is_separated = False
    is_unstructured = False


class RegularCoords(Coords):
    """Regularly spaced coordinates: ``arange(n) * delta + zero`` per axis
    (hcipy/field/coordinates.py:533-675)."""
    is_regular = True
    is_separated = True

    def __init__(self, delta, dims, zero):
        self.dims = tuple(int(n) for n in dims)
        self.delta = np.ones(len(self.dims)) * np.asarray(delta, dtype=float)
        self.zero = np.ones(len(self.dims)) * np.asarray(zero, dtype=float)

    @property
    def separated_coords(self):
        return [np.arange(n, dtype=float) * d + z for d, n, z in zip(self.delta, self.dims, self.zero)]

    @property
    def ndim(self):
        return len(self.dims)

    def scaled(self, s):
        return RegularCoords(self.delta * s, self.dims, self.zero * s)

    def shifted(self, s):
        return RegularCoords(self.delta, self.dims, self.zero + s)

    def key(self):
        return ('regular', tuple(self.delta.tolist()), self.dims, tuple(self.zero.tolist()))


class SeparatedCoords(Coords):
    """Tensor-product coordinates from one 1-D array per axis."""
    is_separated = True

    def __init__(self, separated_coords):
        self._sep = [np.asarray(c, dtype=float).copy() for c in separated_coords]
        self.dims = tuple(len(c) for c in self._sep)

    @property
    def separated_coords(self):
        return [c.copy() for c in self._sep]

    @property
    def ndim(self):
        return len(self.dims)

    def scaled(self, s):
        s = np.ones(self.ndim) * s
        return SeparatedCoords([c * si for c, si in zip(self._sep, s)])

    def shifted(self, s):
        s = np.ones(self.ndim) * s
        return SeparatedCoords([c + si for c, si in zip(self._sep, s)])

    def key(self):
        return ('separated',) + tuple(c.tobytes() for c in self._sep)


class UnstructuredCoords(Coords):
    is_unstructured = True

    def __init__(self, coords):
        self._c = [np.asarray(c, dtype=float).ravel().copy() for c in coords]
        self.dims = (len(self._c[0]),)

    @property
    def ndim(self):
        return len(self._c)

    def scaled(self, s):
        s = np.ones(self.ndim) * s
        return UnstructuredCoords([c * si for c, si in zip(self._c, s)])

    def shifted(self, s):
        s = np.ones(self.ndim) * s
        return UnstructuredCoords([c + si for c, si in zip(self._c, s)])

    def key(self):
        return ('unstructured',) + tuple(c.tobytes() for c in self._c)


# ----------------------------------------------------------------------------------------------
# Grid


class _PolarView:
    def __init__(self, grid):
        self._g = grid

    @property
    def r(self):
        x, y = self._g.coords
        return Field(np.hypot(x, y), self._g)

    @property
    def theta(self):
        x, y = self._g.coords
        return Field(np.arctan2(y, x), self._g)


class CartesianGrid:
    """A Cartesian grid with regular, separated or unstructured coordinates.

    Mirrors the parts of hcipy's Grid/CartesianGrid used by the propagation path:
    ``dims``, ``shape``, ``size``, ``ndim``, ``delta``, ``zero``, ``separated_coords``,
    ``coords``, ``x``, ``y``, ``weights``, ``scaled``, ``shifted``, ``is_regular``,
    ``is_separated``, ``is_('cartesian')``, ``as_('polar').r``, ``zeros``/``ones``/``empty``.
    """

    def __init__(self, coords, weights=None):
        self.coords_obj = coords
        self._weights = weights

    # --- structure
    @property
    def is_regular(self):
        return self.coords_obj.is_regular

    @property
    def is_separated(self):
        return self.coords_obj.is_separated

    @property
    def is_unstructured(self):
        return self.coords_obj.is_unstructured

    def is_(self, system):
        return system == 'cartesian'

    def as_(self, system):
        if system == 'cartesian':
            return self
        if system == 'polar':
            return _PolarView(self)
        raise ValueError('Unknown coordinate system %s' % system)

    @property
    def ndim(self):
        return self.coords_obj.ndim

    @property
    def dims(self):
        return np.asarray(self.coords_obj.dims)

    @property
    def shape(self):
        if self.is_unstructured:
            return (self.size,)
        return tuple(reversed(self.coords_obj.dims))

    @property
    def size(self):
        return int(math.prod(self.coords_obj.dims))

    @property
    def delta(self):
        return self.coords_obj.delta

    @property
    def zero(self):
        return self.coords_obj.zero

    @property
    def separated_coords(self):
        return self.coords_obj.separated_coords

    @property
    def coords(self):
        if self.is_unstructured:
            return [c.copy() for c in self.coords_obj._c]
        sep = self.separated_coords
        if self.ndim == 1:
            return [sep[0]]
        g = np.meshgrid(*sep)
        return [c.ravel() for c in g]

    @property
    def points(self):
        return np.array(self.coords).T

    @property
    def x(self):
        return self.coords[0]

    @property
    def y(self):
        return self.coords[1]

    @property
    def weights(self):
        """Scalar for regular grids (hcipy/field/cartesian_grid.py:171-172), midpoint rule
        per axis for separated grids (:174-189)."""
        if self._weights is not None:
            return self._weights
        if self.is_regular:
            return float(np.prod(self.delta))
        if self.is_separated:
            w = None
            for i, x in enumerate(self.separated_coords):
                wi = np.concatenate(([x[1] - x[0]], (x[2:] - x[:-2]) / 2, [x[-1] - x[-2]]))
                shape = (1,) * (self.ndim - i - 1) + (-1,) + (1,) * i
                wi = wi.reshape(shape)
                w = wi if w is None else w * wi
            return w.ravel()
        return 1.0

    @weights.setter
    def weights(self, w):
        self._weights = w

    # --- transformations (hcipy/field/cartesian_grid.py:83-104)
    def scaled(self, scale):
        w = self.weights
        if np.isscalar(scale):
            w = w * abs(float(scale)) ** self.ndim
        else:
            w = w * np.prod(np.abs(scale))
        return CartesianGrid(self.coords_obj.scaled(scale), w)

    def shifted(self, shift):
        return CartesianGrid(self.coords_obj.shifted(shift), self._weights)

    def rotated(self, angle):
        """2-D rotation of the points by ``angle`` (R = [[cos, -sin], [sin, cos]] applied to the
        points, hcipy/field/cartesian_grid.py:12-32,144-165); the result is unstructured."""
        if self.ndim != 2:
            raise ValueError('Only two-dimensional grids can be rotated here.')
        x, y = self.coords
        c, s = math.cos(angle), math.sin(angle)
        return CartesianGrid(UnstructuredCoords([c * x - s * y, s * x + c * y]), self._weights)

    def copy(self):
        return CartesianGrid(self.coords_obj, self._weights)

    def closest_to(self, p):
        c = np.array(self.coords)
        return int(np.argmin(np.sum((c - np.asarray(p, dtype=float)[:, None]) ** 2, axis=0)))

    # --- Field factories
    def zeros(self, tensor_shape=None, dtype=None, device=None):
        return self._new(torch.zeros, tensor_shape, dtype, device)

    def ones(self, tensor_shape=None, dtype=None, device=None):
        return self._new(torch.ones, tensor_shape, dtype, device)

    def empty(self, tensor_shape=None, dtype=None, device=None):
        return self._new(torch.empty, tensor_shape, dtype, device)

    def _new(self, fn, tensor_shape, dtype, device):
        ts = () if tensor_shape is None else tuple(np.atleast_1d(tensor_shape).tolist())
        t = fn(ts + (self.size,), dtype=field_dtype(dtype) or torch.float64, device=device or default_device())
        return Field(t, self)

    # --- identity
    def _key(self):
        w = self._weights
        wk = None if w is None else (w if np.isscalar(w) else np.asarray(w).tobytes())
        return self.coords_obj.key() + (wk,)

    def __eq__(self, other):
        return isinstance(other, CartesianGrid) and self._key() == other._key()

    def __hash__(self):
        return hash(self._key())

    def __repr__(self):
        return 'CartesianGrid(%s, dims=%s)' % (type(self.coords_obj).__name__, tuple(self.coords_obj.dims))


def make_uniform_grid(dims, extent, center=0, has_center=False):
    """hcipy/field/util.py:68-125."""
    dims = np.atleast_1d(dims)
    extent = np.atleast_1d(np.asarray(extent, dtype=float))
    n = max(len(dims), len(extent), np.size(center))
    dims = np.ones(n, dtype=int) * dims
    extent = np.ones(n) * extent
    center = np.ones(n) * center
    delta = extent / dims
    zero = -extent / 2 + center + delta / 2
    if has_center:
        zero = zero - delta / 2 * (1 - dims % 2)
    return CartesianGrid(RegularCoords(delta, dims, zero))


def make_pupil_grid(dims, diameter=1):
    """hcipy/field/util.py:127-163."""
    return make_uniform_grid(np.ones(2, dtype=int) * dims, np.ones(2) * diameter)


def make_focal_grid(q, num_airy, spatial_resolution=None, pupil_diameter=None, focal_length=None,
                    f_number=None, reference_wavelength=None):
    """hcipy/field/util.py:222-318."""
    if spatial_resolution is None:
        if f_number is None:
            if pupil_diameter is None or focal_length is None:
                if reference_wavelength is None:
                    spatial_resolution = 1.0
                else:
                    raise ValueError('You only supplied a reference wavelength and forgot to supply either an f_number or a (pupil_diameter, focal_length).')
            else:
                f_number = np.asarray(focal_length) / np.asarray(pupil_diameter)
        if spatial_resolution is None:
            if reference_wavelength is None:
                raise ValueError('You supplied an f_number or (pupil_diameter, focal_length), and forgot to supply a reference wavelength.')
            spatial_resolution = f_number * reference_wavelength
    q = np.ones(2) * q
    num_airy = np.ones(2) * num_airy
    sr = np.ones(2) * spatial_resolution
    delta = sr / q
    dims = tuple(int(2 * n * qi) for n, qi in zip(num_airy, q))
    dims_arr = np.asarray(dims, dtype=float)
    zero = delta * (-dims_arr / 2 + (dims_arr % 2) * 0.5)
    return CartesianGrid(RegularCoords(delta, dims, zero))


# ----------------------------------------------------------------------------------------------
# Field


def _find_grid(args):
    for a in args:
        if isinstance(a, Field):
            g = a.__dict__.get('grid')
            if g is not None:
                return g
        elif isinstance(a, (list, tuple)):
            g = _find_grid(a)
            if g is not None:
                return g
    return None


def _wrap(ret, grid):
    if isinstance(ret, torch.Tensor):
        if grid is not None and ret.dim() >= 1 and ret.shape[-1] == grid.size:
            out = ret.as_subclass(Field)
            out.grid = grid
            return out
        return ret.as_subclass(torch.Tensor) if isinstance(ret, Field) else ret
    if isinstance(ret, tuple):
        return tuple(_wrap(r, grid) for r in ret)
    if isinstance(ret, list):
        return [_wrap(r, grid) for r in ret]
    return ret


def _as_tensor(data, device=None):
    if isinstance(data, torch.Tensor):
        t = data.as_subclass(torch.Tensor) if isinstance(data, Field) else data
        return t if device is None else t.to(device)
    arr = np.asarray(data)
    if arr.dtype == np.bool_:
        arr = arr.astype(float)
    return torch.as_tensor(np.ascontiguousarray(arr), device=device or default_device())


class Field(torch.Tensor):
    """A torch tensor of shape ``(..., grid.size)`` that carries its Grid.

    Results of torch operations keep the grid when their last dimension still has
    ``grid.size`` elements (the reference's Field keeps it through ufuncs,
    hcipy/field/field.py:1232).
    """

    @staticmethod
    def __new__(cls, data, grid=None, device=None):
        t = _as_tensor(data, device)
        f = t.as_subclass(cls)
        f.grid = grid if grid is not None else getattr(data, 'grid', None)
        return f

    def __init__(self, data, grid=None, device=None):
        pass

    @classmethod
    def __torch_function__(cls, func, types, args=(), kwargs=None):
        if kwargs is None:
            kwargs = {}
        grid = _find_grid(args)
        if grid is None and kwargs:
            grid = _find_grid(tuple(kwargs.values()))
        with torch._C.DisableTorchFunctionSubclass():
            ret = func(*args, **kwargs)
        if func in _NO_WRAP:
            return ret
        return _wrap(ret, grid)

    # --- HCIPy Field API
    @property
    def tensor_order(self):
        return self.dim() - 1

    @property
    def tensor_shape(self):
        return tuple(self.shape[:-1])

    @property
    def is_scalar_field(self):
        return self.dim() == 1

    @property
    def is_vector_field(self):
        return self.dim() == 2

    @property
    def shaped(self):
        return self.as_subclass(torch.Tensor).reshape(self.tensor_shape + tuple(self.grid.shape))

    def astype(self, dtype, copy=True):
        return Field(self.as_subclass(torch.Tensor).to(field_dtype(dtype), copy=copy), self.grid)

    def numpy(self):
        return self.as_subclass(torch.Tensor).detach().cpu().numpy()

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a if dtype is None else a.astype(dtype)

    def __reduce_ex__(self, proto):
        return (_rebuild_field, (self.numpy(), self.grid))


def _rebuild_field(arr, grid):
    return Field(arr, grid)


_NO_WRAP = {torch.Tensor.__repr__, torch.Tensor.__format__, torch.Tensor.numpy, torch.Tensor.tolist,
            torch.Tensor.item, torch.Tensor.data_ptr, torch.Tensor.__len__, torch.Tensor.dim}


# ----------------------------------------------------------------------------------------------
# Supersampling (hcipy/field/util.py:525-698)


def evaluate_supersampled(field_generator, grid, oversampling, statistic='mean'):
    """Mean (or sum) of ``field_generator`` over dithered copies of a separated grid
    (hcipy/field/util.py:606-698, dithers from ``make_uniform_grid(oversampling, 1)``)."""
    if not grid.is_separated:
        raise ValueError('evaluate_supersampled requires a separated grid.')
    ov = np.ones(grid.ndim, dtype=int) * int(round(oversampling)) if np.isscalar(oversampling) else np.round(oversampling).astype(int)
    dithers = make_uniform_grid(ov, 1)
    sep = grid.separated_coords
    deltas = []
    for x in sep:
        d = (x[2:] - x[:-2]) / 2.
        deltas.append(np.concatenate(([x[1] - x[0]], d, [x[-1] - x[-2]])))
    acc = None
    for p in dithers.points:
        g = CartesianGrid(SeparatedCoords([c + float(p[i]) * dl for i, (c, dl) in enumerate(zip(sep, deltas))]))
        f = field_generator(g)
        f = f.as_subclass(torch.Tensor) if isinstance(f, torch.Tensor) else torch.as_tensor(np.asarray(f), device=default_device())
        acc = f.clone() if acc is None else acc + f
    if statistic == 'mean':
        acc = acc / len(dithers.points)
    elif statistic != 'sum':
        raise ValueError('Unsupported statistic %s' % statistic)
    return Field(acc, grid)


def subsample_field(field, subsampling, new_grid=None, statistic='mean'):
    """Block mean over ``subsampling`` pixels per axis (hcipy/field/util.py:525-604)."""
    sub = int(round(subsampling))
    g = field.grid
    ny, nx = g.shape
    if new_grid is None:
        new_grid = make_uniform_grid([nx // sub, ny // sub], g.delta * g.dims, g.zero + g.delta * g.dims / 2 - g.delta / 2)
    t = field.as_subclass(torch.Tensor)
    lead = t.shape[:-1]
    r = t.reshape(lead + (ny // sub, sub, nx // sub, sub))
    if statistic == 'mean':
        r = r.mean(dim=(-3, -1))
    else:
        r = r.sum(dim=(-3, -1))
    return Field(r.reshape(lead + (-1,)), new_grid)
