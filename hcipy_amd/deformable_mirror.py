"""Deformable mirror with Gaussian or Xinetics influence functions
(hcipy/optics/deformable_mirror.py).

The influence functions are never stored as a matrix: ``GaussianInfluenceFunctions`` and
``XineticsInfluenceFunctions`` describe them (actuator grid, shape, tilts) and evaluate linear
combinations and their adjoint with the ``hcipy_dm_gaussian_*`` / ``hcipy_dm_table_*`` kernels,
in FP64.  They offer the parts of the reference's ModeBasis protocol the mirror uses: ``len``,
``grid``, iteration and indexing (one influence function as a Field), ``linear_combination``
and ``is_sparse``.
"""
import ctypes
import math
import os

import numpy as np
import torch

from . import _lib, pointwise
from .field import Field, default_device, make_uniform_grid
from .optics import OpticalElement, Wavefront

__all__ = ['TipTiltMirror', 'make_actuator_positions', 'make_gaussian_influence_functions', 'GaussianInfluenceFunctions',
           'make_xinetics_influence_functions', 'XineticsInfluenceFunctions', 'DeformableMirror', 'PokeSet']


def make_actuator_positions(num_actuators_across_pupil, actuator_spacing, x_tilt=0, y_tilt=0, z_tilt=0):
    """Actuator positions, BMC convention (deformable_mirror.py:11-41)."""
    extent = actuator_spacing * num_actuators_across_pupil
    grid = make_uniform_grid(num_actuators_across_pupil, [extent, extent]).scaled(np.cos([y_tilt, x_tilt]))
    if z_tilt == 0:
        return grid
    return grid.rotated(z_tilt)


class _SurfaceFn(torch.autograd.Function):
    """surface = IF . c with the transposed influence functions as its backward."""

    @staticmethod
    def forward(ctx, c, basis):
        ctx.basis = basis
        return basis._surface_plain(c)

    @staticmethod
    def backward(ctx, g):
        return ctx.basis.linear_combination_adjoint(g.contiguous()), None


class _InfluenceFunctions:
    """Shared ModeBasis-like interface of the kernel-evaluated influence functions."""

    is_sparse = True

    def __len__(self):
        return self.num_actuators_across_pupil ** 2

    @property
    def num_modes(self):
        return len(self)

    def _coeffs(self, coefficients):
        c = coefficients
        if isinstance(c, torch.Tensor):
            c = c.to(device=default_device() if c.device.type != 'cuda' else c.device, dtype=torch.float64)
        else:
            c = torch.as_tensor(np.asarray(c, dtype=np.float64), device=default_device())
        if c.shape[-1] != len(self):
            raise ValueError('Expected %d coefficients, got %d.' % (len(self), c.shape[-1]))
        return c.contiguous()

    def linear_combination(self, coefficients):
        """sum_a c[a] IF_a as an FP64 Field (mode_basis.py:239-258); leading dimensions of the
        coefficients are a batch.  Differentiable with respect to tensor coefficients (the
        backward is ``linear_combination_adjoint``)."""
        c = self._coeffs(coefficients)
        if c.requires_grad and torch.is_grad_enabled():
            return Field(_SurfaceFn.apply(c, self), self.grid)
        return Field(self._surface_plain(c), self.grid)

    def _surface_plain(self, c):
        lead = tuple(c.shape[:-1])
        batch = int(np.prod(lead)) if lead else 1
        out = torch.empty(lead + (self.grid.size,), dtype=torch.float64, device=c.device)
        fn, _ = self._entry_points
        st = getattr(_lib.load(), fn)(ctypes.byref(self._desc), _lib.dptr(c), _lib.dptr(out), batch,
                                      _lib.stream_ptr(c.device))
        _lib.check(st, fn)
        return out

    def linear_combination_adjoint(self, field):
        """g[a] = sum_pixel field[pixel] IF_a(pixel): the transposed influence matrix applied to
        a pupil-plane array (FP64); leading dimensions are a batch."""
        t = field.as_subclass(torch.Tensor) if isinstance(field, Field) else torch.as_tensor(field)
        t = t.to(device=default_device(), dtype=torch.float64).contiguous()
        if t.shape[-1] != self.grid.size:
            raise ValueError('Field size %d does not match the pupil grid size %d' % (t.shape[-1], self.grid.size))
        lead = tuple(t.shape[:-1])
        batch = int(np.prod(lead)) if lead else 1
        out = torch.empty(lead + (len(self),), dtype=torch.float64, device=t.device)
        _, fn = self._entry_points
        st = getattr(_lib.load(), fn)(ctypes.byref(self._desc), _lib.dptr(t), _lib.dptr(out), batch,
                                      _lib.stream_ptr(t.device))
        _lib.check(st, fn)
        return out

    def __getitem__(self, i):
        if not 0 <= i < len(self):
            raise IndexError(i)
        c = torch.zeros(len(self), dtype=torch.float64, device=default_device())
        c[i] = 1.0
        return self.linear_combination(c)

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]

    @property
    def transformation_matrix(self):
        """Dense (pixels x actuators) FP64 matrix, for small configurations."""
        if self.grid.size * len(self) * 8 > (1 << 31):
            raise ValueError('The dense influence matrix would exceed 2 GiB.')
        eye = torch.eye(len(self), dtype=torch.float64, device=default_device())
        return self.linear_combination(eye).as_subclass(torch.Tensor).T




class GaussianInfluenceFunctions(_InfluenceFunctions):
    """Truncated Gaussian influence functions of an n x n actuator grid on a regular 2-D pupil
    grid (make_gaussian_influence_functions, deformable_mirror.py:43-93, with
    make_gaussian_pokes, mode_basis/gaussian_pokes.py:5-54)."""

    _entry_points = ('hcipy_dm_gaussian_surface', 'hcipy_dm_gaussian_surface_adjoint')

    def __init__(self, pupil_grid, num_actuators_across_pupil, actuator_spacing, crosstalk=0.15, cutoff=3,
                 x_tilt=0, y_tilt=0, z_tilt=0, oversampling=1):
        if pupil_grid.ndim != 2 or not pupil_grid.is_regular:
            raise ValueError('Gaussian influence functions need a regular two-dimensional pupil grid.')
        n = int(num_actuators_across_pupil)
        pos = make_actuator_positions(n, actuator_spacing)
        self.grid = pupil_grid
        self.num_actuators_across_pupil = n
        self.actuator_spacing = float(actuator_spacing)
        self.sigma = actuator_spacing / math.sqrt(-2 * math.log(crosstalk))
        # cutoff in units of sigma as make_gaussian_pokes receives it, times sigma
        self.radius = (actuator_spacing / self.sigma * cutoff) * self.sigma
        d = _lib.DmGaussianDesc()
        d.nx, d.ny = (int(v) for v in pupil_grid.dims)
        d.x0, d.y0 = (float(v) for v in pupil_grid.zero)
        d.dx, d.dy = (float(v) for v in pupil_grid.delta)
        d.n_act = n
        d.act_x0, d.act_y0 = (float(v) for v in pos.zero)
        d.pitch = float(pos.delta[0])
        d.sigma = self.sigma
        d.radius = self.radius
        d.x_tilt, d.y_tilt, d.z_tilt = float(x_tilt), float(y_tilt), float(z_tilt)
        d.oversampling = int(round(oversampling))
        self._desc = d

def make_gaussian_influence_functions(pupil_grid, num_actuators_across_pupil, actuator_spacing, crosstalk=0.15,
                                      cutoff=3, x_tilt=0, y_tilt=0, z_tilt=0, oversampling=1):
    """Gaussian influence functions (deformable_mirror.py:43-93): sigma = spacing /
    sqrt(-2 ln crosstalk), truncated at ``cutoff`` actuator spacings."""
    return GaussianInfluenceFunctions(pupil_grid, num_actuators_across_pupil, actuator_spacing, crosstalk, cutoff,
                                      x_tilt, y_tilt, z_tilt, oversampling)


def _read_fits_f64(path):
    """The primary array of a plain FITS file with BITPIX -64 (big-endian FP64), as an array of
    shape (NAXISn, ..., NAXIS1) with the length-1 axes squeezed."""
    with open(path, 'rb') as fh:
        raw = fh.read()
    cards = {}
    pos = 0
    while True:
        block = raw[pos:pos + 2880]
        pos += 2880
        end = False
        for k in range(0, 2880, 80):
            card = block[k:k + 80].decode('ascii', 'replace')
            key = card[:8].strip()
            if key == 'END':
                end = True
                break
            if card[8:10] == '= ':
                cards[key] = card[10:].split('/')[0].strip()
        if end:
            break
    if int(cards['BITPIX']) != -64:
        raise ValueError('expected a BITPIX -64 FITS array')
    naxis = int(cards['NAXIS'])
    shape = tuple(int(cards['NAXIS%d' % (i + 1)]) for i in range(naxis))[::-1]
    n = int(np.prod(shape))
    data = np.frombuffer(raw, dtype='>f8', count=n, offset=pos).astype(np.float64).reshape(shape)
    return np.squeeze(data)


_XINETICS_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data', 'influence_dm5v2.fits')


class XineticsInfluenceFunctions(_InfluenceFunctions):
    """Influence functions of a Xinetics deformable mirror (make_xinetics_influence_functions,
    deformable_mirror.py:95-138): the measured DM5V2 actuator shape (91 x 91 samples at a tenth
    of the actuator spacing, normalised to a peak of 1), bilinearly interpolated at the tilted
    pupil point relative to each actuator, zero outside the sampled extent, times
    cos(y_tilt) / cos(x_tilt)."""

    _entry_points = ('hcipy_dm_table_surface', 'hcipy_dm_table_surface_adjoint')

    def __init__(self, pupil_grid, num_actuators_across_pupil, actuator_spacing, x_tilt=0, y_tilt=0, z_tilt=0):
        if pupil_grid.ndim != 2 or not pupil_grid.is_regular:
            raise ValueError('Xinetics influence functions need a regular two-dimensional pupil grid.')
        n = int(num_actuators_across_pupil)
        pos = make_actuator_positions(n, actuator_spacing)
        self.grid = pupil_grid
        self.num_actuators_across_pupil = n
        self.actuator_spacing = float(actuator_spacing)
        actuator = _read_fits_f64(_XINETICS_TABLE)
        actuator = actuator / actuator.max()
        # the actuator grid: make_uniform_grid(shape, shape * spacing / 10) (x extent from the
        # first axis of the array's shape, as the reference passes it)
        shape = np.array(actuator.shape)
        tgrid = make_uniform_grid(actuator.shape, shape * actuator_spacing / 10.0)
        self._table = torch.as_tensor(np.ascontiguousarray(actuator), device=default_device())
        d = _lib.DmTableDesc()
        d.nx, d.ny = (int(v) for v in pupil_grid.dims)
        d.x0, d.y0 = (float(v) for v in pupil_grid.zero)
        d.dx, d.dy = (float(v) for v in pupil_grid.delta)
        d.n_act = n
        d.act_x0, d.act_y0 = (float(v) for v in pos.zero)
        d.pitch = float(pos.delta[0])
        d.x_tilt, d.y_tilt, d.z_tilt = float(x_tilt), float(y_tilt), float(z_tilt)
        d.table_nx, d.table_ny = (int(v) for v in tgrid.dims)
        d.table_x0, d.table_y0 = (float(v) for v in tgrid.zero)
        d.table_dx, d.table_dy = (float(v) for v in tgrid.delta)
        d.table = _lib.dptr(self._table)
        self._desc = d


def make_xinetics_influence_functions(pupil_grid, num_actuators_across_pupil, actuator_spacing, x_tilt=0, y_tilt=0,
                                      z_tilt=0):
    """Xinetics influence functions (deformable_mirror.py:95-138)."""
    return XineticsInfluenceFunctions(pupil_grid, num_actuators_across_pupil, actuator_spacing, x_tilt, y_tilt, z_tilt)


class DeformableMirror(OpticalElement):
    """A deformable mirror using influence functions (deformable_mirror.py:168-299).

    ``actuators`` is a host float64 array (as in the reference) and may be changed in place;
    the surface is recomputed on the GPU when it changed.  forward multiplies the field by
    exp(2 i k surface), backward by exp(-2 i k surface).
    """

    def __init__(self, influence_functions):
        self.influence_functions = influence_functions
        self.actuators = np.zeros(len(influence_functions))
        self.input_grid = influence_functions.grid
        self._surface = None

    @property
    def influence_functions(self):
        return self._influence_functions

    @influence_functions.setter
    def influence_functions(self, influence_functions):
        self._influence_functions = influence_functions
        self._actuators_for_cached_surface = None

    @property
    def num_actuators(self):
        return self._actuators.size

    @property
    def actuators(self):
        return self._actuators

    @actuators.setter
    def actuators(self, actuators):
        # a torch tensor (for example one that requires grad) is kept as is, anything else
        # becomes a host float64 array as in the reference
        if isinstance(actuators, torch.Tensor):
            self._actuators = actuators
        else:
            self._actuators = np.asarray(actuators, dtype=np.float64)

    @property
    def surface(self):
        """The mirror surface in meters (FP64 Field on the device).  Cached for host actuator
        arrays; recomputed (and differentiable) for tensor actuators."""
        if isinstance(self._actuators, torch.Tensor):
            return self.influence_functions.linear_combination(self._actuators)
        if self._actuators_for_cached_surface is not None and np.array_equal(self._actuators,
                                                                            self._actuators_for_cached_surface):
            return self._surface
        self._surface = self.influence_functions.linear_combination(self._actuators)
        self._actuators_for_cached_surface = self._actuators.copy()
        return self._surface

    @property
    def opd(self):
        return 2 * self.surface

    def phase_for(self, wavelength):
        return self.surface * (4 * np.pi / wavelength)

    def random(self, rms):
        self._actuators = np.random.randn(self._actuators.size) * rms

    def flatten(self):
        self._actuators = np.zeros(len(self.influence_functions))

    def forward(self, wavefront):
        return Wavefront(pointwise.multiply_phase(wavefront.electric_field, self.surface, 2 * wavefront.wavenumber),
                         wavefront.wavelength, wavefront.input_stokes_vector)

    def backward(self, wavefront):
        return Wavefront(pointwise.multiply_phase(wavefront.electric_field, self.surface, -2 * wavefront.wavenumber),
                         wavefront.wavelength, wavefront.input_stokes_vector)


class TipTiltMirror(OpticalElement):
    """A tip-tilt mirror: a deformable mirror whose two modes are the x and y coordinates
    (tip_tilt_mirror.py:6-23), actuators in meters of surface per meter of x / y.  forward
    multiplies the field by exp(2 i k (a0 x + a1 y)), backward by the conjugate; on a regular
    grid the tilt phase is formed in FP64 by the phase-ramp kernel, with no surface array."""

    def __init__(self, input_grid):
        self.input_grid = input_grid
        self.actuators = np.zeros((2,))

    @property
    def num_actuators(self):
        return 2

    @property
    def actuators(self):
        return self._actuators

    @actuators.setter
    def actuators(self, actuators):
        self._actuators = np.asarray(actuators, dtype=np.float64).reshape(2)

    @property
    def surface(self):
        x, y = self.input_grid.coords
        return Field(torch.as_tensor(self._actuators[0] * x + self._actuators[1] * y, device=default_device()),
                     self.input_grid)

    @property
    def opd(self):
        return 2 * self.surface

    def flatten(self):
        self._actuators = np.zeros(2)

    def _apply(self, wavefront, sign):
        E = wavefront.electric_field
        k2 = 2 * sign * wavefront.wavenumber
        if self.input_grid.is_regular and not pointwise.needs_grad(E):
            out = pointwise.phase_ramp(E, self.input_grid, k2 * self._actuators[0], k2 * self._actuators[1])
            out = Field(out, self.input_grid)
        else:
            out = pointwise.multiply_phase(E, self.surface, k2)
        return Wavefront(out, wavefront.wavelength, wavefront.input_stokes_vector)

    def forward(self, wavefront):
        return self._apply(wavefront, 1)

    def backward(self, wavefront):
        return self._apply(wavefront, -1)


class PokeSet:
    """Focal-plane intensities of every single-actuator poke of Gaussian deformable mirrors, for
    several propagation items at once (``hcipy_poke_set_*``): the DM-Jacobian workload of
    broadband wavefront control, where the reference pokes each actuator
    (DeformableMirror.actuators, hcipy/optics/deformable_mirror.py:200-263) and propagates every
    poked wavefront through the per-wavelength Fraunhofer instance
    (hcipy/optics/optical_element.py:297-380).

    ``items``: sequence of ``(propagator, wavelength, field, influence_functions)`` with a
    FraunhoferPropagator, the base pupil field (complex64, or complex128 for FP64) on its input
    grid and GaussianInfluenceFunctions on that grid.  ``poke_height`` is the surface height of
    one poke (meters; the poke multiplies the field by exp(2 i k h IF_a)).  All items share one
    precision (the propagators' complex64 arithmetic, or FP64 for complex128 fields).

    ``execute()`` recomputes, from the current contents of the field tensors (``self.fields``,
    which may be overwritten in place), ``self.base[k]`` = |E_focal|^2 of item k and
    ``self.pokes[k][a]`` = |E_focal|^2 with actuator a poked, and returns both lists.  With
    ``output='field_change'`` ``self.pokes[k][a]`` is instead the change of the complex focal
    field caused by the poke (the Jacobian column of actuator a times the poke height).  Memory
    (workspaces, launch tables) is allocated once here; executions allocate nothing and can be
    captured in HIP graphs.  ``autotune=True`` times every grouped tile shape for the two poke
    launches here and keeps the fastest (``self.tiles``); ``tiles=(stage1, stage2)`` sets them
    instead (``_lib.POKE_TILE_CFGS`` lists the shapes of each precision)."""

    def __init__(self, items, poke_height, device=None, autotune=False, output='intensity', tiles=None):
        self.device = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
        lib = _lib.load()
        items = list(items)
        if not items:
            raise ValueError('a poke set needs at least one item')
        citems = (_lib.PokeItem * len(items))()
        if output not in ('intensity', 'field_change'):
            raise ValueError("output must be 'intensity' or 'field_change'")
        kind = 0 if output == 'intensity' else 1
        self.output = output
        self.fields, self.base, self.pokes, self._keep = [], [], [], []
        for k, (prop, wavelength, field, infl) in enumerate(items):
            if not isinstance(infl, GaussianInfluenceFunctions):
                raise ValueError('item %d: poke sets take Gaussian influence functions' % k)
            d = prop.instance(wavelength)
            mft = d.fourier_transform
            t = field.as_subclass(torch.Tensor) if isinstance(field, torch.Tensor) else torch.as_tensor(np.asarray(field))
            t = t.to(self.device)
            if not t.is_complex():
                t = t.to(torch.complex128)
            t = t.contiguous()
            if t.numel() != mft.input_grid.size:
                raise ValueError('item %d: field of %d samples on a %d-sample pupil grid' % (k, t.numel(), mft.input_grid.size))
            w = np.asarray(mft.input_grid.weights).ravel()
            if not np.all(w == w[0]):
                raise ValueError('item %d: poke sets need scalar pupil-grid weights' % k)
            plan = mft.plan_for(t.dtype, self.device)
            alpha = complex(float(w[0])) * complex(d.norm_factor)
            fp64 = t.dtype == torch.complex128
            odt = torch.float64 if fp64 else torch.float32
            n_focal = mft.output_grid.size
            base = torch.empty(n_focal, dtype=odt, device=self.device)
            pdt = odt if output == 'intensity' else (torch.complex128 if fp64 else torch.complex64)
            pokes = torch.empty((len(infl), n_focal), dtype=pdt, device=self.device)
            k_wave = 2 * np.pi / wavelength
            citems[k] = _lib.PokeItem(plan.handle.value, t.data_ptr(), alpha.real, alpha.imag, infl._desc,
                                      2 * k_wave * float(poke_height), base.data_ptr(), pokes.data_ptr(), kind)
            self.fields.append(t)
            self.base.append(base)
            self.pokes.append(pokes)
            self._keep.append((plan, mft, prop))
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            st = lib.hcipy_poke_set_create(ctypes.byref(h), ctypes.cast(citems, ctypes.c_void_p), len(items),
                                           _lib.stream_ptr(self.device))
        _lib.check(st, 'hcipy_poke_set_create')
        self.handle = h
        n = ctypes.c_size_t()
        _lib.check(lib.hcipy_poke_set_bytes(h, ctypes.byref(n)), 'hcipy_poke_set_bytes')
        self.nbytes = n.value
        npk, fw, ff = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        _lib.check(lib.hcipy_poke_set_info(h, ctypes.byref(npk), ctypes.byref(fw), ctypes.byref(ff)), 'hcipy_poke_set_info')
        self.num_pokes = npk.value
        self.flops_windowed = fw.value  # real flops of one execution as run
        self.flops_full = ff.value      # the same results by full transforms of every poked field
        self.tiles = (_lib.TILES_AUTO, _lib.TILES_AUTO)
        if tiles is not None:
            if autotune:
                raise ValueError('give either autotune=True or tiles, not both')
            t1, t2 = (int(t) for t in tiles)
            _lib.check(lib.hcipy_poke_set_set_tiles(h, t1, t2), 'hcipy_poke_set_set_tiles')
            self.tiles = (t1, t2)
        if autotune:
            chosen = (ctypes.c_int * 2)()
            with torch.cuda.device(self.device):
                _lib.check(lib.hcipy_poke_set_autotune(h, 2, chosen, _lib.stream_ptr(self.device)), 'hcipy_poke_set_autotune')
            self.tiles = (chosen[0], chosen[1])

    def execute(self):
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().hcipy_poke_set_execute(self.handle, _lib.stream_ptr(self.device)),
                       'hcipy_poke_set_execute')
        return self.base, self.pokes

    def close(self):
        if getattr(self, 'handle', None) is not None and _lib._lib is not None:
            _lib.load().hcipy_poke_set_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
