"""Wavefronts and optical elements (hcipy/optics/wavefront.py, hcipy/optics/optical_element.py,
hcipy/optics/apodization.py).

Protocol: ``OpticalElement.forward(wavefront) -> Wavefront``, ``backward`` the adjoint-style
inverse, ``__call__ = forward`` (optical_element.py:15-66).
"""
import copy
import math
from collections import OrderedDict

import numpy as np
import torch

from . import pointwise
from .field import Field, field_dtype

__all__ = ['Wavefront', 'OpticalElement', 'AgnosticOpticalElement', 'OpticalSystem', 'Apodizer', 'PhaseApodizer',
           'SurfaceApodizer', 'ComplexSurfaceApodizer']


def _tensor(f):
    return f.as_subclass(torch.Tensor) if isinstance(f, Field) else f


class Wavefront:
    """Electric field with a wavelength (wavefront.py:15-360): scalar, a Jones-vector field, or a
    partially polarized wavefront, i.e. the 2 x 2 Jones-matrix field acting on an input Stokes
    vector (I, Q, U, V) (wavefront.py:44-75).

    The electric field is cast to complex64 when it is float32/complex64 and to complex128
    otherwise (wavefront.py:77-88).  With a Stokes vector, a scalar field E becomes E times the
    2 x 2 identity; a field of shape (2, 2, size) is taken as the Jones matrix itself.
    """

    def __init__(self, electric_field, wavelength=1, input_stokes_vector=None):
        self.electric_field = electric_field
        if input_stokes_vector is not None:
            t = _tensor(self._electric_field)
            if t.dim() == 1:
                eye = torch.eye(2, dtype=t.dtype, device=t.device)
                self._electric_field = Field((eye[:, :, None] * t[None, None, :]).contiguous(), self.grid)
            elif tuple(t.shape[:-1]) != (2, 2):
                raise ValueError('When supplying a Stokes vector, the electric field must be either a scalar or 2-tensor field.')
            input_stokes_vector = np.asarray(input_stokes_vector, dtype=float).reshape(4)
        self.wavelength = wavelength
        self._input_stokes_vector = input_stokes_vector

    def copy(self):
        return Wavefront(Field(_tensor(self._electric_field).clone(), self._electric_field.grid), self.wavelength,
                         self._input_stokes_vector)

    @property
    def electric_field(self):
        return self._electric_field

    @electric_field.setter
    def electric_field(self, electric_field):
        if not hasattr(electric_field, 'grid'):
            raise ValueError('The electric field must be a Field.')
        t = _tensor(electric_field) if isinstance(electric_field, torch.Tensor) else Field(electric_field, electric_field.grid).as_subclass(torch.Tensor)
        cdt = torch.complex64 if t.dtype in (torch.float32, torch.complex64) else torch.complex128
        self._electric_field = Field(t.to(cdt), electric_field.grid)

    @property
    def input_stokes_vector(self):
        return self._input_stokes_vector

    @property
    def wavenumber(self):
        return 2 * np.pi / self.wavelength

    @wavenumber.setter
    def wavenumber(self, k):
        self.wavelength = 2 * np.pi / k

    @property
    def grid(self):
        return self._electric_field.grid

    @property
    def is_scalar(self):
        return self._electric_field.is_scalar_field

    @property
    def is_partially_polarized(self):
        return self._input_stokes_vector is not None

    @property
    def is_polarized(self):
        return self._input_stokes_vector is not None or _tensor(self._electric_field).dim() == 2

    def _stokes(self, component):
        t = _tensor(self._electric_field)
        if self._input_stokes_vector is not None:
            return pointwise.stokes(self._electric_field, self._input_stokes_vector, component)
        if t.dim() == 2 and t.shape[0] == 2:
            return pointwise.stokes(self._electric_field, None, component)
        raise ValueError('Stokes parameters need a scalar, Jones-vector (2, n) or Jones-matrix (2, 2, n) field.')

    @property
    def I(self):  # noqa: E743 (HCIPy name)
        """Intensity (wavefront.py:113-143): |E|^2 for a scalar field, and the Stokes I of a
        Jones-vector field or a partially polarized wavefront (``hcipy_pw_stokes``).  Other tensor
        fields sum |E|^2 over their components."""
        t = _tensor(self._electric_field)
        if self._input_stokes_vector is not None or (t.dim() == 2 and t.shape[0] == 2):
            return self._stokes(0)
        I = pointwise.intensity(self._electric_field)
        if I.dim() > 1:
            I = Field(_tensor(I).sum(dim=tuple(range(I.dim() - 1))), self.grid)
        return I

    def _zeros(self):
        t = _tensor(self._electric_field)
        rdt = torch.float32 if t.dtype == torch.complex64 else torch.float64
        return Field(torch.zeros(t.shape[-1], dtype=rdt, device=t.device), self.grid)

    @property
    def Q(self):  # noqa: N802 (HCIPy name)
        """Stokes Q (wavefront.py:145-173); zero for a scalar field."""
        return self._zeros() if self.is_scalar else self._stokes(1)

    @property
    def U(self):  # noqa: N802 (HCIPy name)
        """Stokes U (wavefront.py:175-203); zero for a scalar field."""
        return self._zeros() if self.is_scalar else self._stokes(2)

    @property
    def V(self):  # noqa: N802 (HCIPy name)
        """Stokes V (wavefront.py:205-232); zero for a scalar field."""
        return self._zeros() if self.is_scalar else self._stokes(3)

    @property
    def stokes_vector(self):
        """(4, n) Stokes vector (wavefront.py:234-259); a scalar field is unpolarized (I, 0, 0, 0)."""
        if self.is_scalar:
            I = _tensor(self.I)
            z = torch.zeros_like(I)
            return Field(torch.stack([I, z, z, z]), self.grid)
        return self._stokes(-1)

    @property
    def intensity(self):
        return self.I

    @property
    def amplitude(self):
        return Field(torch.abs(_tensor(self._electric_field)), self.grid)

    @property
    def phase(self):
        return Field(torch.angle(_tensor(self._electric_field)), self.grid)

    @property
    def power(self):
        """Intensity times grid weights (wavefront.py:310-330)."""
        return self.I * torch.as_tensor(np.asarray(self.grid.weights), device=self._electric_field.device)

    @property
    def total_power(self):
        """FP64 sum of intensity x weights (wavefront.py:332-360)."""
        if self._input_stokes_vector is not None:
            return pointwise.real_sum(self.I) * float(np.asarray(self.grid.weights).ravel()[0])
        return pointwise.total_power(self._electric_field, self.grid.weights)

    @total_power.setter
    def total_power(self, p):
        cur = self.total_power
        if cur == 0:
            raise ValueError('Cannot set the total power of a wavefront with zero power.')
        E = _tensor(self._electric_field) * math.sqrt(p / cur)
        self._electric_field = Field(E, self.grid)


class OpticalElement:
    """Base class (optical_element.py:6-151)."""

    def __call__(self, wavefront):
        return self.forward(wavefront)

    def forward(self, wavefront):
        raise NotImplementedError()

    def backward(self, wavefront):
        raise NotImplementedError()


def wavelength_key(wavelength):
    """Cache key of a wavelength: round(log(wl) / log(1 + 1e-9)) (optical_element.py:333)."""
    return int(round(math.log(wavelength) / math.log(1 + 1e-9)))


class AgnosticOpticalElement(OpticalElement):
    """Element whose per-(grid, wavelength) instance data is built on first use and kept in a
    FIFO cache of at most ``max_in_cache`` entries (optical_element.py:261-380, 600-644).

    The default keeps 64 instances where the reference keeps 11: instance data (Fourier
    matrices, filter plans) lives in device memory, and a FIFO of 11 under a cycle of more
    wavelengths (32-wavelength broadband or Jacobian runs) misses on every access, rebuilding
    every instance on every pass.  64 instances of a 1024^2 -> 512^2 propagator hold about
    1 GB."""

    def __init__(self, grid_dependent=True, wavelength_dependent=True, max_in_cache=64):
        self.grid_dependent = grid_dependent
        self.wavelength_dependent = wavelength_dependent
        self.max_in_cache = max_in_cache
        self._cache = OrderedDict()

    class _InstanceData:
        pass

    def clear_cache(self):
        self._cache = OrderedDict()

    def get_instance_data(self, input_grid, output_grid, wavelength):
        if input_grid is None:
            input_grid = self.get_input_grid(output_grid, wavelength)
        if output_grid is None:
            output_grid = self.get_output_grid(input_grid, wavelength)
        key = (input_grid if self.grid_dependent else None,
               output_grid if self.grid_dependent else None,
               wavelength_key(wavelength) if self.wavelength_dependent else None)
        data = self._cache.get(key)
        if data is None:
            data = AgnosticOpticalElement._InstanceData()
            data.input_grid = input_grid
            data.output_grid = output_grid
            data.wavelength = wavelength
            self.make_instance(data, input_grid, output_grid, wavelength)
            self._cache[key] = data
            while len(self._cache) > self.max_in_cache:
                self._cache.popitem(last=False)
        return data

    def evaluate_parameter(self, parameter, input_grid, output_grid, wavelength):
        return parameter(wavelength) if callable(parameter) else parameter

    def get_input_grid(self, output_grid, wavelength):
        return output_grid

    def get_output_grid(self, input_grid, wavelength):
        return input_grid

    def make_instance(self, instance_data, input_grid, output_grid, wavelength):
        raise NotImplementedError()


class OpticalSystem(OpticalElement):
    """Sequence of elements; backward runs the elements' backward in reverse
    (optical_element.py:848-949)."""

    def __init__(self, optical_elements):
        self.optical_elements = list(optical_elements)

    def forward(self, wavefront):
        wf = wavefront
        for el in self.optical_elements:
            wf = el.forward(wf)
        return wf

    def backward(self, wavefront):
        wf = wavefront
        for el in reversed(self.optical_elements):
            wf = el.backward(wf)
        return wf


class Apodizer(OpticalElement):
    """E * a forward, E * conj(a) backward (apodization.py:6-35)."""

    def __init__(self, apodization):
        self.apodization = apodization

    def _get_apodization(self, wavelength):
        return self.apodization(wavelength) if callable(self.apodization) and not isinstance(self.apodization, torch.Tensor) else self.apodization

    def forward(self, wavefront):
        a = self._get_apodization(wavefront.wavelength)
        return Wavefront(pointwise.multiply(wavefront.electric_field, a, conjugate=False), wavefront.wavelength, wavefront.input_stokes_vector)

    def backward(self, wavefront):
        a = self._get_apodization(wavefront.wavelength)
        return Wavefront(pointwise.multiply(wavefront.electric_field, a, conjugate=True), wavefront.wavelength, wavefront.input_stokes_vector)


class SurfaceApodizer(OpticalElement):
    """Thin sagged plate: E exp(i (n - 1) k sag) forward, E exp(-i (n - 1) k sag) backward, with
    the refractive index n a scalar or a function of wavelength (apodization.py:64-102).  The
    phase is formed in FP64 by the phase kernel."""

    def __init__(self, surface_sag, refractive_index):
        self.surface_sag = surface_sag
        self.refractive_index = refractive_index

    def _get_refractive_index(self, wavelength):
        return self.refractive_index(wavelength) if callable(self.refractive_index) else self.refractive_index

    def _scale(self, wavefront):
        return float((self._get_refractive_index(wavefront.wavelength) - 1) * wavefront.wavenumber)

    def forward(self, wavefront):
        E = pointwise.multiply_phase(wavefront.electric_field, self.surface_sag, self._scale(wavefront))
        return Wavefront(E, wavefront.wavelength, wavefront.input_stokes_vector)

    def backward(self, wavefront):
        E = pointwise.multiply_phase(wavefront.electric_field, self.surface_sag, -self._scale(wavefront))
        return Wavefront(E, wavefront.wavelength, wavefront.input_stokes_vector)


class ComplexSurfaceApodizer(SurfaceApodizer):
    """Sagged plate with an amplitude coating: E a exp(i (n - 1) k surface) forward, E conj(a)
    exp(-i (n - 1) k surface) backward; a may be a function of wavelength
    (apodization.py:104-150)."""

    def __init__(self, amplitude, surface, refractive_index):
        SurfaceApodizer.__init__(self, surface, refractive_index)
        self.amplitude = amplitude

    @property
    def surface(self):
        return self.surface_sag

    def _get_amplitude(self, wavelength):
        a = self.amplitude
        return a(wavelength) if callable(a) and not isinstance(a, torch.Tensor) else a

    def forward(self, wavefront):
        E = pointwise.multiply(wavefront.electric_field, self._get_amplitude(wavefront.wavelength))
        E = pointwise.multiply_phase(E, self.surface_sag, self._scale(wavefront))
        return Wavefront(E, wavefront.wavelength, wavefront.input_stokes_vector)

    def backward(self, wavefront):
        E = pointwise.multiply(wavefront.electric_field, self._get_amplitude(wavefront.wavelength), conjugate=True)
        E = pointwise.multiply_phase(E, self.surface_sag, -self._scale(wavefront))
        return Wavefront(E, wavefront.wavelength, wavefront.input_stokes_vector)


class PhaseApodizer(OpticalElement):
    """E * exp(i phi) forward, E * exp(-i phi) backward (apodization.py:37-62)."""

    def __init__(self, phase):
        self._phase = phase

    @property
    def phase(self):
        return self._phase

    def forward(self, wavefront):
        return Wavefront(pointwise.multiply_phase(wavefront.electric_field, self._phase, 1.0), wavefront.wavelength, wavefront.input_stokes_vector)

    def backward(self, wavefront):
        return Wavefront(pointwise.multiply_phase(wavefront.electric_field, self._phase, -1.0), wavefront.wavelength, wavefront.input_stokes_vector)
