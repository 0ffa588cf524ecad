"""Pyramid wavefront-sensor optics on the device MFT and pointwise kernels
(hcipy/wavefront_sensing/pyramid.py:10-200).

The optics are pupil -> focal plane (Fraunhofer MFT), a circular spatial filter, the pyramid
as a refractive surface -(sep / (2 (n - 1))) (|x| + |y|), focal plane -> detector pupil
(Fraunhofer MFT).  The modulated sensor runs all modulation steps as one batch: the tilted
copies of the input (phase-ramp kernel) or the shifted pyramid surfaces (one phase operand per
step) go through one batched MFT per plane instead of one propagation per step.
"""
import numpy as np
import torch

from . import pointwise
from .aperture import make_circular_aperture
from .deformable_mirror import TipTiltMirror
from .field import CartesianGrid, Field, UnstructuredCoords, evaluate_supersampled, make_pupil_grid
from .optics import Apodizer, OpticalElement, SurfaceApodizer, Wavefront
from .propagation import FraunhoferPropagator

__all__ = ['PyramidWavefrontSensorOptics', 'ModulatedPyramidWavefrontSensorOptics']


def _tensor(f):
    return f.as_subclass(torch.Tensor) if isinstance(f, Field) else f


class PyramidWavefrontSensorOptics(OpticalElement):
    """The optics of a pyramid wavefront sensor (pyramid.py:95-200): same arguments, defaults and
    ``ValueError`` conditions.  ``precision`` selects the complex64 MFT arithmetic."""

    def __init__(self, input_grid, output_grid, separation=None, pupil_diameter=None, wavelength_0=1, q=None,
                 num_airy=None, refractive_index=lambda x: 1.5, precision=None):
        if not input_grid.is_regular:
            raise ValueError('The input grid must be a regular grid.')
        self.input_grid = input_grid
        self.output_grid = output_grid
        extent_in = np.asarray(input_grid.delta, dtype=float) * np.asarray(input_grid.dims)
        if pupil_diameter is None:
            pupil_diameter = np.max(extent_in)
        if separation is None:
            separation = pupil_diameter
        # at least 2 focal samples per resolution element, and enough to sample the output
        extent_out = np.asarray(output_grid.delta, dtype=float) * np.asarray(output_grid.dims)
        qmin = np.ceil(max(np.max(extent_out / extent_in), 2))
        if q is None:
            q = qmin
        elif q < qmin:
            raise ValueError('The requested focal plane sampling is too low to sufficiently sample the wavefront sensor output.')
        self.num_airy = np.max(input_grid.shape) / 2 if num_airy is None else num_airy
        num_pixels = 2 * int(self.num_airy * q)
        spatial_resolution = wavelength_0 / pupil_diameter
        self.focal_grid = make_pupil_grid(num_pixels, 2 * spatial_resolution * self.num_airy)
        self._separation = separation
        self._refractive_index = refractive_index
        self._wavelength_0 = wavelength_0
        self.spatial_filter = Apodizer(make_circular_aperture(2 * self.num_airy * wavelength_0 / pupil_diameter)(self.focal_grid))
        self.pyramid = SurfaceApodizer(Field(self._surface(self.focal_grid), self.focal_grid), refractive_index)
        self.pupil_to_focal = FraunhoferPropagator(input_grid, self.focal_grid, precision=precision)
        self.focal_to_pupil = FraunhoferPropagator(self.focal_grid, output_grid, precision=precision)

    def _surface(self, grid):
        x, y = grid.coords
        n0 = self._refractive_index(self._wavelength_0)
        return -self._separation / (2 * (n0 - 1)) * (np.abs(x) + np.abs(y))

    def forward(self, wavefront):
        wf = self.pupil_to_focal.forward(wavefront)
        wf = self.pyramid.forward(self.spatial_filter.forward(wf))
        return self.focal_to_pupil.forward(wf)

    def backward(self, wavefront):
        wf = self.focal_to_pupil.backward(wavefront)
        wf = self.pyramid.backward(self.spatial_filter.backward(wf))
        return self.pupil_to_focal.backward(wf)


class ModulatedPyramidWavefrontSensorOptics(OpticalElement):
    """A modulated pyramid wavefront sensor (pyramid.py:10-93): ``num_steps`` tip-tilt positions
    on a circle of radius ``modulation / 2``.  The slow method tilts the input with a tip-tilt
    mirror per step; the fast method shifts the pyramid surface instead (surfaces precomputed,
    supersampled 4x).  forward returns one Wavefront per step, as the reference does; all
    steps are propagated as one batch."""

    def __init__(self, pyramid_wavefront_sensor, modulation, num_steps=12, fast_modulation_method=False):
        self.modulation = modulation
        self.pyramid_wavefront_sensor = pyramid_wavefront_sensor
        self.tip_tilt_mirror = TipTiltMirror(pyramid_wavefront_sensor.input_grid)
        self.focal_grid = pyramid_wavefront_sensor.focal_grid
        self.pupil_to_focal = pyramid_wavefront_sensor.pupil_to_focal
        self.focal_to_pupil = pyramid_wavefront_sensor.focal_to_pupil
        self._fast_modulation_method = fast_modulation_method
        theta = np.linspace(0, 2 * np.pi, num_steps, endpoint=False)
        x_mod = modulation / 2 * np.cos(theta)
        y_mod = modulation / 2 * np.sin(theta)
        self.modulation_positions = CartesianGrid(UnstructuredCoords((x_mod, y_mod)))
        self._points = np.stack([x_mod, y_mod], axis=1)
        self._surfaces = None
        if fast_modulation_method:
            pws = pyramid_wavefront_sensor
            # the factor 2 compensates the 1/2 of the tilt mirror's surface
            surfs = [evaluate_supersampled(lambda g: Field(pws._surface(g), g), self.focal_grid.shifted(2 * p), 4)
                     for p in self._points]
            self._surfaces = torch.stack([_tensor(s) for s in surfs]).contiguous()

    @staticmethod
    def _split(wf, input_stokes_vector):
        E = _tensor(wf.electric_field)
        return [Wavefront(Field(E[i], wf.electric_field.grid), wf.wavelength, input_stokes_vector)
                for i in range(E.shape[0])]

    def forward(self, wavefront):
        pws = self.pyramid_wavefront_sensor
        E = _tensor(wavefront.electric_field)
        steps = self._points.shape[0]
        if self._fast_modulation_method:
            foc = pws.spatial_filter.forward(self.pupil_to_focal.forward(wavefront))
            F = _tensor(foc.electric_field)
            Fb = F.unsqueeze(0).expand((steps,) + tuple(F.shape)).contiguous()
            S = self._surfaces
            if F.dim() > 1:
                # one surface per step, shared by the fields of a batched input
                S = S.reshape((steps,) + (1,) * (F.dim() - 1) + (F.shape[-1],)).expand(Fb.shape).contiguous()
            n = pws._refractive_index(wavefront.wavelength)
            G = pointwise.multiply_phase(Fb, S, (n - 1) * wavefront.wavenumber)
            # the steps travel as one batch without the Stokes vector, which is restored per step
            out = self.focal_to_pupil.forward(Wavefront(Field(G, self.focal_grid), wavefront.wavelength))
        else:
            T = E.unsqueeze(0).expand((steps,) + tuple(E.shape)).contiguous()
            k2 = 2 * wavefront.wavenumber
            grid = pws.input_grid
            for i, (px, py) in enumerate(self._points):
                # in place on row i: the step's tilt exp(2 i k (px x + py y))
                pointwise.phase_ramp(T[i], grid, k2 * px, k2 * py, out=T[i])
            out = pws.forward(Wavefront(Field(T, grid), wavefront.wavelength))
        return self._split(out, wavefront.input_stokes_vector)

    def backward(self, wavefront):
        raise RuntimeError('This is a non-physical operation.')
