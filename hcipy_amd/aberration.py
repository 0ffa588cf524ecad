"""Surface aberrations built on the spectral-noise screens and the Fourier filter
(hcipy/optics/aberration.py:56-183).

``make_high_pass_power_law_error`` filters a power-law screen with a super-Gaussian high-pass
FourierFilter (q = 2) on the device; ``SurfaceAberration`` applies a power-law screen as a
reflective surface; ``SurfaceAberrationAtDistance`` wraps a surface between a Fresnel
propagation to the surface and back.
"""
import torch

from .aperture import make_circular_aperture, make_super_gaussian_aperture
from .field import Field
from .fourier_filter import FourierFilter
from .optics import OpticalElement, SurfaceApodizer
from .propagation import FresnelPropagator
from .spectral_noise import _host, make_power_law_error

__all__ = ['make_high_pass_power_law_error', 'SurfaceAberration', 'SurfaceAberrationAtDistance']


def make_high_pass_power_law_error(pupil_grid, ptv, diameter, cutoff_frequency, exponent=-2.5, aperture=None,
                                   filter_shape_parameter=15, seed=None):
    """A power-law surface error (peak-to-valley ``ptv`` over a circle of ``diameter`` before
    filtering) through the high-pass filter 1 - super-Gaussian(2 cutoff_frequency), real part,
    times the mask ``aperture > 0`` (aberration.py:56-92).  ``seed`` picks the realization."""
    sg = make_super_gaussian_aperture(2 * cutoff_frequency, filter_shape_parameter)

    def filter_function(fourier_grid):
        return 1 - _host(sg(fourier_grid))

    ff = FourierFilter(pupil_grid, filter_function, q=2)
    if aperture is None:
        aperture = make_circular_aperture(diameter)(pupil_grid)
    mask = (_host(aperture) > 0).astype(float)
    screen = make_power_law_error(pupil_grid, ptv, diameter, exponent=exponent, aperture=None, seed=seed)
    out = ff.forward(screen).as_subclass(torch.Tensor).real
    return Field(out * torch.as_tensor(mask, device=out.device), pupil_grid)


class SurfaceAberration(SurfaceApodizer):
    """A power-law surface error (make_power_law_error) applied as a surface of refractive index
    ``refractive_index`` (-1: a mirror) (aberration.py:94-121)."""

    def __init__(self, pupil_grid, ptv, diameter, exponent=-2.5, refractive_index=-1, aperture=None,
                 remove_modes=None, seed=None):
        surface = make_power_law_error(pupil_grid, ptv, diameter, exponent, aperture, remove_modes, seed=seed)
        SurfaceApodizer.__init__(self, surface, refractive_index)


class SurfaceAberrationAtDistance(OpticalElement):
    """A surface at ``distance`` from the current plane: Fresnel-propagate to it, apply the
    surface, propagate back (aberration.py:123-183).  The adjoint applies the surface's
    backward between the same two propagations, as the reference does."""

    def __init__(self, input_grid, surface_aberration, distance):
        self.fresnel = FresnelPropagator(input_grid, distance)
        self.surface_aberration = surface_aberration

    def forward(self, wavefront):
        wf = self.fresnel.forward(wavefront)
        wf = self.surface_aberration.forward(wf)
        return self.fresnel.backward(wf)

    def backward(self, wavefront):
        wf = self.fresnel.forward(wavefront)
        wf = self.surface_aberration.backward(wf)
        return self.fresnel.backward(wf)
