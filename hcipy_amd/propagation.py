"""Propagators (hcipy/propagation/).

``FraunhoferPropagator`` runs the MFMA matrix Fourier transform with the lens normalisation
1/(i f lambda) folded into the GEMM epilogue (hcipy/propagation/fraunhofer.py:32-88).
"""
import numpy as np

from . import _lib
from .field import Field
from .fourier import MatrixFourierTransform
from .optics import AgnosticOpticalElement, Wavefront

__all__ = ['FraunhoferPropagator']


class FraunhoferPropagator(AgnosticOpticalElement):
    """Monochromatic perfect-lens propagation from the front to the back focal plane.

    forward: ``MFT_{pupil -> uv}(E) / (i f lambda)`` with ``uv = focal * 2 pi / (f lambda)``;
    backward: ``MFT^dagger(E) * (i f lambda)`` (fraunhofer.py:32-38, 56-88).
    ``precision`` selects the complex64 arithmetic ('f16x2' or 'fp32').
    """

    def __init__(self, input_grid, output_grid, focal_length=1, precision=None):
        self._input_grid = input_grid
        self._output_grid = output_grid
        self._focal_length = focal_length
        self.precision = precision
        AgnosticOpticalElement.__init__(self, grid_dependent=True, wavelength_dependent=True)

    def make_instance(self, instance_data, input_grid, output_grid, wavelength):
        focal_length = self.evaluate_parameter(self.focal_length, input_grid, output_grid, wavelength)
        instance_data.uv_grid = output_grid.scaled(2 * np.pi / (focal_length * wavelength))
        instance_data.fourier_transform = MatrixFourierTransform(input_grid, instance_data.uv_grid, precision=self.precision)
        instance_data.norm_factor = 1 / (1j * focal_length * wavelength)

    @property
    def focal_length(self):
        return self._focal_length

    @focal_length.setter
    def focal_length(self, focal_length):
        self._focal_length = focal_length
        self.clear_cache()

    def get_input_grid(self, output_grid, wavelength):
        return self._input_grid

    def get_output_grid(self, input_grid, wavelength):
        return self._output_grid

    def instance(self, wavelength):
        return self.get_instance_data(self._input_grid, self._output_grid, wavelength)

    def forward(self, wavefront):
        d = self.instance(wavefront.wavelength)
        U = d.fourier_transform._apply(wavefront.electric_field, _lib.FORWARD, extra_alpha=d.norm_factor)
        return Wavefront(Field(U, self._output_grid), wavefront.wavelength)

    def backward(self, wavefront):
        d = self.instance(wavefront.wavelength)
        U = d.fourier_transform._apply(wavefront.electric_field, _lib.BACKWARD, extra_alpha=1 / d.norm_factor)
        return Wavefront(Field(U, self._input_grid), wavefront.wavelength)
