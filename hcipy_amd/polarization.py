"""Jones-matrix optical elements (hcipy/optics/polarization.py:22-209).

The Jones matrices are set-up data evaluated on the host in FP64 (like the aperture
generators); applying them to a wavefront is the per-pixel product ``hcipy_pw_field_dot``.
"""
import numpy as np
import torch

from . import pointwise
from .field import Field, default_device
from .optics import AgnosticOpticalElement, Wavefront

__all__ = ['JonesMatrixOpticalElement', 'PhaseRetarder', 'LinearRetarder', 'CircularRetarder', 'phase_retarder_jones']


def phase_retarder_jones(phase_retardation, fast_axis_orientation, circularity):
    """Jones matrix of a general phase retarder (polarization.py:141-160), shape (2, 2) for scalar
    parameters or (2, 2, n) when any parameter is an array of n samples."""
    d = np.asarray(phase_retardation, dtype=float)
    t = np.asarray(fast_axis_orientation, dtype=float)
    c = np.asarray(circularity, dtype=float)
    phi_plus = np.exp(1j * d / 2)
    phi_minus = np.exp(-1j * d / 2)
    ct, st = np.cos(t), np.sin(t)
    j11 = phi_plus * ct ** 2 + phi_minus * st ** 2
    j12 = (phi_plus - phi_minus) * np.exp(-1j * c) * ct * st
    j21 = (phi_plus - phi_minus) * np.exp(1j * c) * ct * st
    j22 = phi_plus * st ** 2 + phi_minus * ct ** 2
    j11, j12, j21, j22 = np.broadcast_arrays(j11, j12, j21, j22)
    return np.array([[j11, j12], [j21, j22]])


def _param(p):
    """Host array of a parameter that may be a scalar, an ndarray or a (device) Field."""
    if isinstance(p, torch.Tensor):
        return p.as_subclass(torch.Tensor).detach().cpu().numpy()
    return p


class JonesMatrixOpticalElement(AgnosticOpticalElement):
    """A Jones matrix, constant (2, 2) or a (2, 2, n) field, or a function of the wavelength
    returning either (polarization.py:22-116).  A scalar wavefront becomes the unpolarized
    wavefront J E (input Stokes vector (1, 0, 0, 0)); otherwise E <- J E (J^H E backward)."""

    def __init__(self, jones_matrix):
        self.jones_matrix = jones_matrix
        AgnosticOpticalElement.__init__(self, True, True)

    def make_instance(self, instance_data, input_grid, output_grid, wavelength):
        J = _param(self.evaluate_parameter(self.jones_matrix, input_grid, output_grid, wavelength))
        J = np.asarray(J, dtype=complex)
        if J.ndim == 2:
            J = np.repeat(J[:, :, np.newaxis], input_grid.size, axis=2)
        instance_data.jones_matrix = torch.as_tensor(np.ascontiguousarray(J), device=default_device())

    def _apply(self, wavefront, adjoint):
        d = self.get_instance_data(wavefront.grid, None, wavefront.wavelength)
        wf = wavefront
        if wf.is_scalar:
            wf = Wavefront(wf.electric_field, wf.wavelength, (1, 0, 0, 0))
        E = pointwise.field_dot(d.jones_matrix, wf.electric_field, adjoint=adjoint)
        return Wavefront(Field(E, wf.grid), wf.wavelength, wf.input_stokes_vector)

    def forward(self, wavefront):
        return self._apply(wavefront, False)

    def backward(self, wavefront):
        return self._apply(wavefront, True)


class PhaseRetarder(JonesMatrixOpticalElement):
    """General phase retarder (polarization.py:121-194): retardation, fast-axis orientation and
    circularity, each a scalar or an array over the grid."""

    def __init__(self, phase_retardation, fast_axis_orientation, circularity):
        self.phase_retardation = phase_retardation
        self.fast_axis_orientation = fast_axis_orientation
        self.circularity = circularity
        JonesMatrixOpticalElement.__init__(self, self._jones)

    def _jones(self, wavelength):
        ev = lambda p: _param(p(wavelength) if callable(p) else p)  # noqa: E731
        return phase_retarder_jones(ev(self.phase_retardation), ev(self.fast_axis_orientation), ev(self.circularity))


class LinearRetarder(PhaseRetarder):
    """Linear retarder (polarization.py:196-207): circularity 0."""

    def __init__(self, phase_retardation, fast_axis_orientation):
        PhaseRetarder.__init__(self, phase_retardation, fast_axis_orientation, 0)


class CircularRetarder(PhaseRetarder):
    """Circular retarder (polarization.py:209-220): fast axis at 45 degrees, circularity pi/2."""

    def __init__(self, phase_retardation):
        PhaseRetarder.__init__(self, phase_retardation, np.pi / 4, np.pi / 2)
