"""Propagators (hcipy/propagation/).

``FraunhoferPropagator`` runs the MFMA matrix Fourier transform with the lens normalisation
1/(i f lambda) folded into the GEMM epilogue (hcipy/propagation/fraunhofer.py:32-88).
"""
import numpy as np

import torch

from . import _lib
from .field import Field, default_device
from .fourier import (FastFourierTransform, MatrixFourierTransform, MftGroup, grouped_transforms,
                      grouped_transforms_stacked, make_fft_grid)
from .optics import AgnosticOpticalElement, Wavefront

__all__ = ['FraunhoferPropagator']


class FraunhoferPropagator(AgnosticOpticalElement):
    """Monochromatic perfect-lens propagation from the front to the back focal plane.

    forward: ``MFT_{pupil -> uv}(E) / (i f lambda)`` with ``uv = focal * 2 pi / (f lambda)``;
    backward: ``MFT^dagger(E) * (i f lambda)`` (fraunhofer.py:32-38, 56-88).
    ``precision`` selects the complex64 arithmetic ('f16x2', 'bf16x3' or 'fp32').
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
        return Wavefront(Field(U, self._output_grid), wavefront.wavelength, wavefront.input_stokes_vector)

    def backward(self, wavefront):
        d = self.instance(wavefront.wavelength)
        U = d.fourier_transform._apply(wavefront.electric_field, _lib.BACKWARD, extra_alpha=1 / d.norm_factor)
        return Wavefront(Field(U, self._input_grid), wavefront.wavelength, wavefront.input_stokes_vector)

    def _group_for(self, n, group):
        if group is not None:
            return group
        g = getattr(self, '_group', None)
        if g is None or g.capacity < n:
            g = self._group = MftGroup(max(64, n))
        return g

    def forward_stacked(self, x, wavelengths, group=None):
        """Row i of the stacked field tensor ``x`` (``(n, ..., input size)``) propagated forward at
        ``wavelengths[i]``, as one grouped launch per MFT stage; returns the stacked output tensor.
        Differentiable (the adjoints are grouped too).  The input grid must have scalar weights."""
        ds = [self.instance(wl) for wl in wavelengths]
        return grouped_transforms_stacked([d.fourier_transform for d in ds], _lib.FORWARD, x,
                                          [d.norm_factor for d in ds], self._group_for(len(ds), group))

    def backward_stacked(self, x, wavelengths, group=None):
        """``backward`` of row i of the stacked ``x`` at ``wavelengths[i]``; as ``forward_stacked``."""
        ds = [self.instance(wl) for wl in wavelengths]
        return grouped_transforms_stacked([d.fourier_transform for d in ds], _lib.BACKWARD, x,
                                          [1 / d.norm_factor for d in ds], self._group_for(len(ds), group))

    @staticmethod
    def _stackable(fields):
        ts = [f.as_subclass(torch.Tensor) if isinstance(f, torch.Tensor) else None for f in fields]
        return (all(t is not None for t in ts) and len({(tuple(t.shape), t.dtype, t.device) for t in ts}) == 1,
                ts)

    def forward_multiple(self, wavefronts, output='field', group=None):
        """Forward propagation of several wavefronts (for example the wavelengths of a broadband
        source, each with its own Fourier matrices) as one grouped launch per MFT stage.
        ``output`` is 'field' (complex, like ``forward``) or 'intensity' (fused |E|^2, like
        ``forward_intensity``).  Returns a list of Wavefronts or intensity Fields.  With gradients
        recorded the result is differentiable like ``forward``, and the backward pass of the
        transforms is one grouped launch per stage as well."""
        wavefronts = list(wavefronts)
        if output not in ('field', 'intensity'):
            raise ValueError("output must be 'field' or 'intensity'")
        if np.size(self._input_grid.weights) != 1:  # non-scalar weights: one call per wavefront
            if output == 'field':
                return [self.forward(wf) for wf in wavefronts]
            return [self.forward_intensity(wf) for wf in wavefronts]
        group = self._group_for(len(wavefronts), group)
        grad = torch.is_grad_enabled() and any(isinstance(wf.electric_field, torch.Tensor) and
                                               wf.electric_field.requires_grad for wf in wavefronts)
        if output == 'field' or grad:
            ok, ts = self._stackable([wf.electric_field for wf in wavefronts])
            if ok and ts[0].is_complex():
                # one stacked tensor: the intensities below run as one launch
                st = self.forward_stacked(torch.stack(ts), [wf.wavelength for wf in wavefronts], group)
                if output == 'intensity':
                    from .pointwise import intensity
                    return [Field(i, self._output_grid) for i in intensity(st).unbind(0)]
                return [Wavefront(Field(o, self._output_grid), wf.wavelength, wf.input_stokes_vector)
                        for o, wf in zip(st.unbind(0), wavefronts)]
            calls = []
            for wf in wavefronts:
                d = self.instance(wf.wavelength)
                calls.append((d.fourier_transform, _lib.FORWARD, wf.electric_field, d.norm_factor))
            outs = grouped_transforms(calls, group)
            if output == 'field':
                return [Wavefront(Field(o, self._output_grid), wf.wavelength, wf.input_stokes_vector)
                        for o, wf in zip(outs, wavefronts)]
            from .pointwise import intensity
            return [Field(intensity(o), self._output_grid) for o in outs]
        kind = _lib.OUTPUT_INTENSITY
        calls = []
        for wf in wavefronts:
            d = self.instance(wf.wavelength)
            w = complex(np.asarray(self._input_grid.weights).ravel()[0])
            calls.append((d.fourier_transform, _lib.FORWARD, wf.electric_field, w * d.norm_factor, kind))
        return [Field(o, self._output_grid) for o in group.execute(calls)]

    def backward_multiple(self, wavefronts, group=None):
        """``backward`` of several wavefronts as one grouped launch per MFT stage; differentiable
        like ``backward``.  Returns a list of Wavefronts."""
        wavefronts = list(wavefronts)
        if np.size(self._output_grid.weights) != 1:  # non-scalar weights: one call per wavefront
            return [self.backward(wf) for wf in wavefronts]
        ok, ts = self._stackable([wf.electric_field for wf in wavefronts])
        if ok and ts[0].is_complex():
            outs = self.backward_stacked(torch.stack(ts), [wf.wavelength for wf in wavefronts], group).unbind(0)
        else:
            calls = []
            for wf in wavefronts:
                d = self.instance(wf.wavelength)
                calls.append((d.fourier_transform, _lib.BACKWARD, wf.electric_field, 1 / d.norm_factor))
            outs = grouped_transforms(calls, self._group_for(len(wavefronts), group))
        return [Wavefront(Field(o, self._input_grid), wf.wavelength, wf.input_stokes_vector)
                for o, wf in zip(outs, wavefronts)]

    def forward_intensity(self, wavefront, float64=False):
        """``forward(wavefront).intensity`` in one pass: |E|^2 is evaluated in the MFT epilogue
        from the FP64 value, so complex64 inputs are not limited by the complex64 rounding of the
        focal field."""
        d = self.instance(wavefront.wavelength)
        out = _lib.OUTPUT_INTENSITY_F64 if float64 else _lib.OUTPUT_INTENSITY
        I = d.fourier_transform._apply(wavefront.electric_field, _lib.FORWARD, extra_alpha=d.norm_factor, output=out)
        return Field(I, self._output_grid)


def _nearfield_branch(input_grid, wavelength, distance):
    """True when the native transfer function applies (angular_spectrum.py:53-55, fresnel.py:48-50)."""
    L_max = max(n * d for n, d in zip(input_grid.dims, input_grid.delta))
    return not any(d < wavelength * abs(distance) / L_max for d in input_grid.delta)


class _NearFieldPropagator(AgnosticOpticalElement):
    KIND = None

    def __init__(self, input_grid, distance, num_oversampling=2, refractive_index=1, zero_padding=2):
        self._distance = distance
        self._num_oversampling = num_oversampling
        self._refractive_index = refractive_index
        self._zero_padding = zero_padding
        AgnosticOpticalElement.__init__(self, grid_dependent=True, wavelength_dependent=True)

    def make_instance(self, instance_data, input_grid, output_grid, wavelength):
        from .fourier_filter import FourierFilter, NearFieldTransferFunction
        if not input_grid.is_regular or not input_grid.is_('cartesian'):
            raise ValueError('The input grid must be a regular, Cartesian grid.')
        n = self.evaluate_parameter(self.refractive_index, input_grid, output_grid, wavelength)
        k = 2 * np.pi / wavelength * n
        if _nearfield_branch(input_grid, wavelength, self.distance):
            tf = NearFieldTransferFunction(self.KIND, k, self.distance, self.num_oversampling)
        else:
            tf = self._impulse_response_tf(input_grid, k, wavelength)
        instance_data.fourier_filter = FourierFilter(input_grid, tf, q=self._q())

    def _impulse_response_tf(self, input_grid, k, wavelength):
        """Impulse-response branch (angular_spectrum.py:55-69, fresnel.py:50-67), taken when the grid
        step is below lambda |z| / L: the impulse response, supersampled on the spatial grid of
        the filter's internal grid (the padded input grid), Fourier transformed onto the internal
        grid.  The impulse response is generated in FP64 on the device; the transform is the
        FP64 MFMA transform onto the FFT grid."""
        internal = make_fft_grid(input_grid, self._q(), 1)
        enlarged = make_fft_grid(internal)
        dev = default_device()
        with torch.cuda.device(dev):
            h = torch.empty(enlarged.size, dtype=torch.complex128, device=dev)
            (x0, y0), (dx, dy) = enlarged.zero, enlarged.delta
            nx, ny = (int(v) for v in enlarged.dims)
            kind = {'angular_spectrum': 0, 'fresnel': 1}[self.KIND]
            st = _lib.load().hcipy_nearfield_impulse_response(kind, nx, ny, float(dx), float(x0), float(dy), float(y0),
                                                             float(k), float(wavelength), float(self.distance),
                                                             int(self.num_oversampling), _lib.dptr(h),
                                                             _lib.stream_ptr(dev))
            _lib.check(st, 'hcipy_nearfield_impulse_response')
            return FastFourierTransform(enlarged).forward(Field(h, enlarged))

    def _q(self):
        return self._zero_padding

    @property
    def distance(self):
        return self._distance

    @distance.setter
    def distance(self, distance):
        self._distance = distance
        self.clear_cache()

    @property
    def num_oversampling(self):
        return self._num_oversampling

    @property
    def refractive_index(self):
        return self._refractive_index

    def get_input_grid(self, output_grid, wavelength):
        return output_grid

    def get_output_grid(self, input_grid, wavelength):
        return input_grid

    def forward(self, wavefront):
        d = self.get_instance_data(wavefront.grid, None, wavefront.wavelength)
        return Wavefront(d.fourier_filter.forward(wavefront.electric_field), wavefront.wavelength,
                         wavefront.input_stokes_vector)

    def backward(self, wavefront):
        d = self.get_instance_data(wavefront.grid, None, wavefront.wavelength)
        return Wavefront(d.fourier_filter.backward(wavefront.electric_field), wavefront.wavelength,
                         wavefront.input_stokes_vector)


class AngularSpectrumPropagator(_NearFieldPropagator):
    """Angular-spectrum propagation, transfer function exp(i sqrt(k^2 - |k|^2 + 0j) z)
    supersampled x num_oversampling, FourierFilter with q = 2 (angular_spectrum.py:7-152)."""
    KIND = 'angular_spectrum'

    def __init__(self, input_grid, distance, num_oversampling=2, refractive_index=1):
        _NearFieldPropagator.__init__(self, input_grid, distance, num_oversampling, refractive_index, 2)


class FresnelPropagator(_NearFieldPropagator):
    """Paraxial Fresnel propagation, transfer function exp(ikz) exp(-i z |k|^2 / 2k)
    (fresnel.py:8-164), FourierFilter with q = zero_padding."""
    KIND = 'fresnel'

    def __init__(self, input_grid, distance, num_oversampling=2, refractive_index=1, zero_padding=2):
        _NearFieldPropagator.__init__(self, input_grid, distance, num_oversampling, refractive_index, zero_padding)


__all__ += ['AngularSpectrumPropagator', 'FresnelPropagator']
