"""Lyot coronagraph in Babinet form (hcipy/coronagraphy/lyot.py:7-122)."""
from . import pointwise
from .field import Field
from .optics import Apodizer, OpticalElement, Wavefront
from .propagation import FraunhoferPropagator

__all__ = ['LyotCoronagraph']


class LyotCoronagraph(OpticalElement):
    """Lyot coronagraph with a small focal-plane mask M on its own focal grid.

    forward (lyot.py:66-92):  F = P E;  F <- F - M F;  L = E - P^dagger F;  L <- L * stop.
    The two subtractions run as single fused passes (``hcipy_pw_sub_mul``): F (1 - M) on the
    focal grid and (E - P^dagger F) * stop on the pupil grid.
    """

    def __init__(self, input_grid, focal_plane_mask, lyot_stop=None, focal_length=1, focal_plane_mask_grid=None,
                 precision=None):
        if isinstance(focal_plane_mask, OpticalElement):
            if focal_plane_mask_grid is None:
                raise ValueError('Could not determine the input grid of the focal plane mask. Please pass focal_plane_mask_grid explicitly.')
            grid = focal_plane_mask_grid
            self.focal_plane_mask = focal_plane_mask
        else:
            grid = focal_plane_mask.grid if focal_plane_mask_grid is None else focal_plane_mask_grid
            self.focal_plane_mask = Apodizer(focal_plane_mask)
        if lyot_stop is not None and not isinstance(lyot_stop, OpticalElement):
            lyot_stop = Apodizer(lyot_stop)
        self.lyot_stop = lyot_stop
        self.prop = FraunhoferPropagator(input_grid, grid, focal_length=focal_length, precision=precision)

    def _fpm_complement(self, wf_foc, conjugate):
        """F - M F (or F - conj(M) F)."""
        if isinstance(self.focal_plane_mask, Apodizer) and not callable(self.focal_plane_mask.apodization):
            a = self.focal_plane_mask.apodization
            masked = pointwise.multiply(wf_foc.electric_field, a, conjugate=conjugate)
        else:
            el = self.focal_plane_mask
            masked = (el.backward(wf_foc) if conjugate else el.forward(wf_foc)).electric_field
        return pointwise.sub_mul(wf_foc.electric_field, masked)

    def _stop(self):
        if isinstance(self.lyot_stop, Apodizer) and not callable(self.lyot_stop.apodization):
            return self.lyot_stop.apodization
        return None

    def forward(self, wavefront):
        wf_foc = self.prop.forward(wavefront)
        wf_foc = Wavefront(self._fpm_complement(wf_foc, False), wavefront.wavelength)
        lyot = self.prop.backward(wf_foc)
        stop = self._stop()
        if self.lyot_stop is None or stop is not None:
            E = pointwise.sub_mul(wavefront.electric_field, lyot.electric_field, stop)
            return Wavefront(Field(E, wavefront.grid), wavefront.wavelength)
        lyot = Wavefront(pointwise.sub_mul(wavefront.electric_field, lyot.electric_field), wavefront.wavelength)
        return self.lyot_stop.forward(lyot)

    def backward(self, wavefront):
        wf = self.lyot_stop.backward(wavefront) if self.lyot_stop is not None else wavefront
        wf_foc = self.prop.forward(wf)
        wf_foc = Wavefront(self._fpm_complement(wf_foc, True), wavefront.wavelength)
        pup = self.prop.backward(wf_foc)
        return Wavefront(pointwise.sub_mul(wf.electric_field, pup.electric_field), wavefront.wavelength)
