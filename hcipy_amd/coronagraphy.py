"""Coronagraphs: the Lyot coronagraph in Babinet form (hcipy/coronagraphy/lyot.py:7-122) and the
multi-scale focal-mask coronagraph with its vortex specialisation
(hcipy/coronagraphy/multi_scale.py:10-182, hcipy/coronagraphy/vortex.py:12-15)."""
import math

import numpy as np
import torch

from . import _lib, pointwise
from .aperture import make_circular_aperture
from .field import Field, make_focal_grid
from .fourier import FastFourierTransform, MatrixFourierTransform, MftGroup
from .fourier_filter import FourierFilter
from .optics import AgnosticOpticalElement, Apodizer, OpticalElement, Wavefront
from .polarization import phase_retarder_jones
from .propagation import FraunhoferPropagator

__all__ = ['LyotCoronagraph', 'MultiScaleCoronagraph', 'VortexCoronagraph', 'VectorVortexCoronagraph',
           'periodic_tukey_window']


def _plain_tensor(t):
    return t.as_subclass(torch.Tensor) if isinstance(t, Field) else t


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
        wf_foc = Wavefront(self._fpm_complement(wf_foc, False), wavefront.wavelength,
                           wavefront.input_stokes_vector)
        lyot = self.prop.backward(wf_foc)
        stop = self._stop()
        if self.lyot_stop is None or stop is not None:
            E = pointwise.sub_mul(wavefront.electric_field, lyot.electric_field, stop)
            return Wavefront(Field(E, wavefront.grid), wavefront.wavelength, wavefront.input_stokes_vector)
        lyot = Wavefront(pointwise.sub_mul(wavefront.electric_field, lyot.electric_field), wavefront.wavelength,
                         wavefront.input_stokes_vector)
        return self.lyot_stop.forward(lyot)

    def forward_multiple(self, wavefronts, group=None):
        """``forward`` of several wavefronts (for example the wavelengths of a broadband source),
        with the transforms to and from the focal-plane mask of all of them as grouped launches
        (``FraunhoferPropagator.forward_multiple`` / ``backward_multiple``).  Differentiable like
        ``forward``; the backward pass groups the adjoint transforms too.  Returns a list of
        Wavefronts."""
        wavefronts = list(wavefronts)
        stop = self._stop()
        ok, ts = self.prop._stackable([wf.electric_field for wf in wavefronts])
        if ok and ts[0].is_complex() and self._stackable_chain():
            wls = [wf.wavelength for wf in wavefronts]
            out = self.forward_stacked(torch.stack(ts), wls, group)
            return [Wavefront(Field(o, wf.grid), wf.wavelength, wf.input_stokes_vector)
                    for o, wf in zip(out.unbind(0), wavefronts)]
        foc = self.prop.forward_multiple(wavefronts, group=group)
        foc = [Wavefront(self._fpm_complement(f, False), wf.wavelength, wf.input_stokes_vector)
               for f, wf in zip(foc, wavefronts)]
        lyot = self.prop.backward_multiple(foc, group=group)
        out = []
        for wf, ly in zip(wavefronts, lyot):
            if self.lyot_stop is None or stop is not None:
                E = pointwise.sub_mul(wf.electric_field, ly.electric_field, stop)
                out.append(Wavefront(Field(E, wf.grid), wf.wavelength, wf.input_stokes_vector))
            else:
                L = Wavefront(pointwise.sub_mul(wf.electric_field, ly.electric_field), wf.wavelength,
                              wf.input_stokes_vector)
                out.append(self.lyot_stop.forward(L))
        return out

    def _stackable_chain(self):
        """True when the chain can run on stacked wavelengths: a fixed focal-plane mask and Lyot
        stop (no wavelength dependence) and grids with scalar weights."""
        fixed_mask = isinstance(self.focal_plane_mask, Apodizer) and not callable(self.focal_plane_mask.apodization)
        return (fixed_mask and np.size(self.prop._input_grid.weights) == 1 and
                np.size(self.prop._output_grid.weights) == 1 and (self.lyot_stop is None or self._stop() is not None))

    def forward_stacked(self, x, wavelengths, group=None):
        """``forward`` of row i of the stacked pupil field ``x`` (``(n, ..., pupil size)``) at
        ``wavelengths[i]``: the two transforms of all rows as grouped launches
        (``FraunhoferPropagator.forward_stacked`` / ``backward_stacked``), the mask and the stop
        as one launch each.  Returns the stacked Lyot-plane field; differentiable.  Needs a fixed
        focal-plane mask (an Apodizer with an array) and Lyot stop."""
        if not self._stackable_chain():
            raise ValueError('forward_stacked needs a fixed focal-plane mask and Lyot stop and scalar grid weights')
        E = _plain_tensor(x).contiguous()
        F = self.prop.forward_stacked(E, wavelengths, group)
        F = self._fpm_complement(Wavefront(Field(F, self.prop._output_grid), wavelengths[0]), False)
        L = self.prop.backward_stacked(_plain_tensor(F), wavelengths, group)
        return _plain_tensor(pointwise.sub_mul(E, L, self._stop()))

    def backward(self, wavefront):
        wf = self.lyot_stop.backward(wavefront) if self.lyot_stop is not None else wavefront
        wf_foc = self.prop.forward(wf)
        wf_foc = Wavefront(self._fpm_complement(wf_foc, True), wavefront.wavelength,
                           wavefront.input_stokes_vector)
        pup = self.prop.backward(wf_foc)
        return Wavefront(pointwise.sub_mul(wf.electric_field, pup.electric_field), wavefront.wavelength,
                         wavefront.input_stokes_vector)


def periodic_tukey_window(n):
    """Periodic Tukey window with taper fraction 1 (a periodic Hann window), i.e.
    ``scipy.signal.windows.tukey(n, 1, False)``: ``0.5 - 0.5 cos(2 pi k / n)``, k = 0 .. n-1."""
    k = np.arange(n, dtype=float)
    return 0.5 - 0.5 * np.cos(2 * np.pi * k / n)


class MultiScaleCoronagraph(OpticalElement):
    """Focal-plane phase-mask coronagraph evaluated on nested focal grids
    (hcipy/coronagraphy/multi_scale.py:10-182).

    Level 0 samples the mask at q = 2 over the whole FFT field and runs as a FourierFilter with
    the (windowed) level-0 mask as transfer function.  Level i samples the central
    ``window_size`` pixels of level i - 1 at ``scaling_factor`` times the resolution; its mask is
    windowed (all but the last level) and has the band-limited content of every coarser level
    (FFT, then adjoint MFT onto its grid) subtracted, so that the levels add up to one mask
    sampled finely near the centre.  The propagation runs at wavelength 1 (focal grids in
    lambda/D):

        L = FF_0(E) + sum_i P_i^dagger (M_i . P_i E),   then the Lyot stop,

    with P_i the Fraunhofer propagator onto level i (conj(M_i) and the adjoint filter for
    ``backward``).  Without gradients the level MFTs run as two grouped launches (all forward
    transforms, then all adjoint transforms); the adjoint transforms are computed with a
    negated scale so that each level is added by a fused subtract pass.

    ``complex_mask(grid)`` returns a scalar mask (size,) or a Jones-matrix mask (2, 2, size); a
    Jones mask acts on the first dimension of a polarized field (field_dot), level 0 then
    being a matrix-transfer-function FourierFilter.
    """

    def __init__(self, input_grid, complex_mask, lyot_stop=None, q=1024, scaling_factor=4, window_size=32,
                 precision=None):
        self.input_grid = input_grid
        pupil_diameter = np.asarray(input_grid.dims) * np.asarray(input_grid.delta)
        if lyot_stop is not None and not isinstance(lyot_stop, OpticalElement):
            lyot_stop = Apodizer(lyot_stop)
        self.lyot_stop = lyot_stop

        levels = int(math.ceil(math.log(q / 2) / math.log(scaling_factor))) + 1
        qs = [2 * scaling_factor ** i for i in range(levels)]
        num_airys = [tuple(n / 2 for n in input_grid.dims)]
        for i in range(1, levels):
            num_airys.append(tuple(n_airy * window_size / (2 * qs[i - 1] * n_airy) for n_airy in num_airys[i - 1]))

        self.focal_grids = []
        self.focal_masks = []
        self.props = []
        w1 = periodic_tukey_window(window_size)
        w = np.outer(w1, w1)
        spectra = [None] * levels
        for i in range(levels):
            focal_grid = make_focal_grid(qs[i], num_airys[i], pupil_diameter=pupil_diameter, reference_wavelength=1,
                                         focal_length=1)
            mask = complex_mask(focal_grid)
            mask = (mask.as_subclass(torch.Tensor) if isinstance(mask, torch.Tensor)
                    else torch.as_tensor(np.asarray(mask))).to(torch.complex128)
            dev = mask.device if mask.is_cuda else torch.device('cuda', torch.cuda.current_device())
            mask = mask.to(dev)
            if i != levels - 1:
                # np.pad with one (before, after) pair built from the two axes, as the reference does
                wp = np.pad(w, tuple((n_foc - n_w) // 2 for n_foc, n_w in zip(focal_grid.shape, w.shape)), 'constant')
                mask = mask * torch.as_tensor(1 - wp.ravel(), device=dev)
            for j in range(i):
                # the FFT of level j's mask does not depend on i: computed once, on first use
                if spectra[j] is None:
                    fft = FastFourierTransform(self.focal_grids[j])
                    spectra[j] = (fft.output_grid, fft.forward(Field(self.focal_masks[j], self.focal_grids[j])))
                fft_grid, spectrum = spectra[j]
                mft = MatrixFourierTransform(focal_grid, fft_grid)
                mask = mask - mft.backward(spectrum).as_subclass(torch.Tensor)
            if i == 0:
                prop = FourierFilter(input_grid, Field(mask, focal_grid), qs[0])
            else:
                prop = FraunhoferPropagator(input_grid, focal_grid, precision=precision)
            self.focal_grids.append(focal_grid)
            self.focal_masks.append(Field(mask, focal_grid))
            self.props.append(prop)
        self._group = None
        # grouped: the level transforms run as two grouped launches; False runs them one by one
        self.grouped = True
        self._masks_h = None

    def _levels(self, E, conjugate):
        """sum over levels >= 1 of P_i^dagger (M_i P_i E), accumulated onto the level-0 result."""
        if (self.focal_masks[0].dim() == 3 and E.dim() == 1 and self.grouped and not pointwise.needs_grad(E)):
            return self._levels_scalar_jones(E, conjugate)
        L = self.props[0].backward(E) if conjugate else self.props[0].forward(E)
        rest = list(zip(self.focal_masks[1:], self.props[1:]))
        if not rest:
            return L
        if pointwise.needs_grad(E) or not self.grouped:
            for mask, prop in rest:
                d = prop.instance(1)
                F = d.fourier_transform._apply(E, _lib.FORWARD, extra_alpha=d.norm_factor)
                F = self._apply_mask(F, mask, conjugate)
                neg = d.fourier_transform._apply(F, _lib.BACKWARD, extra_alpha=-1 / d.norm_factor)
                L = pointwise.sub_mul(L, neg)
            return L
        # Two groups, one per grouped launch: a group's argument table is copied from its host
        # mirror when the launch runs, so under stream capture one group must not serve two
        # launches of the same captured sequence.
        if self._group is None or self._group[0].device != E.device:
            self._group = (MftGroup(max(16, len(rest)), E.device), MftGroup(max(16, len(rest)), E.device))
        w_in = complex(np.asarray(self.input_grid.weights).ravel()[0])
        fwd = []
        for mask, prop in rest:
            d = prop.instance(1)
            fwd.append((d.fourier_transform, _lib.FORWARD, E, w_in * d.norm_factor, _lib.OUTPUT_FIELD))
        Fs = self._group[0].execute(fwd)
        bwd = []
        for F, (mask, prop) in zip(Fs, rest):
            d = prop.instance(1)
            F = self._apply_mask(Field(F, mask.grid), mask, conjugate)
            w_uv = complex(np.asarray(d.uv_grid.weights).ravel()[0]) / (2 * np.pi) ** 2
            bwd.append((d.fourier_transform, _lib.BACKWARD, F, -w_uv / d.norm_factor, _lib.OUTPUT_FIELD))
        return pointwise.sub_many(L, self._group[1].execute(bwd))

    def _levels_scalar_jones(self, E, conjugate):
        """Jones masks acting on a scalar field E, i.e. on E times the 2 x 2 identity: every
        level is linear, so level 0 is each element filter applied to E (T[i][k] = F_ik(E),
        adjoint F_ki^H(E)), and level i forms P_i E once and the masked field J P_i E (J^H
        for the adjoint) per component.  The same sums as the tensor field E (x) I, without the
        transforms of its zero components."""
        f0 = self.props[0]
        if conjugate:
            L = f0.apply_elements(E, adjoint=True).as_subclass(torch.Tensor).transpose(0, 1).contiguous()
        else:
            L = f0.apply_elements(E).as_subclass(torch.Tensor)
        rest = list(zip(self.focal_masks[1:], self.props[1:]))
        if not rest:
            return L
        if self._group is None or self._group[0].device != E.device:
            self._group = (MftGroup(max(16, len(rest)), E.device), MftGroup(max(16, len(rest)), E.device))
        if conjugate and self._masks_h is None:
            self._masks_h = [torch.conj(m.transpose(0, 1)).contiguous() for m in self.focal_masks]
        w_in = complex(np.asarray(self.input_grid.weights).ravel()[0])
        fwd = []
        for mask, prop in rest:
            d = prop.instance(1)
            fwd.append((d.fourier_transform, _lib.FORWARD, E, w_in * d.norm_factor, _lib.OUTPUT_FIELD))
        Fs = self._group[0].execute(fwd)
        bwd = []
        for k, (F, (mask, prop)) in enumerate(zip(Fs, rest)):
            d = prop.instance(1)
            J = self._masks_h[k + 1] if conjugate else mask
            # (J F)[i][k] = J_ik F: the masked E (x) I, in the field's dtype
            F4 = F.as_subclass(torch.Tensor).expand((2, 2) + tuple(F.shape)).contiguous()
            G = pointwise.multiply(F4, J)
            w_uv = complex(np.asarray(d.uv_grid.weights).ravel()[0]) / (2 * np.pi) ** 2
            bwd.append((d.fourier_transform, _lib.BACKWARD, G, -w_uv / d.norm_factor, _lib.OUTPUT_FIELD))
        return pointwise.sub_many(L, self._group[1].execute(bwd))

    @staticmethod
    def _apply_mask(F, mask, conjugate):
        """M F (conj(M) F), or J F (J^H F) for a Jones-matrix mask."""
        if mask.dim() == 3:
            return pointwise.field_dot(mask, F, adjoint=conjugate)
        return pointwise.multiply(F, mask, conjugate=conjugate)

    def forward(self, wavefront):
        """Pupil -> Lyot plane (multi_scale.py:108-145)."""
        E = wavefront.electric_field
        L = Field(self._levels(E, False), self.input_grid)
        lyot = Wavefront(L, wavefront.wavelength, wavefront.input_stokes_vector)
        if self.lyot_stop is not None:
            lyot = self.lyot_stop.forward(lyot)
        return lyot

    def backward(self, wavefront):
        """Lyot plane -> pupil: the stop's adjoint, then the levels with conj(M_i)
        (multi_scale.py:147-182)."""
        if self.lyot_stop is not None:
            wavefront = self.lyot_stop.backward(wavefront)
        E = wavefront.electric_field
        return Wavefront(Field(self._levels(E, True), self.input_grid), wavefront.wavelength,
                         wavefront.input_stokes_vector)


class VortexCoronagraph(MultiScaleCoronagraph):
    """Scalar vortex coronagraph of topological charge ``charge``: the multi-scale coronagraph
    with the mask exp(i charge theta), its sample at the origin removed
    (hcipy/coronagraphy/vortex.py:12-15)."""

    def __init__(self, input_grid, charge, lyot_stop=None, q=1024, scaling_factor=4, window_size=32, precision=None):
        self.charge = charge
        hole = make_circular_aperture(1e-9)

        def complex_mask(grid):
            theta = grid.as_('polar').theta.as_subclass(torch.Tensor)
            return torch.exp(1j * charge * theta) * (1 - hole(grid).as_subclass(torch.Tensor))

        MultiScaleCoronagraph.__init__(self, input_grid, complex_mask, lyot_stop, q, scaling_factor, window_size,
                                       precision)


class VectorVortexCoronagraph(AgnosticOpticalElement):
    """Vector vortex coronagraph (hcipy/coronagraphy/vortex.py:17-199): on every level of the
    multi-scale scheme the focal mask is the Jones matrix of a linear retarder with retardance
    ``phase_retardation`` (a scalar or a function of the wavelength) and fast axis
    charge / 2 * theta, with the sample at the origin removed.  A scalar wavefront enters as the
    unpolarized wavefront (input Stokes vector (1, 0, 0, 0)); the instances are built per input
    grid and wavelength."""

    def __init__(self, charge, lyot_stop=None, phase_retardation=np.pi, q=1024, scaling_factor=4, window_size=32,
                 precision=None):
        self.charge = charge
        if lyot_stop is not None and not isinstance(lyot_stop, OpticalElement):
            lyot_stop = Apodizer(lyot_stop)
        self.lyot_stop = lyot_stop
        self.phase_retardation = phase_retardation
        self.q = q
        self.scaling_factor = scaling_factor
        self.window_size = window_size
        self.precision = precision
        AgnosticOpticalElement.__init__(self, grid_dependent=True, wavelength_dependent=True)

    def make_instance(self, instance_data, input_grid, output_grid, wavelength):
        retardance = float(self.evaluate_parameter(self.phase_retardation, input_grid, output_grid, wavelength))
        charge = self.charge

        def jones_mask(grid):
            x, y = (np.asarray(c) for c in grid.coords)
            theta = np.arctan2(y, x)
            hole = ((x ** 2 + y ** 2) <= (0.5e-9) ** 2).astype(float)
            J = phase_retarder_jones(retardance, charge / 2 * theta, 0) * (1 - hole)
            return torch.as_tensor(np.ascontiguousarray(J))

        instance_data.coronagraph = MultiScaleCoronagraph(input_grid, jones_mask, None, self.q, self.scaling_factor,
                                                          self.window_size, self.precision)

    def _polarized(self, wavefront):
        if wavefront.is_polarized:
            return wavefront
        return Wavefront(wavefront.electric_field, wavefront.wavelength, (1, 0, 0, 0))

    def _run(self, wavefront, conjugate):
        """The levels on the wavefront; a scalar wavefront is the unpolarized E (x) I, whose
        levels are computed from E itself."""
        d = self.get_instance_data(wavefront.grid, None, wavefront.wavelength)
        coro = d.coronagraph
        E = wavefront.electric_field
        if wavefront.is_scalar and not pointwise.needs_grad(E):
            L = coro._levels(E.as_subclass(torch.Tensor), conjugate)
            return Wavefront(Field(L, coro.input_grid), wavefront.wavelength, (1, 0, 0, 0))
        wf = self._polarized(wavefront)
        return coro.backward(wf) if conjugate else coro.forward(wf)

    def forward(self, wavefront):
        """Pupil -> Lyot plane (vortex.py:160-202)."""
        lyot = self._run(wavefront, False)
        if self.lyot_stop is not None:
            lyot = self.lyot_stop.forward(lyot)
        return lyot

    def backward(self, wavefront):
        """Lyot plane -> pupil (vortex.py:204-246): the stop's adjoint, then the levels with the
        conjugate Jones matrices (the retarder's Jones matrix is symmetric, so conj(J) = J^H)."""
        if self.lyot_stop is not None:
            wavefront = self.lyot_stop.backward(wavefront)
        return self._run(wavefront, True)
