"""FFT spectral noise: random screens with a given power spectral density, the phase screens and
surface errors that feed the propagation chains (hcipy/util/spectral_noise.py:99-201,
hcipy/optics/aberration.py:11-54).

Set-up follows the reference on the host in FP64: the PSD is evaluated on the FFT grid of the
output grid and the white noise is drawn with numpy's ``default_rng(seed)``, so that one seed
gives the reference's realization.  The per-realization work runs on the device: the noise times
the PSD amplitude (``hcipy_pw_sub_mul``), the shifts as FP64 phase ramps
(``hcipy_pw_phase_ramp``) and the evaluation as the weighted inverse transform of the FFT grid
onto the output grid (the MFT of ``FastFourierTransform.backward``).
"""
import numpy as np
import torch

from . import pointwise
from .aperture import make_circular_aperture
from .field import Field, default_device
from .fourier import FastFourierTransform

__all__ = ['SpectralNoiseFactoryFFT', 'SpectralNoiseFFT', 'make_power_law_error']


def _host(a):
    if isinstance(a, torch.Tensor):
        return a.as_subclass(torch.Tensor).detach().cpu().numpy()
    return np.asarray(a)


class SpectralNoiseFactoryFFT:
    """Factory of FFT spectral-noise realizations (spectral_noise.py:99-165).

    ``psd(grid)`` returns the power spectral density (per Hertz) on a Fourier grid; the factory
    evaluates it on the FFT grid of ``output_grid`` with zero padding ``oversample`` and keeps
    ``C = sqrt(psd / weights * (2 pi)^ndim)`` on the device.  The screens wrap at the edges of
    the (oversampled) grid, as the reference's do.
    """

    def __init__(self, psd, output_grid, oversample=1):
        if not output_grid.is_regular:
            raise ValueError("Can't make a SpectralNoiseFactoryFFT on a non-regular grid.")
        self.psd = psd
        self.output_grid = output_grid
        self.oversample = oversample
        self.fourier = FastFourierTransform(output_grid, q=oversample)
        self.input_grid = self.fourier.output_grid
        self.period = np.asarray(output_grid.delta) * np.asarray(output_grid.shape)
        w = np.asarray(self.input_grid.weights, dtype=float)
        amp = np.sqrt(_host(psd(self.input_grid)).astype(float) / w * (2 * np.pi) ** self.input_grid.ndim)
        self.C = torch.as_tensor(np.ascontiguousarray(amp), device=default_device())

    def make_random(self, seed=None):
        """One realization (spectral_noise.py:133-165): C times complex white noise drawn with
        ``numpy.random.default_rng(seed)`` (real and imaginary parts from one stream, in that
        order)."""
        rng = np.random.default_rng(seed)
        n = self.input_grid.size
        noise = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        t = torch.as_tensor(noise, device=self.C.device)
        return SpectralNoiseFFT(self, Field(pointwise.multiply(t, self.C), self.input_grid))


class SpectralNoiseFFT:
    """A realization: its Fourier coefficients on the FFT grid (spectral_noise.py:167-201)."""

    def __init__(self, factory, C):
        self.factory = factory
        self.C = C

    def copy(self):
        return SpectralNoiseFFT(self.factory, Field(self.C.as_subclass(torch.Tensor).clone(), self.C.grid))

    def shift(self, shift):
        """In place: C *= exp(-i shift . k), so that the screen moves by ``shift[i]`` along axis i
        (spectral_noise.py:177-187).  The reference sums its per-axis phases with
        np.add.reduce(np.ix_(...)), which raises on numpy >= 1.24 and on older numpy ravels the
        [nx][ny] sum into the [ny][nx] field (the shift components land on the other axes); this
        is the translation its docstring describes."""
        s = np.zeros(2)
        s[:len(np.atleast_1d(shift))] = np.atleast_1d(shift)
        pointwise.phase_ramp(self.C, self.C.grid, -s[0], -s[1], out=self.C)

    def shifted(self, shift):
        a = self.copy()
        a.shift(shift)
        return a

    def __call__(self):
        """The screen on the output grid: the real part of the inverse transform of C
        (spectral_noise.py:189-201)."""
        out = self.factory.fourier.backward(self.C)
        return Field(out.as_subclass(torch.Tensor).real.contiguous(), self.factory.output_grid)


def make_power_law_error(pupil_grid, ptv, diameter, exponent=-2.5, aperture=None, remove_modes=None, seed=None):
    """Surface error with a power-law PSD r^exponent, scaled to a peak-to-valley ``ptv`` over
    ``aperture`` (a circle of ``diameter`` by default) and multiplied by it
    (aberration.py:11-54).  ``seed`` picks the realization (the reference always draws a fresh
    one).  Removing modes (``remove_modes``) needs mode bases and a Tikhonov inverse, which are
    not part of this library."""
    if remove_modes is not None:
        raise NotImplementedError('remove_modes needs mode bases and inverse_tikhonov, which this library does not provide')

    def psd(grid):
        r = np.hypot(*grid.coords) if grid.ndim == 2 else np.abs(grid.coords[0])
        with np.errstate(divide='ignore'):
            res = r ** exponent
        res[r == 0] = 0
        return res

    if aperture is None:
        aperture = make_circular_aperture(diameter)(pupil_grid)
    ap = _host(aperture).astype(float)
    screen = SpectralNoiseFactoryFFT(psd, pupil_grid).make_random(seed)()
    # the peak-to-valley normalisation and the aperture product are set-up on the host copy, in
    # FP64, as the reference computes them
    sh = _host(screen)
    sh = sh * (ptv / np.ptp(sh[ap != 0]))
    return Field(torch.as_tensor(sh * ap, device=screen.device), pupil_grid)
