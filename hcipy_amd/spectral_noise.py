"""FFT and multi-scale spectral noise: random screens with a given power spectral density, the
phase screens and surface errors that feed the propagation chains
(hcipy/util/spectral_noise.py:99-330, hcipy/optics/aberration.py:11-54).

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
from .fourier import FastFourierTransform, MatrixFourierTransform

__all__ = ['SpectralNoiseFactoryFFT', 'SpectralNoiseFFT', 'SpectralNoiseFactoryMultiscale', 'SpectralNoiseMultiscale',
           'make_power_law_error']


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


def _radius(grid):
    c = grid.coords
    return np.hypot(c[0], c[1]) if grid.ndim == 2 else np.abs(c[0])


def _amplitude(psd, grid):
    w = np.asarray(grid.weights, dtype=float)
    return np.sqrt(_host(psd(grid)).astype(float) / w * (2 * np.pi) ** grid.ndim)


class SpectralNoiseFactoryMultiscale:
    """Factory of multi-scale spectral-noise realizations (spectral_noise.py:203-266).

    The high frequencies sit on the FFT grid of ``output_grid`` with |k| below the boundary
    zeroed; the low frequencies on that grid scaled by 1/``oversampling`` (finer sampling, so
    the screen can be shifted by that fraction of the grid extent without wrapping) with |k| at
    or above the boundary zeroed.  The boundary is the largest |k_x| of the fine grid.  The
    screen is the FFT backward of the first part plus the MFT backward of the second, both on
    the device in FP64.
    """

    def __init__(self, psd, output_grid, oversampling):
        if not output_grid.is_regular:
            raise ValueError("Can't make a SpectralNoiseFactoryMultiscale on a non-regular grid.")
        self.psd = psd
        self.output_grid = output_grid
        self.oversampling = oversampling
        self.fourier_1 = FastFourierTransform(output_grid)
        self.input_grid_1 = self.fourier_1.output_grid
        self.input_grid_2 = self.input_grid_1.scaled(1.0 / oversampling)
        self.fourier_2 = MatrixFourierTransform(output_grid, self.input_grid_2)
        boundary = np.abs(self.input_grid_2.coords[0]).max()
        c1 = _amplitude(psd, self.input_grid_1)
        c1[_radius(self.input_grid_1) < boundary] = 0
        c2 = _amplitude(psd, self.input_grid_2)
        c2[_radius(self.input_grid_2) >= boundary] = 0
        dev = default_device()
        self.C_1 = torch.as_tensor(np.ascontiguousarray(c1), device=dev)
        self.C_2 = torch.as_tensor(np.ascontiguousarray(c2), device=dev)

    def make_random(self, seed=None):
        """One realization (spectral_noise.py:243-266): the noise for the FFT part (real, then
        imaginary parts), then for the MFT part, from one ``default_rng(seed)`` stream."""
        rng = np.random.default_rng(seed)
        n1, n2 = self.input_grid_1.size, self.input_grid_2.size
        w1 = rng.standard_normal(n1) + 1j * rng.standard_normal(n1)
        w2 = rng.standard_normal(n2) + 1j * rng.standard_normal(n2)
        dev = self.C_1.device
        C_1 = Field(pointwise.multiply(torch.as_tensor(w1, device=dev), self.C_1), self.input_grid_1)
        C_2 = Field(pointwise.multiply(torch.as_tensor(w2, device=dev), self.C_2), self.input_grid_2)
        return SpectralNoiseMultiscale(self, C_1, C_2)


class SpectralNoiseMultiscale:
    """A multi-scale realization: the coefficients of its FFT and MFT parts
    (spectral_noise.py:268-330)."""

    def __init__(self, factory, C_1, C_2):
        self.factory = factory
        self.C_1 = C_1
        self.C_2 = C_2

    def copy(self):
        return SpectralNoiseMultiscale(self.factory,
                                       Field(self.C_1.as_subclass(torch.Tensor).clone(), self.C_1.grid),
                                       Field(self.C_2.as_subclass(torch.Tensor).clone(), self.C_2.grid))

    def shift(self, shift):
        """In place: both parts times exp(-i shift . k), so that the screen moves by ``shift[i]``
        along axis i (spectral_noise.py:296-314).  The reference sums its per-axis phases as the
        np.ix_ broadcast of [nx][ny] and ravels that into the [ny][nx] coefficients, which on a
        square grid exchanges the two shift components; this is the translation its docstring
        describes."""
        s = np.zeros(2)
        s[:len(np.atleast_1d(shift))] = np.atleast_1d(shift)
        pointwise.phase_ramp(self.C_1, self.C_1.grid, -s[0], -s[1], out=self.C_1)
        pointwise.phase_ramp(self.C_2, self.C_2.grid, -s[0], -s[1], out=self.C_2)

    def shifted(self, shift):
        a = self.copy()
        a.shift(shift)
        return a

    def __call__(self):
        """The screen: real part of the FFT backward of C_1 plus that of the MFT backward of C_2
        (spectral_noise.py:316-330)."""
        a = self.factory.fourier_1.backward(self.C_1).as_subclass(torch.Tensor)
        b = self.factory.fourier_2.backward(self.C_2).as_subclass(torch.Tensor)
        return Field((a.real + b.real).contiguous(), self.factory.output_grid)


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
