"""Seeded synthetic inputs of the benchmark workloads (host numpy, FP64).

* ``plain_mft_256``   : 256^2 circular pupil (D = 1, unit amplitude inside), phase = N(0,1)
  per pixel (default_rng(1)) smoothed by a Gaussian of sigma 4 px, 0.1 rad RMS; focal grid
  q = 4, 16 lambda/D radius (128^2); lambda = f = 1.
* ``plain_mft_1024``  : 1024^2 pupil, circular aperture with 4 spider vanes of width 0.01 D,
  same phase recipe at 0.05 rad RMS (seed 2); focal grid q = 4, 64 lambda/D (512^2).
  Its angular-spectrum companion: lambda = 500 nm, D = 1 m, z = 1 km, 2x zero padding.
* ``lyot_1024``       : 1024^2 circular pupil, power-law phase (PSD exponent -2, 0.01 rad RMS,
  seed 3); focal-plane mask 1 - circ(6 lambda/D) on a 256^2 grid (q = 32, 4 lambda/D);
  Lyot stop circ(0.95 D); final focal grid q = 4, 64 lambda/D (512^2).
"""
import numpy as np

from .field import make_focal_grid, make_pupil_grid


def _circ(grid, d):
    x, y = grid.separated_coords
    return ((x[np.newaxis, :] ** 2 + y[:, np.newaxis] ** 2) <= (d / 2) ** 2).ravel().astype(float)


def _spiders(grid, num, width, length=1.0):
    x, y = grid.separated_coords
    x = x[np.newaxis, :]
    y = y[:, np.newaxis]
    out = np.ones(grid.size)
    for ang in np.linspace(0, 2 * np.pi, num, endpoint=False):
        # make_spider((0, 0), (cos, sin) * D, w): centre at half length (hcipy/aperture/generic.py:301-348)
        p2 = np.array([np.cos(ang), np.sin(ang)]) * length
        shift = p2 / 2
        a = np.arctan2(p2[1], p2[0])
        xs = x - shift[0]
        ys = y - shift[1]
        xn = xs * np.cos(a) + ys * np.sin(a)
        yn = ys * np.cos(a) - xs * np.sin(a)
        s = (xn <= length / 2) * (xn >= -length / 2) * (yn <= width / 2) * (yn >= -width / 2)
        out = out * (1 - s.ravel())
    return out


def smoothed_phase(shape, sigma_px, rms, mask, seed):
    rng = np.random.default_rng(seed)
    n = rng.standard_normal(shape)
    ky = np.fft.fftfreq(shape[0])[:, None]
    kx = np.fft.fftfreq(shape[1])[None, :]
    g = np.exp(-2 * np.pi ** 2 * sigma_px ** 2 * (kx ** 2 + ky ** 2))
    ph = np.real(np.fft.ifft2(np.fft.fft2(n) * g))
    m = np.asarray(mask).reshape(shape) > 0
    ph *= rms / np.sqrt(np.mean(ph[m] ** 2))
    return ph.ravel()


def power_law_phase(shape, exponent, rms, mask, seed):
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(shape[0])[:, None]
    kx = np.fft.fftfreq(shape[1])[None, :]
    f = np.hypot(kx, ky)
    f[0, 0] = np.inf
    c = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * f ** (exponent / 2)
    ph = np.real(np.fft.ifft2(c))
    m = np.asarray(mask).reshape(shape) > 0
    ph -= ph[m].mean()
    ph *= rms / np.sqrt(np.mean(ph[m] ** 2))
    return ph.ravel()


def plain_mft_256():
    pupil = make_pupil_grid(256, 1.0)
    focal = make_focal_grid(4, 16)
    ap = _circ(pupil, 1.0)
    ph = smoothed_phase(pupil.shape, 4.0, 0.1, ap, 1)
    return pupil, focal, ap * np.exp(1j * ph)


def plain_mft_1024():
    pupil = make_pupil_grid(1024, 1.0)
    focal = make_focal_grid(4, 64)
    ap = _circ(pupil, 1.0) * _spiders(pupil, 4, 0.01)
    ph = smoothed_phase(pupil.shape, 4.0, 0.05, ap, 2)
    return pupil, focal, ap * np.exp(1j * ph)


def lyot_1024():
    pupil = make_pupil_grid(1024, 1.0)
    ap = _circ(pupil, 1.0)
    ph = power_law_phase(pupil.shape, -2.0, 0.01, ap, 3)
    fpm_grid = make_focal_grid(32, 4)
    fpm = 1 - _circ(fpm_grid, 6.0)
    stop = _circ(pupil, 0.95)
    focal = make_focal_grid(4, 64)
    return pupil, fpm_grid, focal, ap * np.exp(1j * ph), fpm, stop, ap
