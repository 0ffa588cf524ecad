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
* ``poke_jacobian``    : broadband DM-poke batch.  32 wavelengths lambda_0 (0.95 ... 1.05),
  lambda_0 = 500 nm, D = 1 m, f = 1 m; wavelength i uses the pupil size (256, 512, 1024)[i % 3]
  (focal grid q = 4 at lambda_0, n/16 lambda_0/D radius: 128^2, 256^2, 512^2, fixed in physical
  units, so every wavelength has its own Fourier matrices); circular pupil with the smoothed
  0.05 rad phase (seed 4); 32 x 32 Gaussian actuators (crosstalk 0.15, the reference's sigma =
  pitch / sqrt(-2 ln 0.15)); every actuator poked by a surface height of 1e-3 rad of phase at
  lambda_0.
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


POKE_LAMBDA0 = 500e-9
POKE_PUPILS = (256, 512, 1024)


def poke_jacobian(n_wavelengths=32, n_act=32, pupils=POKE_PUPILS, seed=4):
    """Geometry and base fields of the broadband DM-poke batch: a list of
    (pupil_grid, focal_grid, wavelength, field) with the focal grid in meters (f = 1 m), and the
    poke surface height h (2 k_0 h = 1e-3 rad) and the actuator count across."""
    lam0 = POKE_LAMBDA0
    wls = lam0 * np.linspace(0.95, 1.05, n_wavelengths)
    fields = {}
    for n in pupils:
        pg = make_pupil_grid(n, 1.0)
        ap = _circ(pg, 1.0)
        fields[n] = (pg, make_focal_grid(4, n // 16, spatial_resolution=lam0),
                     ap * np.exp(1j * smoothed_phase(pg.shape, 4.0 * n / 1024, 0.05, ap, seed)))
    items = []
    for i, wl in enumerate(wls):
        pg, fg, E = fields[pupils[i % len(pupils)]]
        items.append((pg, fg, wl, E))
    h = 1e-3 / (2 * 2 * np.pi / lam0)
    return items, h, n_act
