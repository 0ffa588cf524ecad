"""Direct power-of-two FFT (FastFourierTransform, hcipy/fourier/fast_fourier_transform.py:188-406)
on the device against the oracle's numpy FFT with the reference's pad, fftshifts, shift arrays
and crop, and against the dense Fourier sum on grids off the FFT lattice (GPU).

Tolerances: complex128 runs FP64 ramps, twiddles and line FFTs: 1e-12 of the output peak
against sums whose phases are exact.  On FFT-lattice grids (no shift) the test forms such sums
for sampled outputs, with the phase u_k x_n = pi (a2 + 2k)(b2 + 2n) / (2N) reduced in integers
(a2 = 2 u0 / du, b2 = 2 x0 / dx).  The oracle's ramps evaluate exp(-i u x) from rounded FP64
phases; the largest is its fftshift ramp on the internal grid, pi N / 2 radians per axis, so the
comparison with the oracle allows 1e-12 plus 2e-16 of that phase (5.1e-12 more at 8192^2).
complex64 with the default precision 'fp64' rounds only its loads and stores: 1e-6 (the FP32
input rounding of random fields).  precision 'fp32' runs FP32 line FFTs with twiddles rounded
from FP64: a few FP32 roundings per butterfly level relative to the field's norm, held to 2e-6
of the output peak on random fields."""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from hcipy_amd.fourier import _DirectFft
from oracle import hcipy_oracle as o

from helpers import dev, host, og, peak_rel_field, say

pytestmark = pytest.mark.gpu


def _crand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _oracle_phase(N):
    """Largest phase of the oracle's fftshift ramp on the internal grid: (N dx / 2)(N du / 2)."""
    return sum(np.pi * n / 2 for n in N)


def _exact_axis(g, go, N, d, ks):
    """exp(-i u_k x_n) for output indices ks of axis d, phase reduced in integers."""
    a2 = int(round(2 * go.zero[d] / go.delta[d]))
    b2 = int(round(2 * g.zero[d] / g.delta[d]))
    assert abs(2 * go.zero[d] / go.delta[d] - a2) < 1e-9 and abs(2 * g.zero[d] / g.delta[d] - b2) < 1e-9
    n = np.arange(int(g.dims[d]), dtype=np.int64)
    num = ((a2 + 2 * np.asarray(ks, dtype=np.int64))[:, None] * (b2 + 2 * n)[None, :]) % (4 * N[d])
    return np.exp(-1j * np.pi * num / (2 * N[d]))


def _exact_samples(f, g, go, N, seed):
    """w sum_n f[n] exp(-i u_k . x_n) on 8 sampled output columns and rows (on the lattice)."""
    rng = np.random.default_rng(seed)
    w = float(np.asarray(g.weights).ravel()[0])
    ks = [np.sort(rng.choice(int(go.dims[d]), size=min(8, int(go.dims[d])), replace=False)) for d in range(g.ndim)]
    ex = _exact_axis(g, go, N, 0, ks[0])
    if g.ndim == 1:
        return (ex @ f) * w, np.asarray(ks[0])
    ey = _exact_axis(g, go, N, 1, ks[1])
    F = f.reshape(int(g.dims[1]), int(g.dims[0]))
    idx = (ks[1][:, None] * int(go.dims[0]) + ks[0][None, :]).ravel()
    return ((ey @ F) @ ex.T).ravel() * w, idx


def _check_class(g, q, fov, shift, label, fp32=True):
    ft = ha.FastFourierTransform(g, q, fov, shift)
    assert ft.method == 'fft'
    ref_ft = o.FFT(og(g), q, fov, shift)
    f = _crand(g.size, 7)
    ref = ref_ft.forward(f)
    slack = 2e-16 * _oracle_phase(ft._impl._fft_n)
    cases = [(np.complex128, ft, 1e-12 + slack), (np.complex64, ft, 1e-6 + slack)]
    if fp32:
        cases.append((np.complex64, ha.FastFourierTransform(g, q, fov, shift, precision='fp32'), 2e-6 + slack))
    for dtype, t, tol in cases:
        out = host(t.forward(ha.Field(dev(f.astype(dtype)), g)))
        assert out.shape == ref.shape
        if dtype == np.complex128 and np.allclose(shift, 0):
            exact, idx = _exact_samples(f, g, ft.output_grid, ft._impl._fft_n, 3)
            peak = np.abs(ref).max()
            ee = np.abs(out[idx] - exact).max() / peak
            say('direct FFT %s sampled outputs vs exact-phase sums: device %.2e, oracle %.2e' %
                (label, ee, np.abs(ref[idx] - exact).max() / peak))
            assert ee < 1e-12
        e = peak_rel_field(out, ref)
        say('direct FFT %s %s %s: %.2e vs oracle' % (label, dtype.__name__, t._impl.precision or 'default', e))
        assert e < tol
    back = host(ft.backward(ha.Field(dev(ref), ft.output_grid)))
    eb = peak_rel_field(back, ref_ft.backward(ref))
    say('direct FFT %s backward: %.2e' % (label, eb))
    assert eb < 1e-12 + slack


@pytest.mark.parametrize('n,q,fov,shift', [(1024, 2, 1, 0), (512, 2, 0.5, 0.1), (2048, 1, 1, 0), (128, 4, 0.3, 0.25),
                                           (64, 1, 1, 0), (8, 2, 1, 0), (4096, 2, 0.25, 0)])
def test_fft_class_direct_matches_oracle(n, q, fov, shift):
    _check_class(ha.make_pupil_grid(n, 1.0), q, fov, shift, 'n=%d q=%g fov=%g shift=%g' % (n, q, fov, shift))


@pytest.mark.parametrize('dims,q,fov,shift', [([96, 40], (4 / 3, 1.6), (1, 0.5), (0.0, 0.3)),
                                              ([64, 256], 2, 1, (0.1, -0.2))])
def test_fft_class_direct_rectangular(dims, q, fov, shift):
    g = ha.make_uniform_grid(dims, [1.0, 2.0], center=[0.1, -0.05])
    _check_class(g, q, fov, shift, 'dims=%s' % (dims,))


@pytest.mark.parametrize('n,q,fov', [(4096, 2, 1), (100, 2.56, 0.7), (3, 4 / 3, 1)])
def test_fft_class_direct_1d(n, q, fov):
    """1-D grids finish in the row pass; 4096 samples at q = 2 need the 8192-point line FFT,
    beyond the zoom FFT's n_in + n_out - 1 <= 8192."""
    _check_class(ha.make_uniform_grid([n], [float(n)]), q, fov, 0.0, '1-D n=%d q=%g' % (n, q))


def test_direct_fft_off_lattice_grids_match_dense_sum():
    """Output grids whose first coordinate is off the FFT lattice (the ramps then use FP64
    phases, not the integer reduction), input sizes below the FFT length, against the dense
    separable Fourier sum w_in exp(-i u x) f."""
    rng = np.random.default_rng(3)
    N = (64, 128)
    n = (50, 37)
    dx = (0.7, 1.3)
    x0 = (rng.uniform(-20, 20), rng.uniform(-20, 20))
    m = (64, 100)
    du = tuple(2 * np.pi / (N[i] * dx[i]) for i in range(2))
    u0 = (rng.uniform(-3, 3), rng.uniform(-3, 3))
    g = ha.make_uniform_grid(list(n), [n[i] * dx[i] for i in range(2)], center=[x0[i] + (n[i] - 1) * dx[i] / 2 for i in range(2)])
    go = ha.make_uniform_grid(list(m), [m[i] * du[i] for i in range(2)], center=[u0[i] + (m[i] - 1) * du[i] / 2 for i in range(2)])
    ft = _DirectFft(g, go, N)
    f = _crand(g.size, 5)
    Ex = np.exp(-1j * np.outer(go.separated_coords[0], g.separated_coords[0]))
    Ey = np.exp(-1j * np.outer(go.separated_coords[1], g.separated_coords[1]))
    w = float(np.asarray(g.weights).ravel()[0])
    ref = (Ey @ f.reshape(n[1], n[0]) @ Ex.T).ravel() * w
    out = host(ft.forward(ha.Field(dev(f), g)))
    e = peak_rel_field(out, ref)
    phase = max(np.abs(go.separated_coords[i]).max() * np.abs(g.separated_coords[i]).max() for i in range(2))
    say('direct FFT off-lattice forward: %.2e (max phase %.0f rad)' % (e, phase))
    assert e < 1e-12 + 1e-15 * phase
    F = _crand(go.size, 6)
    wo = float(np.asarray(go.weights).ravel()[0]) / (2 * np.pi) ** 2
    refb = (Ey.conj().T @ F.reshape(m[1], m[0]) @ Ex.conj()).ravel() * wo
    eb = peak_rel_field(host(ft.backward(ha.Field(dev(F), go))), refb)
    say('direct FFT off-lattice backward: %.2e' % eb)
    assert eb < 1e-12 + 1e-15 * phase


def test_direct_fft_batches_and_autograd():
    """A (3, 2) batch of fields equals the fields one at a time, and the gradient of
    Re sum conj(v) F(x) is the dense adjoint of the forward sum."""
    g = ha.make_pupil_grid(48, 1.0)
    ft = ha.FastFourierTransform(g, 4 / 3, 1, 0.2)  # 64-point internal grid
    assert ft.method == 'fft'
    f = _crand((3, 2, g.size), 11)
    out = host(ft.forward(ha.Field(dev(f), g)))
    for i in range(3):
        for j in range(2):
            one = host(ft.forward(ha.Field(dev(f[i, j]), g)))
            assert np.array_equal(out[i, j], one)
    x = dev(f[0, 0]).requires_grad_(True)
    v = _crand(ft.output_grid.size, 12)
    y = ft.forward(ha.Field(x, g)).as_subclass(torch.Tensor)
    (y * dev(v).conj()).sum().real.backward()
    gx, gy = g.separated_coords
    ux, uy = ft.output_grid.separated_coords
    Ex = np.exp(-1j * np.outer(ux, gx))
    Ey = np.exp(-1j * np.outer(uy, gy))
    w = float(np.asarray(g.weights).ravel()[0])
    ref = (Ey.conj().T @ v.reshape(len(uy), len(ux)) @ Ex.conj()).ravel() * w
    e = peak_rel_field(host(x.grad), ref)
    say('direct FFT gradient vs dense adjoint: %.2e' % e)
    assert e < 1e-12
