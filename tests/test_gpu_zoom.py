"""Zoom FFT (Bluestein chirp-z per axis) on the device against the oracle, the matrix Fourier
transform and its own adjoint (GPU).

Tolerances: complex128 runs FP64 line FFTs and FP64 chirps, 1e-12 of the output peak against
the FP64 dense transform (the oracle's MFT, pinned to the dense DFT).  The oracle's restatement
of the reference's chirp-z (powers w^(k^2 / 2) of a unit complex number rounded to FP64) drifts
by ~k^2 * 1e-16 in modulus: 2e-10 at k = 5000, so the device is held to it at 1e-12 plus that
bound.  complex64 runs FP32 line FFTs of length L <= 8192 with chirps rounded from
FP64 (precision 'fp32'); its error is a few FP32 roundings per butterfly level relative to the
field's norm, held here to 2e-6 of the output peak on random fields (whose peak is close to
their norm) and 1e-6 peak-relative intensity on the 1024^2 -> 512^2 Fraunhofer workload.  The
default 'fp64' complex64 path rounds only its loads and stores: 1e-6 on random fields (the
FP32 input rounding) and 1e-7 peak-relative intensity on the Fraunhofer workload."""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from oracle import hcipy_oracle as o

from helpers import dev, host, og, peak_rel_field, say

pytestmark = pytest.mark.gpu

_CASES = [
    ([12, 9], [1.1, 0.8], [0.05, -0.02], [7, 15], [20.0, 30.0], [1.0, -0.5]),
    ([10], [1.0], [0.0], [23], [40.0], [0.3]),
    ([64, 48], [1.0, 0.75], [0.0, 0.0], [200, 130], [90.0, 60.0], [0.2, -0.4]),
    ([5000], [1.0], [0.1], [3000], [300.0], [-2.0]),  # line FFT length 8192
]


def _grids(c):
    di, ei, ci, do, eo, co = c
    return ha.make_uniform_grid(di, ei, center=ci), ha.make_uniform_grid(do, eo, center=co)


def _crand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize('case', range(len(_CASES)))
@pytest.mark.parametrize('dtype,precision,tol', [(np.complex128, None, 1e-12), (np.complex64, 'fp64', 1e-6),
                                                 (np.complex64, 'fp32', 2e-6)])
def test_zoom_matches_oracle(case, dtype, precision, tol):
    pg, fg = _grids(_CASES[case])
    say('zoom case', case, 'dims', tuple(pg.dims), '->', tuple(fg.dims), dtype.__name__, precision)
    zf = ha.ZoomFastFourierTransform(pg, fg, precision=precision)
    f = _crand((3, pg.size), case)
    out = host(zf.forward(ha.Field(dev(f.astype(dtype)), pg)))
    f128 = f.astype(dtype).astype(np.complex128)
    ref = o.mft_forward(f128, og(pg), og(fg))
    assert out.shape == (3, fg.size) and out.dtype == dtype
    err = peak_rel_field(out, ref)
    drift = 1e-16 * max(max(pg.dims), max(fg.dims)) ** 2
    err_zoom = peak_rel_field(out, o.zoom_fft(f128, og(pg), og(fg)))
    say('  forward error vs dense %.2e, vs the chirp-z restatement %.2e' % (err, err_zoom))
    assert err < tol
    assert err_zoom < tol + 10 * drift
    F = _crand((2, 2, fg.size), case + 10)
    back = host(zf.backward(ha.Field(dev(F.astype(dtype)), fg)))
    F128 = F.astype(dtype).astype(np.complex128)
    refb = o.mft_backward(F128, og(pg), og(fg))
    assert back.shape == (2, 2, pg.size)
    assert peak_rel_field(back, refb) < tol
    assert peak_rel_field(back, o.zoom_fft(F128, og(pg), og(fg), adjoint=True)) < tol + 10 * drift


def test_zoom_equals_mft_and_adjoint():
    """The zoom FFT and the MFT compute the same sums; <Z f, g> = <f, Z^dagger g> with the
    dagger the backward transform up to the weights; the autograd gradient is that adjoint."""
    pupil = ha.make_pupil_grid(96)
    focal = ha.make_focal_grid(3, 10)
    zf = ha.ZoomFastFourierTransform(pupil, focal)
    mft = ha.MatrixFourierTransform(pupil, focal)
    f = _crand(pupil.size, 1)
    a = host(zf.forward(ha.Field(dev(f), pupil)))
    b = host(mft.forward(ha.Field(dev(f), pupil)))
    assert peak_rel_field(a, b) < 1e-12
    g = _crand(focal.size, 2)
    assert peak_rel_field(host(zf.backward(ha.Field(dev(g), focal))), host(mft.backward(ha.Field(dev(g), focal)))) < 1e-12
    wi = float(np.asarray(pupil.weights).ravel()[0])
    wo = float(np.asarray(focal.weights).ravel()[0]) / (2 * np.pi) ** 2
    lhs = np.vdot(g, a)
    rhs = np.vdot(host(zf.backward(ha.Field(dev(g), focal))), f) * wi / wo
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    x = dev(f).requires_grad_(True)
    y = zf.forward(ha.Field(x, pupil)).as_subclass(torch.Tensor)
    torch.real(torch.sum(torch.conj(dev(g)) * y)).backward()
    assert peak_rel_field(host(x.grad), host(zf.backward(ha.Field(dev(g), focal))) * wi / wo) < 1e-12


def test_zoom_fraunhofer_1024_to_512():
    """1024^2 pupil -> 512^2 focal plane (q = 4, 64 lambda/D): complex128 against the FP64 MFT,
    complex64 at <= 1e-7 peak-relative intensity error."""
    pupil = ha.make_pupil_grid(1024)
    focal = ha.make_focal_grid(4, 64)
    ap = ((pupil.x ** 2 + pupil.y ** 2) <= 0.25).astype(float)
    rng = np.random.default_rng(3)
    E = ap * np.exp(1j * 0.2 * rng.standard_normal(pupil.size))
    ref = host(ha.MatrixFourierTransform(pupil, focal).forward(ha.Field(dev(E), pupil)))
    zf = ha.ZoomFastFourierTransform(pupil, focal)
    z128 = host(zf.forward(ha.Field(dev(E), pupil)))
    assert peak_rel_field(z128, ref) < 1e-12
    Ir = np.abs(ref) ** 2
    errs = {}
    for prec in ('fp64', 'fp32'):
        z64 = host(ha.ZoomFastFourierTransform(pupil, focal, precision=prec).forward(ha.Field(dev(E.astype(np.complex64)), pupil)))
        errs[prec] = np.abs(np.abs(z64.astype(np.complex128)) ** 2 - Ir).max() / Ir.max()
        say('zoom 1024->512 complex64 %s peak-relative intensity error %.3e' % (prec, errs[prec]))
    assert errs['fp64'] < 1e-7
    assert errs['fp32'] < 1e-6


def test_zoom_empty_batch_and_planner():
    pg, fg = _grids(_CASES[0])
    zf = ha.ZoomFastFourierTransform(pg, fg)
    y = zf.forward(ha.Field(torch.zeros((0, pg.size), dtype=torch.complex64, device='cuda'), pg))
    assert tuple(y.shape) == (0, fg.size)


def test_planner_measure():
    """planner='measure' times each candidate on the device and returns one that computes the
    same transform."""
    pupil = ha.make_pupil_grid(128)
    uv = ha.FraunhoferPropagator(pupil, ha.make_focal_grid(4, 8)).instance(1.0).uv_grid
    ft = ha.make_fourier_transform(pupil, uv, planner='measure')
    say('planner measure chose', type(ft).__name__)
    f = _crand(pupil.size, 4)
    a = host(ft.forward(ha.Field(dev(f), pupil)))
    b = host(ha.MatrixFourierTransform(pupil, uv).forward(ha.Field(dev(f), pupil)))
    assert peak_rel_field(a, b) < 1e-12


def _all_transforms(g, go):
    return [ha.MatrixFourierTransform(g, go), ha.FastFourierTransform(g, *ha.get_fft_parameters(go, g)),
            ha.ZoomFastFourierTransform(g, go)]


@pytest.mark.parametrize('dims,shift_in,scale,shift_out', [([9, 18], [0.1, 0.1], 2, [0.1, 0.1]), ([8, 16], [0, 0], 1, [0, 0]),
                                                          ([65], [0.1], 2, [0.1])])
@pytest.mark.parametrize('dtype', ['complex128', 'complex64'])
def test_energy_conservation_across_methods(dims, shift_in, scale, shift_out, dtype):
    """tests/test_fourier.py:22-96 with the zoom FFT among the methods: at q = 1, fov = 1 every
    forward (MFT, FFT, zoom FFT) followed by every backward is the identity and conserves
    energy (c128 1e-12; c64 1e-6 pattern, 1e-5 energy)."""
    g = ha.make_uniform_grid(dims, 1, has_center=True).shifted(shift_in).scaled(scale)
    go = ha.make_fft_grid(g, 1, 1, shift_out)
    f = np.random.default_rng(0).standard_normal(g.size).astype(dtype)
    e_in = np.sum(np.abs(f) ** 2) * np.asarray(g.weights).ravel()[0]
    fts = _all_transforms(g, go)
    for a in fts:
        F = a.forward(ha.Field(dev(f), g))
        for b in fts:
            back = host(b.backward(F))
            assert back.dtype == np.dtype(dtype)
            e_out = np.sum(np.abs(back) ** 2) * np.asarray(g.weights).ravel()[0]
            pat = np.std(np.abs(back - f)) / np.abs(f).mean()
            if dtype == 'complex128':
                assert abs(e_out / e_in - 1) < 1e-12 and pat < 1e-12, (type(a).__name__, type(b).__name__)
            else:
                assert abs(e_out / e_in - 1) < 1e-5 and pat < 1e-6, (type(a).__name__, type(b).__name__)


@pytest.mark.parametrize('dtype', ['complex128', 'complex64'])
def test_zoom_symmetry_and_linearity(dtype):
    """tests/test_fourier.py:98-144 for the zoom FFT: even inputs give real, odd inputs imaginary
    transforms on a symmetric grid; forward is linear."""
    tol = 1e-12 if dtype == 'complex128' else 1e-6
    g = ha.make_uniform_grid([9, 9], 1, has_center=False).scaled(2)
    go = ha.make_fft_grid(g, 3, 0.8)
    f = np.random.default_rng(0).standard_normal(g.size)
    fe = ((f + f[::-1]) / 2).astype(dtype)
    fo = ((f - f[::-1]) / 2).astype(dtype)
    zf = ha.ZoomFastFourierTransform(g, go)
    Fe = host(zf.forward(ha.Field(dev(fe), g)))
    Fo = host(zf.forward(ha.Field(dev(fo), g)))
    assert np.abs(Fe.imag).max() < tol * max(1, np.abs(Fe).max())
    assert np.abs(Fo.real).max() < tol * max(1, np.abs(Fo).max())
    Fb = host(zf.forward(ha.Field(dev(f.astype(dtype)), g)))
    assert np.allclose(Fb, Fe + Fo, atol=tol * np.abs(Fb).max())
    assert np.allclose(host(zf.forward(ha.Field(dev((2 * f).astype(dtype)), g))), 2 * Fb, atol=tol * np.abs(Fb).max())


def test_zoom_1d_matches_scipy_czt():
    """The 1-D zoom FFT is a chirp-z transform times the output phase shift
    (zoom_fast_fourier_transform.py:42-68): against the dense sum at 1e-12 and against
    scipy.signal.CZT (the reference's own yardstick, tests/test_fourier.py:389-424) within
    scipy's k^2 drift, for random lengths and spacings."""
    import scipy.signal
    rng = np.random.default_rng(1)
    for _ in range(4):
        n = int(rng.integers(50, 3000))
        m = int(rng.integers(50, 3000))
        dx, du = rng.uniform(0.5, 2.0), rng.uniform(0.001, 0.02)
        x0, u0 = rng.uniform(-3, 3), rng.uniform(-5, 5)
        g = ha.make_uniform_grid([n], [n * dx], center=[x0 + (n - 1) * dx / 2])
        go = ha.make_uniform_grid([m], [m * du], center=[u0 + (m - 1) * du / 2])
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out = host(ha.ZoomFastFourierTransform(g, go).forward(ha.Field(dev(x), g)))
        w = np.exp(-1j * du * dx)
        a = np.exp(1j * go.zero[0] * dx)
        ref = scipy.signal.CZT(n, m, w, a)(x) * np.exp(-1j * go.separated_coords[0] * g.zero[0]) * dx
        dense = np.exp(-1j * np.outer(go.separated_coords[0], g.separated_coords[0])) @ x * dx
        peak = np.abs(dense).max()
        # phases u x reach max|u| max|x| radians, rounded at 1e-16 of that in both the device's
        # chirps and the dense sum; scipy forms w^(k^2/2) by powers and drifts by ~k^2 * 1e-16
        phase = np.abs(go.separated_coords[0]).max() * np.abs(g.separated_coords[0]).max()
        assert np.abs(out - dense).max() < (1e-12 + 1e-15 * phase) * peak, (n, m)
        assert np.abs(out - ref).max() < (1e-12 + 1e-15 * (n + m) ** 2) * peak, (n, m)


@pytest.mark.parametrize('n,q,fov,shift', [(1000, 2, 1, 0), (500, 2, 0.5, 0.1), (96, 3, 0.7, 0.0), (2000, 1, 1, 0)])
def test_fft_class_on_zoom_backend_matches_oracle(n, q, fov, shift):
    """FastFourierTransform (fast_fourier_transform.py:188-406) on the zoom-FFT backend against the
    oracle's numpy FFT with the reference's pad, shift and crop: complex128 at 1e-12 of the
    peak plus the oracle's own phase rounding, complex64 (FP64 internal) at 1e-6; backward
    against the oracle too.  The split-FP16
    MFT backend gives the same transform within its own bound."""
    g = ha.make_pupil_grid(n, 1.0)
    ft = ha.FastFourierTransform(g, q, fov, shift)
    assert ft.method == 'zft'
    ref_ft = o.FFT(og(g), q, fov, shift)
    f = _crand(g.size, 7)
    ref = ref_ft.forward(f)
    # the oracle's fftshift ramp rounds phases of up to pi N / 2 per axis at 1e-16
    slack = 2e-16 * sum(np.pi * int(d) / 2 for d in ft.internal_grid.dims)
    for dtype, tol in ((np.complex128, 1e-12 + slack), (np.complex64, 1e-6)):
        out = host(ft.forward(ha.Field(dev(f.astype(dtype)), g)))
        e = peak_rel_field(out, ref)
        say('FFT on zoom n=%d q=%g fov=%g %s: %.2e' % (n, q, fov, dtype.__name__, e))
        assert e < tol
    back = host(ft.backward(ha.Field(dev(ref), ft.output_grid)))
    assert peak_rel_field(back, ref_ft.backward(ref)) < 1e-12 + slack
    if n <= 512:
        m = host(ha.FastFourierTransform(g, q, fov, shift, precision='f16x2').forward(ha.Field(dev(f.astype(np.complex64)), g)))
        assert peak_rel_field(m, ref) < 2e-6
