"""FourierFilter / angular-spectrum / Fresnel propagation against the oracle (GPU).

Tolerances: complex128 1e-12 of the field maximum (1e-9 where the transfer-function phase
k_z z exceeds 1e6 rad: one ulp of k_z z is then above 1e-10 rad, and the device and host libm
hypot/sin/cos may differ by an ulp); complex64 1e-5 (FP32 butterflies over log2 N passes).
"""
import numpy as np
import pytest
import torch

import hcipy_amd as ha
from hcipy_amd import workloads
from oracle import hcipy_oracle as o

from helpers import dev, host, og, peak_rel_field, say

pytestmark = pytest.mark.gpu


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize('n,q', [(16, 1), (16, 2), (17, 2), (64, 4), (16, 3), (32, 2)])
@pytest.mark.parametrize('dtype', [np.complex128, np.complex64])
def test_fourier_filter_random_tf(n, q, dtype):
    g = ha.make_pupil_grid(n, 2.0)
    internal = ha.make_fft_grid(g, q, 1)
    tf = _rand(internal.size, 1)
    f = _rand((3, g.size), 2).astype(dtype)
    filt = ha.FourierFilter(g, ha.Field(dev(tf), internal), q)
    tol = 1e-12 if dtype == np.complex128 else 1e-5
    for adjoint in [False, True]:
        a = host(filt.backward(ha.Field(dev(f), g)) if adjoint else filt.forward(ha.Field(dev(f), g)))
        assert a.dtype == dtype
        b = o.fourier_filter(f.astype(np.complex128), og(g), tf, q, adjoint=adjoint)
        say('filter n=%d q=%d %s adjoint=%s err %.2e' % (n, q, dtype.__name__, adjoint, peak_rel_field(a, b)))
        assert peak_rel_field(a, b) < tol


@pytest.mark.parametrize('which', ['angular_spectrum', 'fresnel'])
@pytest.mark.parametrize('dtype', [np.complex128, np.complex64])
@pytest.mark.parametrize('rel', [0.2, 5])
@pytest.mark.parametrize('n', [256, 96])
def test_nearfield_propagators_vs_oracle(which, dtype, rel, n):
    """rel = 0.2: transfer-function branch; rel = 5: impulse-response branch (the grid step is
    below lambda z / L).  n = 96 pads to 192^2, not a power of two: the filter then runs on
    FastFourierTransform with the device-generated transfer function in grid order."""
    wl, a, b = 500e-9, 0.001, 0.001
    g = ha.make_pupil_grid(n, [16 * a, 16 * b])
    zc = np.min(g.delta * np.array([16 * a, 16 * b]) / wl)
    z = rel * zc
    ap = o.evaluate_supersampled(lambda gg: o.rectangular_aperture(gg, [2 * a, 2 * b]), og(g), 2)
    cls = ha.AngularSpectrumPropagator if which == 'angular_spectrum' else ha.FresnelPropagator
    prop = cls(g, z)
    wf = ha.Wavefront(ha.Field(dev(ap.astype(dtype)), g), wl)
    fwd = host(prop.forward(wf).electric_field)
    bwd = host(prop.backward(wf).electric_field)
    if which == 'angular_spectrum':
        ref_f = o.angular_spectrum_forward(ap, og(g), wl, z)
        ref_b = o.angular_spectrum_forward(ap, og(g), wl, z, adjoint=True)
    else:
        ref_f = o.fresnel_forward(ap, og(g), wl, z)
        ref_b = o.fresnel_forward(ap, og(g), wl, z, adjoint=True)
    tol = 1e-9 if dtype == np.complex128 else 2e-5
    say('%s %s fwd %.2e bwd %.2e' % (which, dtype.__name__, peak_rel_field(fwd, ref_f), peak_rel_field(bwd, ref_b)))
    assert peak_rel_field(fwd, ref_f) < tol
    assert peak_rel_field(bwd, ref_b) < tol


@pytest.mark.parametrize('rel,thr', [(0.2, 0.02), (5, 0.01)])
def test_nearfield_rectangle_closed_form(rel, thr):
    """Fresnel-integral closed form of a rectangle, bars 0.02 at z = 0.2 z_c and 0.01 at 5 z_c
    (impulse-response branch) (tests/test_propagation.py:64-138)."""
    import scipy.special
    n, wl, a, b = 512, 500e-9, 0.001, 0.001
    g = ha.make_pupil_grid(n, [16 * a, 16 * b])
    zc = np.min(g.delta * np.array([16 * a, 16 * b]) / wl)
    z = rel * zc
    ap = ha.evaluate_supersampled(ha.make_rectangular_aperture([2 * a, 2 * b]), g, 2)
    for cls in [ha.FresnelPropagator, ha.AngularSpectrumPropagator]:
        img = ha.subsample_field(cls(g, z).forward(ha.Wavefront(ap, wl)).intensity, 8)
        k = 2 * np.pi / wl

        def gen(gg):
            s = np.sqrt(2 / (z * wl))
            fr = scipy.special.fresnel
            S1, C1 = fr(s * (gg.x - a))
            S2, C2 = fr(s * (gg.x + a))
            S3, C3 = fr(s * (gg.y - b))
            S4, C4 = fr(s * (gg.y + b))
            r = ((C2 + 1j * S2) - (C1 + 1j * S1)) * ((C4 + 1j * S4) - (C3 + 1j * S3)) * 0.5
            return np.abs(r) ** 2
        ref = o.evaluate_supersampled(gen, og(g), 2)
        ref = ref.reshape(n // 8, 8, n // 8, 8).mean(axis=(1, 3)).ravel()
        err = np.abs(host(img) - ref).max()
        say('%s z = %g z_c: closed-form error %.3e' % (cls.__name__, rel, err))
        assert err < thr


def test_angular_spectrum_1024_1km():
    """1024^2 field, z = 1 km, 2x zero padding (2048^2 internal), complex64 and complex128."""
    pupil, focal, E = workloads.plain_mft_1024()
    wl, z = 500e-9, 1000.0
    assert o.angular_spectrum_uses_native_tf(og(pupil), wl, z)
    ref = o.angular_spectrum_forward(E, og(pupil), wl, z)
    for dtype, tol in [(np.complex128, 1e-9), (np.complex64, 2e-5)]:
        prop = ha.AngularSpectrumPropagator(pupil, z)
        out = host(prop(ha.Wavefront(ha.Field(dev(E.astype(dtype)), pupil), wl)).electric_field)
        e = peak_rel_field(out, ref)
        say('AS 1024 1 km %s: field peak-rel %.3e' % (dtype.__name__, e))
        assert e < tol


@pytest.mark.parametrize('dtype,tol', [(np.complex128, 1e-12), (np.complex64, 2e-6)])
@pytest.mark.parametrize('n,q', [(16, 1), (32, 2)])
def test_matrix_transfer_function(n, q, dtype, tol):
    """A (2 x 3) matrix transfer function (fourier_operations.py:120-131) on a 2-tensor field of
    3 x 4 components: the first index is contracted (field_dot), against the oracle; the adjoint
    maps the 2 x 4 result back to 3 x 4."""
    g = ha.make_pupil_grid(n, 2)
    internal = o.make_fft_grid(og(g), q)
    rng = np.random.default_rng(8)
    tf = rng.standard_normal((2, 3, internal.size)) + 1j * rng.standard_normal((2, 3, internal.size))
    f = (rng.standard_normal((3, 4, g.size)) + 1j * rng.standard_normal((3, 4, g.size))).astype(dtype)
    filt = ha.FourierFilter(g, dev(tf), q)
    out = host(filt.forward(ha.Field(dev(f), g)))
    ref = o.fourier_filter(f.astype(np.complex128), og(g), tf, q)
    assert out.shape == (2, 4, g.size)
    say('matrix filter n=%d q=%d %s fwd %.2e' % (n, q, dtype.__name__, peak_rel_field(out, ref)))
    assert peak_rel_field(out, ref) < tol
    back = host(filt.backward(ha.Field(dev(out.astype(dtype)), g)))
    refb = o.fourier_filter(out.astype(np.complex128), og(g), tf, q, adjoint=True)
    assert back.shape == (3, 4, g.size)
    assert peak_rel_field(back, refb) < tol
    if dtype == np.complex128:
        # gradient through the recorded path is the adjoint filter
        x = dev(f).requires_grad_(True)
        y = filt.forward(ha.Field(x, g)).as_subclass(torch.Tensor)
        u = dev(rng.standard_normal(y.shape) + 1j * rng.standard_normal(y.shape))
        torch.real(torch.sum(torch.conj(u) * y)).backward()
        assert peak_rel_field(host(x.grad), host(filt.backward(ha.Field(u, g)))) < 1e-12


@pytest.mark.parametrize('dtype,tol', [(np.complex128, 1e-12), (np.complex64, 2e-6)])
@pytest.mark.parametrize('shape', [(1, 4), (4, 1), (4, 4), (5, 3), (6, 6)])
def test_matrix_transfer_function_shapes(shape, dtype, tol):
    """Matrix transfer functions from 1 x 4 to 6 x 6 against the oracle, both directions, with a
    batch dimension: up to 4 x 4 is one hcipy_filter_execute_matrix call (each input transformed
    once), larger ones run in 4 x 4 blocks whose partial sums are added."""
    p_, q_ = shape
    g = ha.make_pupil_grid(32, 2)
    internal = o.make_fft_grid(og(g), 2)
    rng = np.random.default_rng(10 + p_ * 7 + q_)
    tf = rng.standard_normal((p_, q_, internal.size)) + 1j * rng.standard_normal((p_, q_, internal.size))
    f = (rng.standard_normal((q_, 3, g.size)) + 1j * rng.standard_normal((q_, 3, g.size))).astype(dtype)
    filt = ha.FourierFilter(g, dev(tf), 2)
    out = host(filt.forward(ha.Field(dev(f), g)))
    ref = o.fourier_filter(f.astype(np.complex128), og(g), tf, 2)
    assert out.shape == (p_, 3, g.size)
    e = peak_rel_field(out, ref)
    u = (rng.standard_normal((p_, 3, g.size)) + 1j * rng.standard_normal((p_, 3, g.size))).astype(dtype)
    back = host(filt.backward(ha.Field(dev(u), g)))
    refb = o.fourier_filter(u.astype(np.complex128), og(g), tf, 2, adjoint=True)
    assert back.shape == (q_, 3, g.size)
    eb = peak_rel_field(back, refb)
    say('matrix filter %dx%d %s fwd %.2e adj %.2e' % (p_, q_, dtype.__name__, e, eb))
    assert e < tol and eb < tol


def test_matrix_filter_capi_vs_element_filters():
    """hcipy_filter_execute_matrix through the C-ABI (2 x 2, 1024 x 1024 on the 2048 x 2048
    grid, a batch of 2) against the sums of the scalar element filters of the same plans, and
    with conj against the scalar adjoints."""
    from hcipy_amd.fourier_filter import _FilterPlan
    g = ha.make_pupil_grid(1024, 2)
    internal = ha.make_fft_grid(g, 2, 1)
    rng = np.random.default_rng(12)
    tf = np.exp(2j * np.pi * rng.random((2, 2, internal.size)))
    filt = ha.FourierFilter(g, dev(tf), 2)
    filt._is_matrix(dev(np.zeros(1)).device)
    x = dev(rng.standard_normal((2, 2, g.size)) + 1j * rng.standard_normal((2, 2, g.size)))
    plans = [filt._element_plan(torch.complex128, x.device, i, j) for i in range(2) for j in range(2)]
    for conj in (0, 1):
        out = torch.empty_like(x)
        _FilterPlan.execute_matrix(plans, 2, 2, conj, x, out, 2, {})
        for i in range(2):
            acc = torch.zeros_like(x[0])
            for j in range(2):
                # element (i, j) forward, or for conj the adjoint of element (i, j)
                t = torch.empty_like(x[0])
                plans[i * 2 + j].execute(conj, x[j].contiguous(), t, 2)
                acc += t
            e = peak_rel_field(host(out[i]), host(acc))
            say('matrix C-ABI conj=%d output %d: %.2e' % (conj, i, e))
            assert e < 1e-12


@pytest.mark.parametrize('dtype,tol', [(np.complex128, 1e-12), (np.complex64, 2e-6)])
def test_matrix_filter_elements_non_pow2(dtype, tol):
    """FourierFilter.apply_elements on a 34^2 padded grid (not a power of two): one forward
    transform, the element products, one batched backward transform; each element filter against
    the oracle, both directions, with a batch."""
    g = ha.make_pupil_grid(17, 2)
    internal = o.make_fft_grid(og(g), 2)
    rng = np.random.default_rng(14)
    tf = rng.standard_normal((2, 2, internal.size)) + 1j * rng.standard_normal((2, 2, internal.size))
    f = (rng.standard_normal((3, g.size)) + 1j * rng.standard_normal((3, g.size))).astype(dtype)
    filt = ha.FourierFilter(g, dev(tf), 2)
    for adjoint in (False, True):
        out = host(filt.apply_elements(ha.Field(dev(f), g), adjoint=adjoint))
        assert out.shape == (2, 2, 3, g.size)
        for i in range(2):
            for j in range(2):
                ref = o.fourier_filter(f.astype(np.complex128), og(g), tf[i, j], q=2, adjoint=adjoint)
                assert peak_rel_field(out[i, j], ref) < tol, (i, j, adjoint)


@pytest.mark.parametrize('dtype,tol', [(np.complex128, 1e-12), (np.complex64, 2e-6)])
def test_matrix_filter_elements_shared_passes(dtype, tol):
    """FourierFilter.apply_elements (hcipy_filter_execute_multi: one row pass and one forward
    column FFT shared by every element filter) against each element filter of the oracle, both
    directions, with a batch; more than 8 elements run as several shared passes."""
    g = ha.make_pupil_grid(32, 2)
    internal = o.make_fft_grid(og(g), 2)
    rng = np.random.default_rng(9)
    for p_, q_ in ((2, 2), (3, 3)):
        tf = rng.standard_normal((p_, q_, internal.size)) + 1j * rng.standard_normal((p_, q_, internal.size))
        f = (rng.standard_normal((2, g.size)) + 1j * rng.standard_normal((2, g.size))).astype(dtype)
        filt = ha.FourierFilter(g, dev(tf), 2)
        for adjoint in (False, True):
            out = host(filt.apply_elements(ha.Field(dev(f), g), adjoint=adjoint))
            assert out.shape == (p_, q_, 2, g.size)
            for i in range(p_):
                for j in range(q_):
                    ref = o.fourier_filter(f.astype(np.complex128), og(g), tf[i, j], q=2, adjoint=adjoint)
                    assert peak_rel_field(out[i, j], ref) < tol, (p_, q_, i, j, adjoint)


@pytest.mark.parametrize('dtype,tol', [(np.complex128, 1e-12), (np.complex64, 2e-6)])
@pytest.mark.parametrize('dims,q', [([17, 17], 2), ([24, 16], 3)])
def test_matrix_transfer_function_non_pow2(dims, q, dtype, tol):
    """A (2 x 3) matrix transfer function on padded grids that are not powers of two (34^2,
    72 x 48): FastFourierTransform per component, the per-frequency field_dot, the backward
    transform; forward and adjoint against the oracle, with a batch dimension."""
    g = ha.make_uniform_grid(dims, [1.0, dims[1] / dims[0]])
    internal = o.make_fft_grid(og(g), q)
    rng = np.random.default_rng(13)
    tf = rng.standard_normal((2, 3, internal.size)) + 1j * rng.standard_normal((2, 3, internal.size))
    filt = ha.FourierFilter(g, dev(tf), q)
    f = (rng.standard_normal((3, 2, g.size)) + 1j * rng.standard_normal((3, 2, g.size))).astype(dtype)
    out = host(filt.forward(ha.Field(dev(f), g)))
    ref = o.fourier_filter(f.astype(np.complex128), og(g), tf, q)
    assert out.shape == (2, 2, g.size)
    u = (rng.standard_normal((2, 2, g.size)) + 1j * rng.standard_normal((2, 2, g.size))).astype(dtype)
    back = host(filt.backward(ha.Field(dev(u), g)))
    refb = o.fourier_filter(u.astype(np.complex128), og(g), tf, q, adjoint=True)
    e, eb = peak_rel_field(out, ref), peak_rel_field(back, refb)
    say('matrix filter non-pow2 %s q=%d %s: fwd %.2e adj %.2e' % (dims, q, dtype.__name__, e, eb))
    assert e < tol and eb < tol


@pytest.mark.parametrize('dtype,tol', [(np.complex128, 1e-12), (np.complex64, 2e-6)])
@pytest.mark.parametrize('dims', [[4096, 8], [8, 4096]])
def test_filter_8192_point_lines(dims, dtype, tol):
    """8192-point row (4096 x 8 input, q = 2: 8192 x 16) and column (8 x 4096: 16 x 8192) passes of
    the filter kernels: a random scalar transfer function and a 2 x 2 matrix one against the
    oracle, forward and adjoint."""
    g = ha.make_uniform_grid(dims, [1.0, dims[1] / dims[0]])
    internal = o.make_fft_grid(og(g), 2)
    rng = np.random.default_rng(15)
    f = (rng.standard_normal((2, g.size)) + 1j * rng.standard_normal((2, g.size))).astype(dtype)
    tf = rng.standard_normal(internal.size) + 1j * rng.standard_normal(internal.size)
    filt = ha.FourierFilter(g, lambda gg: ha.Field(dev(tf), gg), 2)
    for adjoint in (False, True):
        out = host(filt.backward(ha.Field(dev(f), g)) if adjoint else filt.forward(ha.Field(dev(f), g)))
        ref = o.fourier_filter(f.astype(np.complex128), og(g), tf, q=2, adjoint=adjoint)
        assert peak_rel_field(out, ref) < tol, adjoint
    tfm = rng.standard_normal((2, 2, internal.size)) + 1j * rng.standard_normal((2, 2, internal.size))
    fm = ha.FourierFilter(g, dev(tfm), 2)
    out = host(fm.forward(ha.Field(dev(f), g)))
    ref = o.fourier_filter(f.astype(np.complex128), og(g), tfm, q=2)
    e = peak_rel_field(out, ref)
    say('8192-point filter %s %s: matrix %.2e' % (dims, dtype.__name__, e))
    assert e < tol


@pytest.mark.parametrize('dtype,tol', [(np.complex128, 1e-12), (np.complex64, 2e-6)])
@pytest.mark.parametrize('shape', [(3, 3), (4, 4), (1, 3)])
def test_filter_8192_point_matrix_shapes(shape, dtype, tol):
    """Matrix transfer functions of 3 and 4 inputs on 8192-point columns (8 x 4096 input, q = 2:
    16 x 8192), forward and adjoint, against the oracle: the column kernels that hold 3 or 4
    transforms per thread there run with 8 points per thread and spill registers (slower, still
    exact)."""
    p, q = shape
    g = ha.make_uniform_grid([8, 4096], [1.0, 4096 / 8])
    internal = o.make_fft_grid(og(g), 2)
    rng = np.random.default_rng(16)
    f = (rng.standard_normal((q, g.size)) + 1j * rng.standard_normal((q, g.size))).astype(dtype)
    tfm = rng.standard_normal((p, q, internal.size)) + 1j * rng.standard_normal((p, q, internal.size))
    fm = ha.FourierFilter(g, dev(tfm), 2)
    out = host(fm.forward(ha.Field(dev(f), g)))
    ref = o.fourier_filter(f.astype(np.complex128), og(g), tfm, q=2)
    assert out.shape == (p, g.size)
    u = (rng.standard_normal((p, g.size)) + 1j * rng.standard_normal((p, g.size))).astype(dtype)
    back = host(fm.backward(ha.Field(dev(u), g)))
    refb = o.fourier_filter(u.astype(np.complex128), og(g), tfm, q=2, adjoint=True)
    e, eb = peak_rel_field(out, ref), peak_rel_field(back, refb)
    say('8192-point matrix filter %dx%d %s: fwd %.2e adj %.2e' % (p, q, dtype.__name__, e, eb))
    assert e < tol and eb < tol
