"""Host-side logic of the package that needs no GPU."""
import numpy as np
import pytest
from scipy.signal import windows

from hcipy_amd.coronagraphy import periodic_tukey_window


def test_periodic_tukey_window_matches_scipy():
    for n in (4, 8, 31, 32, 64):
        np.testing.assert_allclose(periodic_tukey_window(n), windows.tukey(n, 1, False), rtol=0, atol=1e-15)


def test_zoom_fft_support_and_complexity():
    """check_if_supported messages (zoom_fast_fourier_transform.py:118-141) and the planner's
    estimate at 1024^2 -> 512^2 (the reference's coefficients give ~21 ms, the MFT's CPU model
    ~20 ms)."""
    import hcipy_amd as ha
    pupil = ha.make_pupil_grid(1024)
    focal = ha.make_focal_grid(4, 64)
    assert tuple(focal.dims) == (512, 512)
    c = ha.ZoomFastFourierTransform.compute_complexity(pupil, focal)
    assert 0.019 < c.expected_execution_time < 0.023
    with pytest.raises(ValueError, match='same dimensions'):
        ha.ZoomFastFourierTransform(pupil, ha.make_uniform_grid([8], [1.0]))
    with pytest.raises(ValueError, match='8192'):
        ha.ZoomFastFourierTransform(ha.make_uniform_grid([5000], [1.0]), ha.make_uniform_grid([4000], [1.0]))


def test_planner_estimates_and_choice():
    """make_fourier_transform(planner='estimate') (fourier_transform.py:217-300): at 1024^2 ->
    512^2 (q = 4, 64 lambda/D) the reference's time models predict 78.5 ms (FFT), 20.1 ms (MFT)
    and 21.0 ms (zoom FFT), so the MFT is chosen; a 4096^2 -> 4000^2 non-FFT output grid goes to
    the zoom FFT; an FFT grid given by q goes to the FFT."""
    import hcipy_amd as ha
    p = ha.make_pupil_grid(1024)
    uv = ha.FraunhoferPropagator(p, ha.make_focal_grid(4, 64)).instance(1.0).uv_grid
    est = {k: c.compute_complexity(p, uv).expected_execution_time * 1e3
           for k, c in (('fft', ha.FastFourierTransform), ('mft', ha.MatrixFourierTransform),
                        ('zft', ha.ZoomFastFourierTransform))}
    assert abs(est['fft'] - 78.5) < 0.1 and abs(est['mft'] - 20.1) < 0.1 and abs(est['zft'] - 21.0) < 0.1
    assert type(ha.make_fourier_transform(p, uv)) is ha.MatrixFourierTransform
    big = ha.make_pupil_grid(4096)
    assert type(ha.make_fourier_transform(big, ha.make_uniform_grid([4000, 4000], [300, 300]))) is ha.ZoomFastFourierTransform
    assert type(ha.make_fourier_transform(p, q=2)) is ha.FastFourierTransform
    with pytest.raises(ValueError):
        ha.make_fourier_transform(p, uv, planner='guess')


def test_planner_reference_choices():
    """tests/test_fourier.py:146-182: an FFT grid (q = 1, fov = 1, shift 0.1) goes to the FFT,
    q = 8 / fov = 0.3 to the MFT, a 1024^2 pupil at q = 8 / fov = 0.1 to the zoom FFT; an
    unstructured output grid has no device transform here (ValueError)."""
    import hcipy_amd as ha
    g = ha.make_pupil_grid(128)
    ft = ha.make_fourier_transform(g, q=1, fov=1, shift=0.1)
    assert type(ft) is ha.FastFourierTransform and ft.input_grid is g
    assert type(ha.make_fourier_transform(g, ha.make_fft_grid(g, q=1, fov=1, shift=0.1))) is ha.FastFourierTransform
    assert type(ha.make_fourier_transform(g, q=8, fov=0.3, shift=0.1)) is ha.MatrixFourierTransform
    assert type(ha.make_fourier_transform(g, ha.make_fft_grid(g, q=8, fov=0.3, shift=0.1))) is ha.MatrixFourierTransform
    big = ha.make_pupil_grid(1024)
    assert type(ha.make_fourier_transform(big, ha.make_fft_grid(big, q=8, fov=0.1))) is ha.ZoomFastFourierTransform
    from hcipy_amd.field import CartesianGrid, UnstructuredCoords
    rng = np.random.default_rng(0)
    un = CartesianGrid(UnstructuredCoords([rng.standard_normal(100), rng.standard_normal(100)]))
    with pytest.raises(ValueError):
        ha.make_fourier_transform(big, un)


def test_fft_backend_choice():
    """FastFourierTransform evaluates its sum with the direct FFT where every internal length is
    a power of two in [4, 8192], with the zoom FFT where each axis fits the 8192-point line FFT,
    with the MFT beyond that or when a split-precision MFT is asked for."""
    import hcipy_amd as ha
    g = ha.make_pupil_grid(256)
    assert ha.FastFourierTransform(g, q=2).method == 'fft'
    assert ha.FastFourierTransform(g, q=2, fov=0.5, shift=0.1, precision='fp32').method == 'fft'
    assert ha.FastFourierTransform(ha.make_uniform_grid([4096], [1.0]), q=2).method == 'fft'
    assert ha.FastFourierTransform(ha.make_pupil_grid(200), q=2).method == 'zft'
    assert ha.FastFourierTransform(ha.make_pupil_grid(200), q=2, precision='fp32').method == 'zft'
    assert ha.FastFourierTransform(g, precision='f16x2').method == 'mft'
    assert ha.FastFourierTransform(g, precision='bf16x3').method == 'mft'
    assert ha.FastFourierTransform(ha.make_uniform_grid([5000], [1.0])).method == 'mft'
    assert ha.FastFourierTransform(ha.make_uniform_grid([4096], [1.0]), q=4).method == 'mft'
