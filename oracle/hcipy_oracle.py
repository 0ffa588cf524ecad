"""FP64 numpy/scipy restatement of HCIPy's Fourier-propagation path (the oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``; never by the product package.

Each function cites the reference file:line it restates.  How each part is pinned, so that a
dropped term, a sign error, a swapped index or a transposed operand fails a check
(``tests/test_oracle.py``):

* grids (``make_uniform_grid``/``make_pupil_grid``/``make_focal_grid``/``make_fft_grid``/
  ``get_fft_parameters``): the reference's own grid-reconstruction round trip
  (tests/test_fourier.py:185-229) and hand-computed coordinates;
* ``mft_forward``/``mft_backward``: brute-force dense DFT ``naive_ft`` (the reference's
  NaiveFourierTransform, hcipy/fourier/naive_fourier_transform.py:82-124) on small grids,
  Parseval/round trip at fov=1 (tests/test_fourier.py:51-59), even/odd symmetry
  (tests/test_fourier.py:89-124), the weighted-adjoint identity, and numpy's FFT on FFT grids;
* ``fraunhofer_forward``: closed-form Airy 2 J1(x)/x and sinc*sinc fields at the reference's
  own tolerances (tests/test_propagation.py:6-62);
* ``fft_forward``/``fft_backward``/``fourier_filter``: numpy.fft as library routine, the
  MFT on the same grids, and FourierFilter == FFT.backward(FFT.forward(f) * tf)
  (tests/test_fourier.py:273-309);
* ``angular_spectrum_tf``/``fresnel_tf``/``impulse_response_tf``: the Fresnel-integral closed
  form of a rectangular aperture and +-z symmetry (tests/test_propagation.py:64-138), at 0.2 of
  the critical distance (transfer-function branch) and at 5 (impulse-response branch), as the
  reference's own test does;
* apertures: the reference's FITS golden rasters (tests/baseline_for_apertures/circular,
  obstructed_circular);
* ``lyot_forward``: the explicit (1-M) Babinet algebra and the reference thresholds
  (tests/test_coronagraphy.py:272-304);
* ``multiscale_masks``/``multiscale_forward``/``vortex_mask``/``vector_vortex_mask``: the
  reference's rejection thresholds for charges 2-8, scalar and vector
  (tests/test_coronagraphy.py:4-28, 81-105), the unit-mask identity, the adjoint identity, and
  the circular-basis identity tying the vector vortex to the scalar one;
* ``czt``/``zoom_fft``: the dense DFT (``naive_ft``) on regular grids with arbitrary spacings,
  offsets, odd sizes and output longer than input, in 1-D and 2-D, forward and backward; the
  chirp-z transform against a direct sum over its z_k (arbitrary complex a and w) and against
  numpy's FFT when it is one (a = 1, w = exp(-2 pi i / n), m = n);
* ``spectral_noise_fft``/``power_law_error``: the ensemble variance of the screens equals the
  integral of the PSD over the FFT grid (sum psd w / (2 pi)^2: the (2 pi)^ndim and weight
  factors), a one-pixel shift is np.roll of the screen by one pixel along that axis, two half
  shifts compose, and the reference's peak-to-valley thresholds (tests/test_optics.py:823-838);
* ``spectral_noise_multiscale``: the ensemble variance equals the PSD integral over g1 outside
  the boundary plus over g2 inside it (the masks, their sides and both amplitude factors); a
  one-pixel shift moves the screen by one pixel (the periodic FFT part and the non-periodic
  MFT part alike, so only the overlapping samples are compared);
* ``xinetics_influence_functions``: at zero tilt, on a pupil grid whose samples fall on the
  table's samples, each influence function equals the table entries (the shift, axis order
  and normalisation); the reference's tilted-centroid test (tests/test_optics.py:119-152);
* ``stokes_vector``: for pure input states the Mueller rows must agree with the Stokes
  parameters of the propagated Jones vector J e (an independent closed form), and those four
  states span every input Stokes vector; Malus's law for a rotated polarizer;
* ``field_dot``/``phase_retarder_jones``: the half- and quarter-wave plate matrices and the
  circular-basis identity of the vector vortex;
* ``pyramid_wfs_forward``/``modulated_pyramid_wfs_forward``: a flat pupil makes four equal
  pupil images centred at (+-sep/2, +-sep/2) (the surface's scale); a pupil tilted into the
  (+, +) facet lands in the (-, -) quadrant (the sign chain of two forward MFTs and the
  surface); backward is the weighted adjoint of forward; the reference's slow/fast modulation
  agreement (tests/test_wavefront_sensing.py:29-52);
* ``gaussian_influence_functions``/``dm_surface``: crosstalk definition (value at the
  neighbouring actuator) and cutoff (hcipy/optics/deformable_mirror.py:78-79).
"""
import math

import numpy as np
from scipy.linalg import blas

# ----------------------------------------------------------------------------------------------
# Grids (hcipy/field/coordinates.py:533-675, hcipy/field/cartesian_grid.py:83-104,167-189)


class Grid:
    """Cartesian grid with separated coordinates; regular when delta/zero are set.

    Layout follows HCIPy: ``dims`` = (nx, ny), ``shape`` = (ny, nx), x fastest
    (hcipy/field/coordinates.py:671-675).  ``weights`` is a scalar for regular grids
    (hcipy/field/cartesian_grid.py:171-172) and is multiplied by |s|^ndim on scaling
    (hcipy/field/cartesian_grid.py:83-104).
    """

    def __init__(self, delta=None, dims=None, zero=None, separated=None, weights=None):
        if separated is not None:
            self.sep = [np.asarray(s, dtype=float) for s in separated]
            self.delta = None
            self.zero = None
            self.dims = tuple(len(s) for s in self.sep)
        else:
            self.delta = np.asarray(delta, dtype=float)
            self.dims = tuple(int(d) for d in dims)
            self.zero = np.asarray(zero, dtype=float)
            # RegularCoords.separated_coords: arange(n) * delta + zero (coordinates.py:646-652)
            self.sep = [np.arange(n, dtype=float) * d + z for d, n, z in zip(self.delta, self.dims, self.zero)]
        self._weights = weights

    @property
    def is_regular(self):
        return self.delta is not None

    @property
    def ndim(self):
        return len(self.dims)

    @property
    def shape(self):
        return tuple(reversed(self.dims))

    @property
    def size(self):
        return int(np.prod(self.dims))

    @property
    def weights(self):
        if self._weights is not None:
            return self._weights
        if self.is_regular:
            return float(np.prod(self.delta))
        # cartesian_grid.py:174-189: midpoint spacing per axis, outer product
        w = None
        for i, x in enumerate(self.sep):
            wi = np.concatenate(([x[1] - x[0]], (x[2:] - x[:-2]) / 2, [x[-1] - x[-2]]))
            shape = (1,) * (self.ndim - i - 1) + (-1,) + (1,) * i
            wi = wi.reshape(shape)
            w = wi if w is None else w * wi
        return w.ravel()

    def scaled(self, s):
        """Grid.scaled: coords *= s, weights *= |s|^ndim (cartesian_grid.py:83-104)."""
        w = self.weights
        if np.isscalar(s):
            w = w * abs(float(s)) ** self.ndim
        else:
            w = w * np.prod(np.abs(s))
        s = np.ones(self.ndim) * s
        if self.is_regular:
            return Grid(self.delta * s, self.dims, self.zero * s, weights=w)
        return Grid(separated=[x * si for x, si in zip(self.sep, s)], weights=w)

    def shifted(self, shift):
        shift = np.ones(self.ndim) * shift
        if self.is_regular:
            return Grid(self.delta, self.dims, self.zero + shift, weights=self._weights)
        return Grid(separated=[x + si for x, si in zip(self.sep, shift)], weights=self._weights)

    @property
    def coords(self):
        if self.ndim == 1:
            return [self.sep[0]]
        g = np.meshgrid(*self.sep)
        return [c.ravel() for c in g]

    @property
    def x(self):
        return self.coords[0]

    @property
    def y(self):
        return self.coords[1]


def make_uniform_grid(dims, extent, center=0, has_center=False):
    """hcipy/field/util.py:68-125."""
    dims = np.atleast_1d(dims)
    extent = np.atleast_1d(np.asarray(extent, dtype=float))
    n = max(len(dims), len(extent), np.size(center))
    dims = np.ones(n, dtype=int) * dims
    extent = np.ones(n) * extent
    center = np.ones(n) * center
    delta = extent / dims
    zero = -extent / 2 + center + delta / 2
    if has_center:
        zero -= delta / 2 * (1 - dims % 2)
    return Grid(delta, tuple(int(d) for d in dims), zero)


def make_pupil_grid(dims, diameter=1):
    """hcipy/field/util.py:127-163 (always 2-D)."""
    return make_uniform_grid(np.ones(2, dtype=int) * dims, np.ones(2) * diameter)


def make_focal_grid(q, num_airy, spatial_resolution=1.0):
    """hcipy/field/util.py:222-318 (normalised or given spatial resolution)."""
    q = np.ones(2) * q
    num_airy = np.ones(2) * num_airy
    sr = np.ones(2) * spatial_resolution
    delta = sr / q
    dims = tuple(int(2 * n * qi) for n, qi in zip(num_airy, q))
    dims_arr = np.asarray(dims, dtype=float)
    zero = delta * (-dims_arr / 2 + (dims_arr % 2) * 0.5)
    return Grid(delta, dims, zero)


def make_fft_grid(input_grid, q=1, fov=1, shift=0):
    """hcipy/fourier/fast_fourier_transform.py:23-62."""
    nd = input_grid.ndim
    q = tuple(np.ones(nd) * q)
    fov = tuple(np.ones(nd) * fov)
    shift = tuple(np.ones(nd) * shift)
    q = tuple(round(q_i * d_in) / d_in for q_i, d_in in zip(q, input_grid.dims))
    delta = tuple(float((2 * math.pi / (d_in * s_in)) / q_i) for d_in, s_in, q_i in zip(input_grid.delta, input_grid.dims, q))
    dims = tuple(int(d_in * f_i * q_i) for d_in, f_i, q_i in zip(input_grid.dims, fov, q))
    zero = tuple(float(d_i * (-dim / 2 + (dim % 2) * 0.5) + s_i) for d_i, dim, s_i in zip(delta, dims, shift))
    return Grid(delta, dims, zero)


# ----------------------------------------------------------------------------------------------
# Matrix Fourier transform (hcipy/fourier/matrix_fourier_transform.py)


def mft_matrices(input_grid, output_grid, complex_dtype='complex128'):
    """_compute_matrices (matrix_fourier_transform.py:67-111): FP64 phases, exp, cast."""
    if input_grid.ndim == 1:
        xu = np.outer(output_grid.sep[0], input_grid.sep[0])
        return (np.exp(-1j * xu).astype(complex_dtype),)
    x, y = input_grid.sep
    u, v = output_grid.sep
    M1 = np.exp(-1j * np.outer(v, y)).astype(complex_dtype)
    M2 = np.exp(-1j * np.outer(x, u)).astype(complex_dtype)
    return M1, M2


def _weights(grid, div=1.0):
    w = grid.weights / div
    if not np.isscalar(w) and np.all(w == w[0]):
        w = w[0]
    return w


def mft_forward(field, input_grid, output_grid):
    """MatrixFourierTransform.forward (matrix_fourier_transform.py:132-176).

    ``field`` is (..., Ny*Nx); complex64 input is computed with cgemm on complex64 matrices,
    anything else with zgemm (fourier_transform.py:323-343).
    """
    field = np.asarray(field)
    cdt = 'complex64' if field.dtype in (np.float32, np.complex64) else 'complex128'
    lead = field.shape[:-1]
    f_all = field.reshape((-1, field.shape[-1])).astype(cdt)
    out = []
    mats = mft_matrices(input_grid, output_grid, cdt)
    w_in = _weights(input_grid)
    for f in f_all:
        if input_grid.ndim == 1:
            out.append(mats[0].dot(f * w_in))
            continue
        M1, M2 = mats
        gemm = blas.cgemm if cdt == 'complex64' else blas.zgemm
        if np.isscalar(w_in):
            fs = f.reshape(input_grid.shape)
            alpha = w_in
        else:
            fs = (f * w_in).reshape(input_grid.shape)
            alpha = 1
        inter = gemm(1, M2.T, fs.T).T  # (Ny, Nu) = f . M2          (matrix_fourier_transform.py:171)
        res = gemm(alpha, inter.T, M1.T).T  # (Nv, Nu) = alpha M1 . inter  (:172)
        out.append(res.reshape(-1))
    return np.asarray(out).reshape(lead + (output_grid.size,))


def mft_backward(field, input_grid, output_grid):
    """MatrixFourierTransform.backward (matrix_fourier_transform.py:178-224)."""
    field = np.asarray(field)
    cdt = 'complex64' if field.dtype in (np.float32, np.complex64) else 'complex128'
    lead = field.shape[:-1]
    f_all = field.reshape((-1, field.shape[-1])).astype(cdt)
    mats = mft_matrices(input_grid, output_grid, cdt)
    w_out = _weights(output_grid, (2 * np.pi) ** input_grid.ndim)
    out = []
    for f in f_all:
        if input_grid.ndim == 1:
            out.append(mats[0].conj().T.dot(f * w_out))
            continue
        M1, M2 = mats
        gemm = blas.cgemm if cdt == 'complex64' else blas.zgemm
        if np.isscalar(w_out):
            fs = f.reshape(output_grid.shape)
            alpha = w_out
        else:
            fs = (f * w_out).reshape(output_grid.shape)
            alpha = 1
        inter = gemm(1, fs.T, M1.T, trans_b=2).T  # (Ny, Nu) = M1^H . f   (:219)
        res = gemm(alpha, M2.T, inter.T, trans_a=2).T  # (Ny, Nx) = inter . M2^H  (:220)
        out.append(res.reshape(-1))
    return np.asarray(out).reshape(lead + (input_grid.size,))


def czt(x, n, m, w, a):
    """ChirpZTransform.__call__ (hcipy/fourier/chirp_z_transform.py:49-105): X_k = sum_j x_j z_k^-j,
    z_k = a w^-k, along the last axis by Bluestein's algorithm with a numpy FFT of length
    next_fast_len(n + m - 1); the chirp filter holds w^-(d^2/2) for d = -(n-1) .. m-1 in that
    order and the result is read from offset n - 1."""
    from scipy.fft import next_fast_len
    w = complex(w)
    a = complex(a)
    k = np.arange(max(m, n))
    wk2 = w ** (k ** 2 / 2)
    nfft = next_fast_len(n + m - 1)
    Awk2 = a ** -k[:n] * wk2[:n]
    Fwk2 = np.fft.fft(1 / np.hstack((wk2[n - 1:0:-1], wk2[:m])), n=nfft)
    y = np.fft.ifft(np.fft.fft(np.asarray(x) * Awk2, n=nfft) * Fwk2)
    return y[..., n - 1:n + m - 1] * wk2[:m]


def zoom_fft(field, input_grid, output_grid, adjoint=False):
    """ZoomFastFourierTransform.forward / backward (hcipy/fourier/zoom_fast_fourier_transform.py:
    42-116): a chirp-z transform along x (the last axis of the shaped field) with
    w = exp(-i du dx), a = exp(i u0 dx) and the shift exp(-i u x0), then along y; backward
    swaps the grids, flips the signs and weights with output_grid.weights / (2 pi)^ndim."""
    f = np.asarray(field)
    if adjoint:
        input_grid, output_grid = output_grid, input_grid
        sgn = 1
        weight = np.asarray(input_grid.weights).ravel()[0] / (2 * np.pi) ** input_grid.ndim
    else:
        sgn = -1
        weight = np.asarray(input_grid.weights).ravel()[0]
    lead = f.shape[:-1]
    g = (f * weight).reshape(lead + tuple(input_grid.shape))
    nd = input_grid.ndim
    for i in range(nd):
        n, m = int(input_grid.dims[i]), int(output_grid.dims[i])
        dx, du = input_grid.delta[i], output_grid.delta[i]
        x0, u0 = input_grid.zero[i], output_grid.zero[i]
        w = np.exp(sgn * 1j * du * dx)
        a = np.exp(-sgn * 1j * u0 * dx)
        shift = np.exp(sgn * 1j * output_grid.sep[i] * x0)
        axis = g.ndim - 1 - i  # x is the last axis
        g = np.moveaxis(czt(np.moveaxis(g, axis, -1), n, m, w, a) * shift, -1, axis)
    return g.reshape(lead + (-1,))


def naive_ft(field, input_grid, output_grid, adjoint=False):
    """Dense DFT on arbitrary grids (NaiveFourierTransform, naive_fourier_transform.py:82-124)."""
    cin = np.array(input_grid.coords)
    cout = np.array(output_grid.coords)
    if not adjoint:
        A = np.exp(-1j * np.dot(cout.T, cin)) * input_grid.weights
    else:
        A = np.exp(1j * np.dot(cin.T, cout)) * output_grid.weights / (2 * np.pi) ** input_grid.ndim
    return A.dot(np.asarray(field, dtype=complex))


def mft_complexity_ops(input_grid, output_grid):
    """compute_complexity (matrix_fourier_transform.py:256-309): real mults + adds."""
    if input_grid.ndim == 1:
        cm = input_grid.size * output_grid.size
        ca = (input_grid.size - 1) * output_grid.size
    else:
        nox, noy = output_grid.shape
        nix, niy = input_grid.shape
        cm = niy * nix * noy
        ca = noy * nix * (niy - 1)
        cm += nix * nox * noy
        ca += nox * noy * (nix - 1)
        cm += input_grid.size
    return 4 * cm, 2 * cm + 2 * ca


def predict_time(num_operations, a, b, c):
    """FourierTransform._predict_execution_time (fourier_transform.py:159-183), seconds."""
    return (math.exp(a) * (num_operations / 1e9) ** b + math.exp(c)) * 1e-3


# ----------------------------------------------------------------------------------------------
# Fraunhofer propagation (hcipy/propagation/fraunhofer.py:32-88)


def fraunhofer_forward(field, pupil_grid, focal_grid, wavelength=1.0, focal_length=1.0):
    uv = focal_grid.scaled(2 * np.pi / (focal_length * wavelength))
    norm = 1 / (1j * focal_length * wavelength)
    return mft_forward(field, pupil_grid, uv) * norm


def fraunhofer_backward(field, pupil_grid, focal_grid, wavelength=1.0, focal_length=1.0):
    uv = focal_grid.scaled(2 * np.pi / (focal_length * wavelength))
    norm = 1 / (1j * focal_length * wavelength)
    return mft_backward(field, pupil_grid, uv) / norm


# ----------------------------------------------------------------------------------------------
# FFT and FourierFilter (hcipy/fourier/fast_fourier_transform.py:188-406,
# hcipy/fourier/fourier_operations.py:8-151)


def _grid_shift(shift, grid):
    c = grid.coords
    return np.exp(1j * sum(s * ci for s, ci in zip(shift, c)))


class FFT:
    """FastFourierTransform with emulate_fftshifts as configured (default True)."""

    def __init__(self, input_grid, q=1, fov=1, shift=0, emulate_fftshifts=True):
        self.input_grid = input_grid
        self.emulate = emulate_fftshifts
        self.shape_in = input_grid.shape
        self.weights = input_grid.weights
        self.output_grid = make_fft_grid(input_grid, q, fov, shift)
        self.internal_grid = make_fft_grid(input_grid, q, 1)
        self.shape_out = self.output_grid.shape
        self.internal_shape = self.internal_grid.shape
        if self.internal_shape == self.shape_in:
            self.cutout_input = None
        else:
            st = tuple(int(i / 2) - int(n / 2) for i, n in zip(self.internal_shape, self.shape_in))
            self.cutout_input = tuple(slice(s, s + n) for s, n in zip(st, self.shape_in))
        if self.internal_shape == self.shape_out:
            self.cutout_output = None
        else:
            st = tuple(int(i / 2) - int(n / 2) for i, n in zip(self.internal_shape, self.shape_out))
            self.cutout_output = tuple(slice(s, s + n) for s, n in zip(st, self.shape_out))
        center = input_grid.zero + input_grid.delta * (np.asarray(input_grid.dims) // 2)
        self.shift_input = _grid_shift(-center, self.output_grid)
        self.shift_input /= np.fft.ifftshift(self.shift_input.reshape(self.shape_out)).ravel()[0]
        if self.emulate:
            f_shift = input_grid.delta * (np.asarray(self.internal_shape[::-1]) // 2)
            fftshift = _grid_shift(f_shift, self.internal_grid)
            if self.cutout_output:
                self.shift_input *= fftshift.reshape(self.internal_shape)[self.cutout_output].ravel()
            else:
                self.shift_input *= fftshift
        self.shift_input *= self.weights
        shift = np.ones(input_grid.ndim) * shift
        if np.allclose(shift, 0):
            self.shift_output = 1
        else:
            self.shift_output = _grid_shift(-shift, input_grid)
        if self.emulate:
            f_shift = input_grid.delta * (np.asarray(self.internal_shape[::-1]) // 2)
            fftshift = _grid_shift(f_shift, self.internal_grid)
            fftshift *= np.exp(-1j * np.dot(f_shift, self.internal_grid.zero))
            if self.cutout_input:
                self.shift_output = self.shift_output * fftshift.reshape(self.internal_shape)[self.cutout_input].ravel()
            else:
                self.shift_output = self.shift_output * fftshift
        if np.isscalar(self.shift_output) and np.allclose(self.shift_output, 1):
            self.shift_output = None

    def forward(self, field):
        """fast_fourier_transform.py:325-360."""
        arr = np.zeros(self.internal_shape, dtype=complex)
        f = np.asarray(field, dtype=complex).reshape(self.shape_in)
        if self.shift_output is not None:
            f = f * self.shift_output.reshape(self.shape_in)
        if self.cutout_input is None:
            arr[:] = f
        else:
            arr[self.cutout_input] = f
        if not self.emulate:
            arr = np.fft.ifftshift(arr)
        r = np.fft.fftn(arr)
        if not self.emulate:
            r = np.fft.fftshift(r)
        res = r.ravel() if self.cutout_output is None else r[self.cutout_output].ravel()
        return res * self.shift_input

    def backward(self, field):
        """fast_fourier_transform.py:362-404."""
        arr = np.zeros(self.internal_shape, dtype=complex)
        f = np.asarray(field, dtype=complex).reshape(self.shape_out) / self.shift_input.reshape(self.shape_out)
        if self.cutout_output is None:
            arr[:] = f
        else:
            arr[self.cutout_output] = f
        if not self.emulate:
            arr = np.fft.ifftshift(arr)
        r = np.fft.ifftn(arr)
        if not self.emulate:
            r = np.fft.fftshift(r)
        res = r.ravel() if self.cutout_input is None else r[self.cutout_input].ravel()
        if self.shift_output is not None:
            res = res / self.shift_output
        return res


def fourier_filter(field, input_grid, tf_internal, q=1, adjoint=False):
    """FourierFilter._operation (fourier_operations.py:89-151).

    ``tf_internal`` is the transfer function sampled on the internal (FFT output, q, fov=1) grid
    in that grid's natural (centred) order; it is ifftshifted here as in :41-52.
    ``field`` may carry leading batch dimensions.  A transfer function of shape (p, q, size) is a
    matrix field (:120-131): it contracts with the field's first tensor dimension (q components)
    as field_dot does, and the result has p there (the adjoint uses the conjugate transpose).
    """
    fft = FFT(input_grid, q)
    internal_grid = fft.output_grid
    tf_arr = np.asarray(tf_internal)
    if tf_arr.ndim == 3:
        # matrix field: field_dot(tf, F) contracts tf's second index with the field's first
        # tensor index; the field's other tensor indices are carried along
        tf = np.fft.ifftshift(tf_arr.reshape(tf_arr.shape[:2] + internal_grid.shape), axes=(-2, -1))
        if adjoint:
            tf = np.conj(np.swapaxes(tf, 0, 1))
        field = np.asarray(field)
        f = field.reshape(field.shape[:-1] + input_grid.shape)
        arr = np.zeros(field.shape[:-1] + internal_grid.shape, dtype=np.result_type(field.dtype, np.complex64))
        cut = (slice(None),) * (field.ndim - 1) + (fft.cutout_input or (slice(None), slice(None)))
        arr[cut] = f
        F = np.fft.fftn(arr, axes=(-2, -1))
        G = np.einsum('ijyx,j...yx->i...yx', tf.astype(arr.dtype), F)  # field_dot per frequency
        r = np.fft.ifftn(G, axes=(-2, -1))[cut]
        return r.reshape(r.shape[:-2] + (-1,)).astype(arr.dtype)
    tf = np.fft.ifftshift(tf_arr.reshape(internal_grid.shape))
    field = np.asarray(field)
    lead = field.shape[:-1]
    f = field.reshape(lead + input_grid.shape)
    if fft.cutout_input is not None:
        arr = np.zeros(lead + internal_grid.shape, dtype=np.result_type(field.dtype, np.complex64))
        c = (slice(None),) * len(lead) + fft.cutout_input
        arr[c] = f
    else:
        arr = f.astype(np.result_type(field.dtype, np.complex64))
        c = None
    axes = tuple(range(-input_grid.ndim, 0))
    tf = tf.astype(arr.dtype)
    F = np.fft.fftn(arr, axes=axes)
    F = F * (tf.conj() if adjoint else tf)
    r = np.fft.ifftn(F, axes=axes)
    if c is not None:
        r = r[c]
    return r.reshape(lead + (-1,)).astype(arr.dtype)


def evaluate_supersampled(fn, grid, oversampling):
    """evaluate_supersampled for separated grids (hcipy/field/util.py:606-698): mean over
    the dithered sub-grids, dithers from make_uniform_grid(oversampling, 1)."""
    dith = make_uniform_grid(np.ones(grid.ndim, dtype=int) * int(round(oversampling)), 1)
    pts = np.array(dith.coords).T
    deltas = []
    for x in grid.sep:
        d = (x[2:] - x[:-2]) / 2.
        deltas.append(np.concatenate(([x[1] - x[0]], d, [x[-1] - x[-2]])))
    acc = 0
    for p in pts:
        g = Grid(separated=[c + float(p[i]) * dl for i, (c, dl) in enumerate(zip(grid.sep, deltas))])
        acc = acc + fn(g)
    return acc / len(pts)


def angular_spectrum_tf(internal_grid, wavelength, distance, num_oversampling=2, refractive_index=1):
    """Native transfer function branch (hcipy/propagation/angular_spectrum.py:71-78)."""
    k = 2 * np.pi / wavelength * refractive_index

    def tf_native(g):
        k_squared = np.hypot(g.x, g.y) ** 2
        k_z = np.sqrt(k ** 2 - k_squared + 0j)
        return np.exp(1j * k_z * distance)

    return evaluate_supersampled(tf_native, internal_grid, num_oversampling)


def angular_spectrum_uses_native_tf(input_grid, wavelength, distance):
    """Branch rule of angular_spectrum.py:53-55 (True: native transfer function)."""
    L_max = np.max(np.asarray(input_grid.dims) * input_grid.delta)
    return not np.any(input_grid.delta < wavelength * abs(distance) / L_max)


def fresnel_tf(internal_grid, wavelength, distance, num_oversampling=2, refractive_index=1):
    """Native transfer function branch (hcipy/propagation/fresnel.py:69-80)."""
    k = 2 * np.pi / wavelength * refractive_index
    phase_factor = np.exp(1j * k * distance)
    alpha = -0.5j * distance / k

    def tf_native(g):
        x = g.x
        y = g.y
        return phase_factor * np.exp(alpha * (x * x + y * y))

    return evaluate_supersampled(tf_native, internal_grid, num_oversampling)


def impulse_response_tf(kind, internal_grid, wavelength, distance, num_oversampling=2, refractive_index=1):
    """Impulse-response branch (hcipy/propagation/angular_spectrum.py:55-69,
    hcipy/propagation/fresnel.py:50-67): the impulse response supersampled on the FFT grid of the
    internal grid (the padded spatial grid), then FFT'd onto the internal grid."""
    k = 2 * np.pi / wavelength * refractive_index
    enlarged = make_fft_grid(internal_grid)
    fft = FFT(enlarged)

    if kind == 'angular_spectrum':
        def ir(g):
            r_squared = g.x ** 2 + g.y ** 2 + distance ** 2
            r = np.sqrt(r_squared)
            cos_theta = distance / r
            return cos_theta / (2 * np.pi) * np.exp(1j * k * r) * (1 / r_squared - 1j * k / r)
    else:
        phase_factor = np.exp(1j * k * distance) / (1j * wavelength * distance)

        def ir(g):
            return phase_factor * np.exp(0.5j * k / distance * (g.x ** 2 + g.y ** 2))

    return fft.forward(evaluate_supersampled(ir, enlarged, num_oversampling))


def angular_spectrum_forward(field, input_grid, wavelength, distance, num_oversampling=2, adjoint=False):
    """AngularSpectrumPropagator (angular_spectrum.py:45-80): native transfer function, or the
    impulse-response branch when the grid step is below lambda |z| / L."""
    internal = make_fft_grid(input_grid, 2, 1)
    if angular_spectrum_uses_native_tf(input_grid, wavelength, distance):
        tf = angular_spectrum_tf(internal, wavelength, distance, num_oversampling)
    else:
        tf = impulse_response_tf('angular_spectrum', internal, wavelength, distance, num_oversampling)
    return fourier_filter(field, input_grid, tf, q=2, adjoint=adjoint)


def fresnel_forward(field, input_grid, wavelength, distance, num_oversampling=2, zero_padding=2, adjoint=False):
    """FresnelPropagator (fresnel.py:43-82), with the same branch rule."""
    internal = make_fft_grid(input_grid, zero_padding, 1)
    if angular_spectrum_uses_native_tf(input_grid, wavelength, distance):
        tf = fresnel_tf(internal, wavelength, distance, num_oversampling)
    else:
        tf = impulse_response_tf('fresnel', internal, wavelength, distance, num_oversampling)
    return fourier_filter(field, input_grid, tf, q=zero_padding, adjoint=adjoint)


# ----------------------------------------------------------------------------------------------
# Apertures (hcipy/aperture/generic.py:11-46, 301-348, 394-429)


def circular_aperture(grid, diameter, center=(0, 0)):
    x = grid.sep[0][np.newaxis, :]
    y = grid.sep[1][:, np.newaxis]
    return (((x - center[0]) ** 2 + (y - center[1]) ** 2) <= (diameter / 2) ** 2).ravel().astype(float)


def rectangular_aperture(grid, size):
    size = np.ones(2) * size
    x = grid.sep[0][np.newaxis, :]
    y = grid.sep[1][:, np.newaxis]
    return ((np.abs(x) <= size[0] / 2) * (np.abs(y) <= size[1] / 2)).ravel().astype(float)


def spider(grid, p1, p2, width):
    delta = np.array(p2) - np.array(p1)
    shift = delta / 2 + np.array(p1)
    ang = np.arctan2(delta[1], delta[0])
    length = np.linalg.norm(delta)
    x = grid.sep[0][np.newaxis, :] - shift[0]
    y = grid.sep[1][:, np.newaxis] - shift[1]
    xn = x * np.cos(ang) + y * np.sin(ang)
    yn = y * np.cos(ang) - x * np.sin(ang)
    s = (xn <= length / 2) * (xn >= -length / 2) * (yn <= width / 2) * (yn >= -width / 2)
    return (1 - s.ravel()).astype(float)


def obstructed_circular_aperture(grid, pupil_diameter, obscuration_ratio, num_spiders=0, spider_width=0.01):
    outer = circular_aperture(grid, pupil_diameter)
    inner = circular_aperture(grid, pupil_diameter * obscuration_ratio)
    sp = 1
    for angle in np.linspace(0, 2 * np.pi, num_spiders, endpoint=False):
        sp = sp * spider(grid, (0, 0), (pupil_diameter * np.cos(angle), pupil_diameter * np.sin(angle)), spider_width)
    return (outer - inner) * sp


# ----------------------------------------------------------------------------------------------
# Optics: apodizers, Lyot coronagraph, DM, intensity (hcipy/optics/apodization.py:6-62,
# hcipy/coronagraphy/lyot.py:66-122, hcipy/optics/deformable_mirror.py:43-93,200-263,
# hcipy/optics/wavefront.py:113-143,310-360, hcipy/metrics/contrast.py:37-52)


def lyot_forward(E, pupil_grid, fpm_grid, fpm, lyot_stop=None, wavelength=1.0, focal_length=1.0):
    """lyot.py:66-92: F = P E; F -= M F; L = E - P^dagger F; L *= stop."""
    F = fraunhofer_forward(E, pupil_grid, fpm_grid, wavelength, focal_length)
    F = F - F * fpm
    L = fraunhofer_backward(F, pupil_grid, fpm_grid, wavelength, focal_length)
    L = E - L
    if lyot_stop is not None:
        L = L * lyot_stop
    return L


def lyot_backward(E, pupil_grid, fpm_grid, fpm, lyot_stop=None, wavelength=1.0, focal_length=1.0):
    """lyot.py:94-122 (forward propagator, conj(M), backward propagator, subtract)."""
    wf = E * np.conj(lyot_stop) if lyot_stop is not None else E
    F = fraunhofer_forward(wf, pupil_grid, fpm_grid, wavelength, focal_length)
    F = F - F * np.conj(fpm)
    P = fraunhofer_backward(F, pupil_grid, fpm_grid, wavelength, focal_length)
    return wf - P


def vortex_mask(charge):
    """VortexCoronagraph's focal mask (hcipy/coronagraphy/vortex.py:12-15): exp(i l theta)
    with the sample at the origin removed (times 1 - circ(1e-9))."""
    def fn(grid):
        x, y = grid.coords
        return np.exp(1j * charge * np.arctan2(y, x)) * (1 - circular_aperture(grid, 1e-9))
    return fn


def multiscale_window(window_size, shape):
    """Centred 2-D periodic Tukey(alpha = 1) window, padded as np.pad(w, ((n - w) // 2 per axis))
    does (hcipy/coronagraphy/multi_scale.py:73-78)."""
    from scipy.signal import windows
    w1 = windows.tukey(window_size, 1, False)
    w = np.outer(w1, w1)
    return np.pad(w, tuple((n_foc - n_w) // 2 for n_foc, n_w in zip(shape, w.shape)), 'constant').ravel()


def multiscale_masks(pupil_grid, complex_mask, q=1024, scaling_factor=4, window_size=32):
    """MultiScaleCoronagraph.__init__ (hcipy/coronagraphy/multi_scale.py:44-106): the focal grids
    of the levels and their corrected masks.  Level 0 is sampled at q = 2 over the full FFT
    field; level i at q = 2 s^i over a window of ``window_size`` pixels of level i - 1.  Each
    level's mask is windowed (all but the last) and has the band-limited content of every
    coarser level, resampled onto its grid (FFT then adjoint MFT), subtracted."""
    pupil_diameter = np.asarray(pupil_grid.dims) * pupil_grid.delta
    levels = int(math.ceil(math.log(q / 2) / math.log(scaling_factor))) + 1
    qs = [2 * scaling_factor ** i for i in range(levels)]
    num_airys = [tuple(n / 2 for n in pupil_grid.dims)]
    for i in range(1, levels):
        num_airys.append(tuple(n_airy * window_size / (2 * qs[i - 1] * n_airy) for n_airy in num_airys[i - 1]))
    grids, masks = [], []
    for i in range(levels):
        # make_focal_grid(q, num_airy, pupil_diameter, reference_wavelength=1, focal_length=1)
        g = make_focal_grid(qs[i], num_airys[i], spatial_resolution=1.0 / pupil_diameter)
        m = np.asarray(complex_mask(g), dtype=complex)
        if i != levels - 1:
            m = m * (1 - multiscale_window(window_size, g.shape))
        for j in range(i):
            fft = FFT(grids[j])
            m = m - mft_backward(_fft_forward_tensor(fft, masks[j]), g, fft.output_grid)
        grids.append(g)
        masks.append(m)
    return grids, masks


def multiscale_forward(E, pupil_grid, grids, masks, adjoint=False):
    """MultiScaleCoronagraph.forward / backward without the Lyot stop
    (hcipy/coronagraphy/multi_scale.py:108-182), at wavelength 1: level 0 is a FourierFilter
    with q = 2 and the level-0 mask as transfer function, level i adds
    P_i^dagger (M_i P_i E) with the Fraunhofer propagator P_i onto the level's grid (conj(M_i)
    for the backward direction)."""
    L = fourier_filter(E, pupil_grid, masks[0], q=2, adjoint=adjoint)
    for g, m in zip(grids[1:], masks[1:]):
        F = fraunhofer_forward(E, pupil_grid, g)
        if np.ndim(m) == 3:  # Jones-matrix mask (vortex.py:188-189, 237-238): field_dot
            F = field_dot(np.conj(m) if adjoint else m, F)
        else:
            F = F * (np.conj(m) if adjoint else m)
        L = L + fraunhofer_backward(F, pupil_grid, g)
    return L


def field_dot(J, E):
    """field_dot of a (p, q, n) matrix field with a (q, ..., n) field (hcipy/field/field.py)."""
    return np.einsum('ijn,j...n->i...n', np.asarray(J), np.asarray(E))


def stokes_vector(E, stokes=None):
    """Wavefront.I/.Q/.U/.V (hcipy/optics/wavefront.py:112-232) as a (4, n) array.  A Jones-vector
    field E (2, n) uses the fully polarized forms; a (2, 2, n) Jones-matrix field of a partially
    polarized wavefront is weighted by the input Stokes vector (a, b, c, d) with the rows of
    its Mueller matrix as the reference writes them."""
    E = np.asarray(E)
    if E.ndim == 2:
        e0, e1 = E
        return np.array([abs(e0) ** 2 + abs(e1) ** 2, abs(e0) ** 2 - abs(e1) ** 2,
                         2 * np.real(e0 * np.conj(e1)), -2 * np.imag(e0 * np.conj(e1))])
    x, y, z, w = E[0, 0], E[0, 1], E[1, 0], E[1, 1]
    xr, xi, yr, yi, zr, zi, wr, wi = x.real, x.imag, y.real, y.imag, z.real, z.imag, w.real, w.imag
    X, Y, Z, W = abs(x) ** 2, abs(y) ** 2, abs(z) ** 2, abs(w) ** 2
    rows = [
        (X + Y + Z + W, X - Y + Z - W, 2 * (xr * yr + xi * yi + zr * wr + zi * wi),
         2 * (-xr * yi + xi * yr - zr * wi + zi * wr)),
        (X + Y - Z - W, X - Y - Z + W, 2 * (xr * yr + xi * yi - zr * wr - zi * wi),
         2 * (-xr * yi + xi * yr + zr * wi - zi * wr)),
        (2 * (xr * zr + xi * zi + yr * wr + yi * wi), 2 * (xr * zr + xi * zi - yr * wr - yi * wi),
         2 * (xr * wr + xi * wi + yr * zr + yi * zi), 2 * (xi * wr - xr * wi + yr * zi - yi * zr)),
        (2 * (xr * zi - xi * zr + yr * wi - yi * wr), 2 * (xr * zi - xi * zr - yr * wi + yi * wr),
         2 * (xr * wi - xi * wr + yr * zi - yi * zr), 2 * (xr * wr + xi * wi - yr * zr - yi * zi)),
    ]
    a, b, c, d = stokes
    return np.array([0.5 * (r[0] * a + r[1] * b + r[2] * c + r[3] * d) for r in rows])


def jones_vector_of_stokes(stokes):
    """A Jones vector e (2,) whose fully polarized Stokes vector is ``stokes`` (s0 = |s|)."""
    s0, s1, s2, s3 = stokes
    ex = np.sqrt((s0 + s1) / 2)
    if ex == 0:
        return np.array([0, np.sqrt(s0)], dtype=complex)
    # e0 conj(e1) = (s2 - i s3) / 2  ->  e1 = conj((s2 - i s3) / 2) / ex
    return np.array([ex, np.conj((s2 - 1j * s3) / 2) / ex], dtype=complex)


def phase_retarder_jones(phase_retardation, fast_axis_orientation, circularity):
    """PhaseRetarder.jones_matrix (hcipy/optics/polarization.py:141-160)."""
    phi_plus = np.exp(1j * phase_retardation / 2)
    phi_minus = np.exp(-1j * phase_retardation / 2)
    c, s = np.cos(fast_axis_orientation), np.sin(fast_axis_orientation)
    j11 = phi_plus * c ** 2 + phi_minus * s ** 2
    j12 = (phi_plus - phi_minus) * np.exp(-1j * circularity) * c * s
    j21 = (phi_plus - phi_minus) * np.exp(1j * circularity) * c * s
    j22 = phi_plus * s ** 2 + phi_minus * c ** 2
    j11, j12, j21, j22 = np.broadcast_arrays(j11, j12, j21, j22)
    return np.array([[j11, j12], [j21, j22]])


def vector_vortex_mask(charge, phase_retardation=np.pi):
    """VectorVortexCoronagraph's level masks before windowing (hcipy/coronagraphy/vortex.py:
    84-94): LinearRetarder(retardance, charge / 2 theta) Jones matrix, origin sample removed."""
    def fn(grid):
        x, y = grid.coords
        J = phase_retarder_jones(phase_retardation, charge / 2 * np.arctan2(y, x), 0)
        return J * (1 - circular_aperture(grid, 1e-9))
    return fn


def _fft_forward_tensor(fft, f):
    """FFT.forward over the leading (tensor) dimensions of f."""
    f = np.asarray(f)
    lead = f.shape[:-1]
    flat = f.reshape((-1, f.shape[-1]))
    return np.array([fft.forward(v) for v in flat]).reshape(lead + (-1,))


def spectral_noise_fft(psd, output_grid, seed, oversample=1, shift=None):
    """SpectralNoiseFactoryFFT(psd, output_grid, oversample).make_random(seed), optionally
    shifted, evaluated (hcipy/util/spectral_noise.py:99-201): C = sqrt(psd / w (2 pi)^ndim) on the
    FFT grid times complex white noise from default_rng(seed), times exp(-i shift . k), then
    the real part of the FFT backward onto the output grid.  ``shift[i]`` moves along axis i."""
    fft = FFT(output_grid, q=oversample)
    g = fft.output_grid
    C = np.sqrt(np.asarray(psd(g), dtype=float) / g.weights * (2 * np.pi) ** g.ndim)
    rng = np.random.default_rng(seed)
    C = C * (rng.standard_normal(g.size) + 1j * rng.standard_normal(g.size))
    if shift is not None:
        C = C * np.exp(-1j * sum(shift[i] * g.coords[i] for i in range(g.ndim)))
    return fft.backward(C).real


def _radius(g):
    return np.hypot(*g.coords) if g.ndim == 2 else np.abs(g.coords[0])


def spectral_noise_multiscale(psd, output_grid, seed, oversampling, shift=None):
    """SpectralNoiseFactoryMultiscale(psd, output_grid, oversampling).make_random(seed),
    optionally shifted, evaluated (hcipy/util/spectral_noise.py:203-330).  The high frequencies
    sit on the FFT grid g1 of the output grid with |k| < boundary zeroed; the low ones on
    g2 = g1 scaled by 1/oversampling with |k| >= boundary zeroed, boundary = max |x| of g2.
    Noise is drawn for g1 (real, imaginary), then for g2; the screen is the real part of the
    FFT backward of the first plus the MFT backward (output grid <- g2) of the second.
    ``shift[i]`` moves along axis i."""
    fft = FFT(output_grid)
    g1 = fft.output_grid
    g2 = g1.scaled(1.0 / oversampling)
    boundary = np.abs(g2.coords[0]).max()
    nd = g1.ndim
    C1 = np.sqrt(np.asarray(psd(g1), dtype=float) / g1.weights * (2 * np.pi) ** nd)
    C1[_radius(g1) < boundary] = 0
    C2 = np.sqrt(np.asarray(psd(g2), dtype=float) / g2.weights * (2 * np.pi) ** nd)
    C2[_radius(g2) >= boundary] = 0
    rng = np.random.default_rng(seed)
    C1 = C1 * (rng.standard_normal(g1.size) + 1j * rng.standard_normal(g1.size))
    C2 = C2 * (rng.standard_normal(g2.size) + 1j * rng.standard_normal(g2.size))
    if shift is not None:
        C1 = C1 * np.exp(-1j * sum(shift[i] * g1.coords[i] for i in range(nd)))
        C2 = C2 * np.exp(-1j * sum(shift[i] * g2.coords[i] for i in range(nd)))
    return fft.backward(C1).real + mft_backward(C2, output_grid, g2).real


def power_law_psd(exponent):
    """The PSD of make_power_law_error (hcipy/optics/aberration.py:37-41): r^exponent, 0 at r = 0."""
    def psd(grid):
        r = np.hypot(*grid.coords) if grid.ndim == 2 else np.abs(grid.coords[0])
        with np.errstate(divide='ignore'):
            res = r ** exponent
        res[r == 0] = 0
        return res
    return psd


def power_law_error(pupil_grid, ptv, diameter, exponent=-2.5, aperture=None, seed=None):
    """make_power_law_error without mode removal (hcipy/optics/aberration.py:11-54)."""
    if aperture is None:
        aperture = circular_aperture(pupil_grid, diameter)
    screen = spectral_noise_fft(power_law_psd(exponent), pupil_grid, seed)
    screen = screen * (ptv / np.ptp(screen[np.asarray(aperture) != 0]))
    return screen * aperture


def super_gaussian_aperture(grid, fwhm, shape_parameter=5, center=(0, 0)):
    """make_super_gaussian_aperture (hcipy/aperture/generic.py:870-895): exp(-(r/sigma)^(2p)/2),
    sigma = fwhm / 2 / (2 ln 2)^(1/(2p))."""
    x = grid.sep[0][np.newaxis, :] - center[0]
    y = grid.sep[1][:, np.newaxis] - center[1]
    sigma = 0.5 * fwhm / (2 * np.log(2)) ** (1 / (2 * shape_parameter))
    r = np.hypot(x, y) / sigma
    return np.exp(-0.5 * r ** (2 * shape_parameter)).ravel()


def high_pass_power_law_error(pupil_grid, ptv, diameter, cutoff_frequency, exponent=-2.5, aperture=None,
                              filter_shape_parameter=15, seed=None):
    """make_high_pass_power_law_error (hcipy/optics/aberration.py:56-92): the power-law error
    filtered by 1 - super_gaussian(2 cutoff) with FourierFilter(q=2), real part, times the mask
    aperture > 0."""
    internal = make_fft_grid(pupil_grid, 2)
    tf = 1 - super_gaussian_aperture(internal, 2 * cutoff_frequency, filter_shape_parameter)
    if aperture is None:
        aperture = circular_aperture(pupil_grid, diameter)
    mask = (np.asarray(aperture) > 0).astype(float)
    screen = power_law_error(pupil_grid, ptv, diameter, exponent, aperture=None, seed=seed)
    return fourier_filter(screen.astype(complex), pupil_grid, tf, q=2).real * mask


# ----------------------------------------------------------------------------------------------
# Pyramid wavefront-sensor optics (hcipy/wavefront_sensing/pyramid.py:10-200)


class PyramidGeometry:
    """The focal grid, spatial filter and pyramid surface of PyramidWavefrontSensorOptics
    (pyramid.py:118-168)."""

    def __init__(self, input_grid, output_grid, separation=None, pupil_diameter=None, wavelength_0=1, q=None,
                 num_airy=None, refractive_index=lambda w: 1.5):
        ext_in = np.asarray(input_grid.delta) * np.asarray(input_grid.dims)
        ext_out = np.asarray(output_grid.delta) * np.asarray(output_grid.dims)
        self.pupil_diameter = np.max(ext_in) if pupil_diameter is None else pupil_diameter
        self.separation = self.pupil_diameter if separation is None else separation
        qmin = np.ceil(max(np.max(ext_out / ext_in), 2))
        q = qmin if q is None else q
        self.num_airy = np.max(input_grid.shape) / 2 if num_airy is None else num_airy
        num_pixels = 2 * int(self.num_airy * q)
        self.focal_grid = make_pupil_grid(num_pixels, 2 * wavelength_0 / self.pupil_diameter * self.num_airy)
        self.filter = circular_aperture(self.focal_grid, 2 * self.num_airy * wavelength_0 / self.pupil_diameter)
        self.refractive_index = refractive_index
        self.wavelength_0 = wavelength_0
        self.input_grid, self.output_grid = input_grid, output_grid

    def surface(self, grid):
        x = grid.sep[0][np.newaxis, :]
        y = grid.sep[1][:, np.newaxis]
        n0 = self.refractive_index(self.wavelength_0)
        return (-self.separation / (2 * (n0 - 1)) * (np.abs(x) + np.abs(y))).ravel()


def pyramid_wfs_forward(E, geom, wavelength=1.0, adjoint=False):
    """PyramidWavefrontSensorOptics.forward / backward (pyramid.py:170-200): Fraunhofer to the
    focal grid, the spatial filter, exp(+-i (n - 1) k surface), Fraunhofer to the output grid
    (backward: the adjoint chain in reverse)."""
    k = 2 * np.pi / wavelength
    n = geom.refractive_index(wavelength)
    ph = np.exp((-1j if adjoint else 1j) * (n - 1) * k * geom.surface(geom.focal_grid))
    if not adjoint:
        F = fraunhofer_forward(E, geom.input_grid, geom.focal_grid, wavelength)
        return fraunhofer_forward(F * geom.filter * ph, geom.focal_grid, geom.output_grid, wavelength)
    F = fraunhofer_backward(E, geom.focal_grid, geom.output_grid, wavelength)
    return fraunhofer_backward(F * geom.filter * ph, geom.input_grid, geom.focal_grid, wavelength)


def modulated_pyramid_wfs_forward(E, geom, modulation, num_steps=12, fast=False, wavelength=1.0):
    """ModulatedPyramidWavefrontSensorOptics.forward (pyramid.py:29-89): ``num_steps`` positions
    p on a circle of radius modulation/2.  Slow: the input times the tip-tilt mirror's
    exp(2 i k (p_x x + p_y y)), then the pyramid optics.  Fast: the filtered focal field times
    exp(i (n - 1) k s_p) with s_p the surface on the focal grid shifted by 2 p, supersampled 4x."""
    theta = np.linspace(0, 2 * np.pi, num_steps, endpoint=False)
    pts = np.stack([modulation / 2 * np.cos(theta), modulation / 2 * np.sin(theta)], axis=1)
    k = 2 * np.pi / wavelength
    out = []
    if fast:
        F = fraunhofer_forward(E, geom.input_grid, geom.focal_grid, wavelength) * geom.filter
        n = geom.refractive_index(wavelength)
        for p in pts:
            s = evaluate_supersampled(geom.surface, geom.focal_grid.shifted(2 * p), 4)
            out.append(fraunhofer_forward(F * np.exp(1j * (n - 1) * k * s), geom.focal_grid, geom.output_grid, wavelength))
        return out
    x, y = geom.input_grid.coords
    for p in pts:
        out.append(pyramid_wfs_forward(E * np.exp(2j * k * (p[0] * x + p[1] * y)), geom, wavelength))
    return out


def inverse_tikhonov(M, rcond=1e-15):
    """inverse_tikhonov (hcipy/util/matrix_inversion.py:71-100): from the SVD M = U S V^T,
    V diag(s / (s^2 + (rcond max s)^2)) U^T."""
    U, S, Vt = np.linalg.svd(np.asarray(M, dtype=float), full_matrices=False)
    S_inv = S / (S ** 2 + (rcond * S.max()) ** 2)
    return (Vt.T * S_inv).dot(U.T)


def actuator_positions(n_act, pitch):
    """make_actuator_positions without tilts (deformable_mirror.py:11-41)."""
    return make_uniform_grid(n_act, [pitch * n_act, pitch * n_act])


def gaussian_influence_functions(pupil_grid, n_act, pitch, crosstalk=0.15, cutoff=3, x_tilt=0, y_tilt=0, z_tilt=0,
                                 oversampling=1, only=None):
    """make_gaussian_influence_functions (deformable_mirror.py:43-93) with make_gaussian_pokes
    (mode_basis/gaussian_pokes.py:5-54) and evaluate_supersampled (field/util.py:606-698):
    each poke is evaluated on the pupil grid scaled by 1/cos([y_tilt, x_tilt]) and rotated by
    -z_tilt (R = [[cos, -sin], [sin, cos]] on the points, cartesian_grid.py:12-32), divided by
    cos(x_tilt) cos(y_tilt), and averaged over the oversampling dithers ((k + 0.5)/ovs - 0.5)
    times the grid step.  Returns a dense (n_act^2, Npix) array, or the rows of the actuator
    indices ``only``."""
    pos = actuator_positions(n_act, pitch)
    sigma = pitch / np.sqrt(-2 * np.log(crosstalk))
    cut = pitch / sigma * cutoff
    px, py = pos.coords
    if only is not None:
        px, py = px[np.asarray(only)], py[np.asarray(only)]
    x0 = pupil_grid.x
    y0 = pupil_grid.y
    dx, dy = pupil_grid.delta
    ovs = int(round(oversampling))
    dith = (np.arange(ovs) + 0.5) / ovs - 0.5
    out = np.zeros((len(px), pupil_grid.size))
    ca, sa = np.cos(-z_tilt), np.sin(-z_tilt)
    for fy in dith:
        for fx in dith:
            xs = (x0 + fx * dx) / np.cos(y_tilt)
            ys = (y0 + fy * dy) / np.cos(x_tilt)
            qx = ca * xs - sa * ys
            qy = sa * xs + ca * ys
            for i, (mx, my) in enumerate(zip(px, py)):
                r2 = (qx - mx) ** 2 + (qy - my) ** 2
                res = np.exp(-0.5 * r2 / sigma ** 2)
                res -= np.exp(-0.5 * cut ** 2)
                res[r2 > (cut * sigma) ** 2] = 0
                out[i] += res / (np.cos(x_tilt) * np.cos(y_tilt))
    return out / ovs ** 2


def xinetics_influence_functions(pupil_grid, n_act, pitch, table, x_tilt=0, y_tilt=0, z_tilt=0):
    """make_xinetics_influence_functions (hcipy/optics/deformable_mirror.py:95-138): the actuator
    shape ``table`` ([ny][nx], the DM5V2 FITS array) normalised to its maximum, on
    make_uniform_grid(shape, shape * pitch / 10), interpolated linearly (scipy
    RegularGridInterpolator, fill 0 outside, field/interpolation linear.py:5-50) at the pupil
    points scaled by 1/cos([y_tilt, x_tilt]), rotated by -z_tilt and shifted by -m_a, times
    cos(y_tilt) / cos(x_tilt) (the reference's "/ cos(x_tilt) * cos(y_tilt)").  Returns a dense
    (n_act^2, Npix) array."""
    from scipy.interpolate import RegularGridInterpolator
    actuator = np.asarray(table, dtype=float)
    actuator = actuator / actuator.max()
    shape = np.array(actuator.shape)
    tg = make_uniform_grid(list(actuator.shape), shape * pitch / 10.0)
    interp = RegularGridInterpolator((tg.sep[0], tg.sep[1]), actuator.T, 'linear', False, 0)
    xs = pupil_grid.x / np.cos(y_tilt)
    ys = pupil_grid.y / np.cos(x_tilt)
    ca, sa = np.cos(-z_tilt), np.sin(-z_tilt)
    qx = ca * xs - sa * ys
    qy = sa * xs + ca * ys
    px, py = actuator_positions(n_act, pitch).coords
    out = np.zeros((len(px), pupil_grid.size))
    for i, (mx, my) in enumerate(zip(px, py)):
        out[i] = interp(np.stack([qx - mx, qy - my], axis=-1)) / np.cos(x_tilt) * np.cos(y_tilt)
    return out


def dm_surface(influence, actuators):
    """ModeBasis.linear_combination (mode_basis/mode_basis.py:239-258)."""
    return np.asarray(actuators) @ influence


def dm_forward(E, surface, wavelength):
    """deformable_mirror.py:200-218: E * exp(2i k s)."""
    k = 2 * np.pi / wavelength
    return E * np.exp(2j * k * surface)


def intensity(E):
    return np.abs(E) ** 2


def total_power(E, grid):
    return np.sum(np.abs(E) ** 2 * grid.weights)


def mean_intensity_in_roi(I, mask):
    """get_mean_intensity_in_roi (metrics/contrast.py:37-52)."""
    return np.mean(I[np.asarray(mask) > 0])


# ----------------------------------------------------------------------------------------------
# Synthetic inputs of the benchmark workloads (seeded).


def gaussian_smoothed_phase(shape, sigma_px, rms, mask, seed):
    """i.i.d. N(0,1) per pixel (numpy default_rng(seed)), Gaussian smoothing (sigma in px,
    periodic, via FFT), scaled to `rms` radians over mask>0."""
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
    """Power-law PSD phase screen (PSD ~ f^exponent), FFT synthesis, rms over mask."""
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(shape[0])[:, None]
    kx = np.fft.fftfreq(shape[1])[None, :]
    f = np.hypot(kx, ky)
    f[0, 0] = np.inf
    amp = f ** (exponent / 2)
    c = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * amp
    ph = np.real(np.fft.ifft2(c))
    m = np.asarray(mask).reshape(shape) > 0
    ph -= ph[m].mean()
    ph *= rms / np.sqrt(np.mean(ph[m] ** 2))
    return ph.ravel()


def read_fits_image(path):
    """Reads the primary image of a (gzipped or plain) FITS file: 2880-byte header blocks,
    big-endian data of BITPIX -64 / -32 / 8 / 16 / 32."""
    import gzip
    with open(path, 'rb') as fh:
        raw = fh.read()
    if raw[:2] == b'\x1f\x8b':
        raw = gzip.decompress(raw)
    hdr = {}
    pos = 0
    while True:
        block = raw[pos:pos + 2880]
        pos += 2880
        end = False
        for i in range(0, 2880, 80):
            card = block[i:i + 80].decode('ascii')
            key = card[:8].strip()
            if key == 'END':
                end = True
                break
            if card[8:10] == '= ':
                val = card[10:].split('/')[0].strip()
                hdr[key] = val
        if end:
            break
    bitpix = int(hdr['BITPIX'])
    naxis = int(hdr['NAXIS'])
    dims = [int(hdr['NAXIS%d' % (i + 1)]) for i in range(naxis)]
    dt = {-64: '>f8', -32: '>f4', 8: 'u1', 16: '>i2', 32: '>i4'}[bitpix]
    n = int(np.prod(dims))
    data = np.frombuffer(raw[pos:pos + n * np.dtype(dt).itemsize], dtype=dt).astype(float)
    return data.reshape(dims[::-1])


# ----------------------------------------------------------------------------------------------
# CPU baseline: the reference's MFT with precomputed matrices (precompute_matrices=True,
# allocate_intermediate=True; hcipy/fourier/matrix_fourier_transform.py:67-111, 132-176).


class MftCpu:
    """Two BLAS gemm calls with the reference's transpose pattern (matrix_fourier_transform.py:171-172)."""

    def __init__(self, input_grid, output_grid, complex_dtype='complex64'):
        self.M1, self.M2 = mft_matrices(input_grid, output_grid, complex_dtype)
        self.shape_in = input_grid.shape
        self.alpha = input_grid.weights
        self.gemm = blas.cgemm if complex_dtype == 'complex64' else blas.zgemm
        self.inter = np.empty((input_grid.shape[0], self.M2.shape[1]), dtype=complex_dtype)

    def forward(self, f):
        fs = f.reshape(self.shape_in)
        self.gemm(1, self.M2.T, fs.T, c=self.inter.T, overwrite_c=True)
        return self.gemm(self.alpha, self.inter.T, self.M1.T).T.reshape(-1)


def time_it(function, t_max=0.1, repeat_max=5):
    """_time_it (hcipy/fourier/fourier_transform.py:185-215): double the iteration count until
    the run exceeds t_max, then the median of `repeat_max` runs (seconds per call)."""
    import time

    def run(n):
        t0 = time.perf_counter()
        for _ in range(n):
            function()
        return (time.perf_counter() - t0) / n

    n = 1
    while True:
        t = run(n)
        if t * n > t_max:
            break
        n *= 2
    if n == 1 and t > repeat_max * t_max:
        return t
    times = [t] + [run(n) for _ in range(repeat_max - 1)]
    return float(np.median(times))
