// Direct power-of-two FFT between a regular grid and its FFT grid (FastFourierTransform,
// hcipy/fourier/fast_fourier_transform.py:188-406), without the reference's padding buffer,
// fftshifts and shift arrays.
//
// With input samples x_n = x0 + n dx (n < n_in) and output samples u_k = u0 + k du (k < m) of an
// FFT grid (du dx = 2 pi / N, N the internal FFT length), the Fourier sum factors per axis as
//   sum_n f[n] exp(-i u_k x_n) = exp(-i u0 x0) exp(-i du x0 k) sum_n (f[n] exp(-i u0 dx n)) exp(-2 pi i k n / N),
// i.e. a ramp on the input, a plain length-N DFT of the input zero-padded at its end, the first
// m bins, and a ramp on the output.  The backward transform (the adjoint sum with exp(+i u x)) is
// the same with conjugate ramps and the inverse DFT.  On an FFT grid the ramp phases are rational
// multiples of 2 pi (a = u0 / du and 2 x0 / dx integral), which the plan reduces in integers
// before the FP64 sine and cosine, so large N loses no phase accuracy.
//
// Two passes over HBM: rows (input ramp, zero-padded row FFT, first m_x bins into a
// column-blocked intermediate) and columns (column FFT, first m_y bins, output ramps, scale),
// on the register-resident line FFT of line_fft.hpp.  1-D transforms finish in the row pass.
#include "common.hpp"
#include "line_fft.hpp"
#include "../../include/hcipy_amd.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

namespace hcipy {
namespace {

// intermediate layout: columns grouped by 8, each group a contiguous [rows][8] block; a row of the
// intermediate holds round_up(m, 8) columns
__host__ __device__ inline int64_t dblk(int64_t row, int col, int64_t rows) {
  return ((int64_t)(col >> 3) * rows + row) * 8 + (col & 7);
}
__host__ __device__ inline int64_t round_up_dev(int64_t m) { return (m + 7) / 8 * 8; }

// Workgroup b of n in an XCD-aware order: the hardware deals workgroups to the 8 XCDs round robin,
// and this bijection gives XCD x a contiguous range of logical indices, so workgroups that share
// cache lines (neighbouring columns of one 8-column block, or of one output row) meet in one L2.
__device__ inline int xcd_order(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

// LDS line of N points: FP64 lines exchange real and imaginary parts in turn (half the LDS, so
// more lines in flight per CU), FP32 lines exchange complex values.  LINE_T is the line stride
// in values of TC (even, so every line starts 16-byte aligned).
template <typename TC, int N>
struct DLine {
  static constexpr bool SPLIT = sizeof(TC) == 8;
  static constexpr int LINE_T = SPLIT ? hr_lds_points(N) + 2 : 2 * (hr_lds_points(N) + 1);
};

template <typename TC, int N>
struct DRowCfg {
  static constexpr int E = N >= 16 ? 16 : N;
  static constexpr int P = N / E;
  static constexpr int WG = P > 256 ? P : 256;
  static constexpr int LINES = WG / P;
  static constexpr size_t LDS_BYTES = (size_t)LINES * DLine<TC, N>::LINE_T * sizeof(TC);
  static_assert(LDS_BYTES <= 160 * 1024, "row pass exceeds the 160 KiB of LDS per CU");
};

template <typename TC, int N>
struct DColCfg {
  static constexpr int E = N >= 16 ? 16 : N;
  static constexpr int P = N / E;
  static constexpr int CF = (int)(71680 / ((size_t)DLine<TC, N>::LINE_T * sizeof(TC)));
  // FP64 columns need 160 VGPRs (3 waves per SIMD): workgroups of 256 threads fit three per CU
  static constexpr int CT = (sizeof(TC) == 8 ? 256 : 512) / P > 0 ? (sizeof(TC) == 8 ? 256 : 512) / P : 1;
  static constexpr int C0 = CF < CT ? CF : CT;
  static constexpr int C = C0 > 16 ? 16 : (C0 < 1 ? 1 : C0);
  static constexpr int NT = C * P;
  static constexpr size_t LDS_BYTES = (size_t)C * DLine<TC, N>::LINE_T * sizeof(TC);
  static_assert(LDS_BYTES <= 160 * 1024, "column pass exceeds the 160 KiB of LDS per CU");
};

template <typename TC, int N>
__device__ inline typename Cplx<TC>::V* dline(unsigned char* raw, int i) {
  return reinterpret_cast<typename Cplx<TC>::V*>(reinterpret_cast<TC*>(raw) + (size_t)i * DLine<TC, N>::LINE_T);
}

template <typename TC>
__device__ inline typename Cplx<TC>::V from_d2(double2 a) {
  typename Cplx<TC>::V r;
  r.x = (TC)a.x;
  r.y = (TC)a.y;
  return r;
}

// Row pass: line j of batch item b = in[j][:] * ax * ay[j], zero beyond nx, FFT of NX points,
// bins k < mx.  final != 0 (1-D): the bins times cx * cy[0] * alpha go straight to the output.
template <typename TI, typename TC, int NX>
__global__ __launch_bounds__((DRowCfg<TC, NX>::WG)) void dfft_rows_kernel(
    const typename Cplx<TI>::V* __restrict__ in, typename Cplx<TC>::V* __restrict__ w1,
    typename Cplx<TI>::V* __restrict__ out, const typename Cplx<TC>::V* __restrict__ tw,
    const double2* __restrict__ ax, const double2* __restrict__ ay, const double2* __restrict__ cx,
    const double2* __restrict__ cy, double alpha_re, double alpha_im, int nx, int ny, int mx, int sign, int final_) {
  using V = typename Cplx<TC>::V;
  using VI = typename Cplx<TI>::V;
  using Cfg = DRowCfg<TC, NX>;
  constexpr int E = Cfg::E;
  constexpr int P = Cfg::P;
  extern __shared__ __align__(16) unsigned char dfft_lds_raw[];
  const int line = threadIdx.x / P;
  const int tl = threadIdx.x % P;
  const int j = blockIdx.x * Cfg::LINES + line;
  const int64_t b = blockIdx.y;
  const bool valid = j < ny;
  // point tl + q P of the line exists for q < q_in (a count, so the guards need no per-slot index)
  const int q_in = valid ? (nx - tl + P - 1) / P : 0;
  const VI* src = in + (b * ny + (valid ? j : 0)) * (int64_t)nx + tl;
  const double2* axp = ax + tl;
  const V ryj = from_d2<TC>(ay[valid ? j : 0]);
  V v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    V x{0, 0};
    if (q < q_in) {
      const VI f = src[q * P];
      V fv;
      fv.x = (TC)f.x;
      fv.y = (TC)f.y;
      x = cmul(cmul(fv, from_d2<TC>(axp[q * P])), ryj);
    }
    v[q] = x;
  }
  hr_fft<TC, NX, P, DLine<TC, NX>::SPLIT>(v, dline<TC, NX>(dfft_lds_raw, line), tl, tw, sign);
  if (!valid) return;
  const int q_out = (mx - tl + P - 1) / P;
  if (final_) {
    V a;
    a.x = (TC)alpha_re;
    a.y = (TC)alpha_im;
    const V c0 = cmul(from_d2<TC>(cy[0]), a);
    VI* o = out + b * (int64_t)mx + tl;
    const double2* cxp = cx + tl;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      if (q < q_out) {
        const V r = cmul(cmul(v[q], from_d2<TC>(cxp[q * P])), c0);
        o[q * P] = VI{(TI)r.x, (TI)r.y};
      }
    }
    return;
  }
  V* o = w1 + b * (int64_t)ny * round_up_dev(mx);
  if constexpr (P % 8 == 0) {
    // bin tl + q P sits in 8-column block tl / 8 + q P / 8
    V* ot = o + dblk(j, tl, ny);
#pragma unroll
    for (int q = 0; q < E; ++q)
      if (q < q_out) ot[(int64_t)q * (P / 8) * ny * 8] = v[q];
  } else {
#pragma unroll
    for (int q = 0; q < E; ++q)
      if (q < q_out) o[dblk(j, tl + q * P, ny)] = v[q];
  }
}

// Column pass: column kx of the intermediate (ny rows, zero to NY), FFT of NY points, bins
// ky < my times cy[ky] cx[kx] alpha, written row-major (my x mx).  C adjacent columns per
// workgroup (in the XCD-aware order), thread (c, tl) = (tid % C, tid / C).
template <typename TI, typename TC, int NY>
__global__ __launch_bounds__((DColCfg<TC, NY>::NT)) void dfft_cols_kernel(
    const typename Cplx<TC>::V* __restrict__ w1, typename Cplx<TI>::V* __restrict__ out,
    const typename Cplx<TC>::V* __restrict__ tw, const double2* cx, const double2* cy,
    double alpha_re, double alpha_im, int ny, int mx, int my, int sign) {
  using V = typename Cplx<TC>::V;
  using VI = typename Cplx<TI>::V;
  using Cf = DColCfg<TC, NY>;
  constexpr int C = Cf::C;
  constexpr int E = Cf::E;
  constexpr int P = Cf::P;
  extern __shared__ __align__(16) unsigned char dfft_lds_raw[];
  const int c = threadIdx.x % C;
  const int tl = threadIdx.x / C;
  const int kx = xcd_order(blockIdx.x, gridDim.x) * C + c;
  const int64_t b = blockIdx.y;
  const bool col = kx < mx;
  // row tl + q P of the column exists for q < q_in
  const int q_in = col ? (ny - tl + P - 1) / P : 0;
  const V* src = w1 + b * (int64_t)ny * round_up_dev(mx) + dblk(tl, col ? kx : 0, ny);
  V v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) v[q] = q < q_in ? src[q * P * 8] : V{0, 0};
  hr_fft<TC, NY, P, DLine<TC, NY>::SPLIT>(v, dline<TC, NY>(dfft_lds_raw, c), tl, tw, sign);
  if (!col) return;
  V a;
  a.x = (TC)alpha_re;
  a.y = (TC)alpha_im;
  const V cxa = cmul(from_d2<TC>(cx[kx]), a);
  const int q_out = (my - tl + P - 1) / P;
  VI* o = out + (b * (int64_t)my + tl) * mx + kx;
  const double2* cyp = cy + tl;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    if (q < q_out) {
      const V r = cmul(cmul(v[q], from_d2<TC>(cyp[q * P])), cxa);
      o[(int64_t)q * P * mx] = VI{(TI)r.x, (TI)r.y};
    }
  }
}

__global__ void dfft_twiddle_kernel(int n, int f64, void* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double s, c;
  sincospi(-2.0 * (double)t / (double)n, &s, &c);
  if (f64) static_cast<double2*>(out)[t] = make_double2(c, s);
  else static_cast<float2*>(out)[t] = make_float2((float)c, (float)s);
}

// Host: exp(i pi t) for t given as num / den with integer num, den (exact reduction mod 2 den).
double2 expi_pi_rational(int64_t num, int64_t den) {
  int64_t r = num % (2 * den);
  if (r < 0) r += 2 * den;
  if (r > den) r -= 2 * den;  // t in (-1, 1]
  const double t = (double)r / (double)den;
  return make_double2(std::cos(M_PI * t), std::sin(M_PI * t));
}

bool near_int(double v, int64_t* iv) {
  const double r = std::nearbyint(v);
  // within the rounding of the grid's own coordinates (a few ulps); products stay in int64
  if (std::fabs(v - r) <= 1e-13 * std::max(1.0, std::fabs(v)) && std::fabs(r) < 1e9) {
    *iv = (int64_t)r;
    return true;
  }
  return false;
}

// Ramp tables of one axis: forward input ramp a_f[n] = exp(-i u0 dx n), forward output ramp
// c_f[k] = exp(-i u0 x0) exp(-i du x0 k); backward input ramp a_b[k] = exp(+i du x0 k), backward
// output ramp c_b[n] = exp(+i u0 x0) exp(+i u0 dx n).
void axis_ramps(int n, double dx, double x0, int m, double du, double u0, int N, std::vector<double2>& af,
                std::vector<double2>& cf, std::vector<double2>& ab, std::vector<double2>& cb) {
  // u0 dx = pi a2 / N, du x0 = pi b2 / N, u0 x0 = pi a2 b2 / (2 N) with a2 = 2 u0 / du and
  // b2 = 2 x0 / dx, integers on centred FFT grids of either parity (half-sample origins)
  int64_t a2 = 0, b2 = 0;
  const bool ia = near_int(2.0 * u0 / du, &a2);
  const bool ib = near_int(2.0 * x0 / dx, &b2);
  auto lin = [&](bool exact, int64_t num_per_step, double phase_per_step, int64_t den, int64_t t, double sign) {
    if (exact) return expi_pi_rational((int64_t)sign * num_per_step * t, den);
    const double ph = sign * phase_per_step * (double)t;
    return make_double2(std::cos(ph), std::sin(ph));
  };
  double2 c0;  // exp(-i u0 x0)
  if (ia && ib) c0 = expi_pi_rational(-a2 * b2, 2 * (int64_t)N);
  else c0 = make_double2(std::cos(u0 * x0), -std::sin(u0 * x0));
  af.resize(n);
  cb.resize(n);
  for (int t = 0; t < n; ++t) {
    af[t] = lin(ia, a2, u0 * dx, N, t, -1.0);
    const double2 e = lin(ia, a2, u0 * dx, N, t, +1.0);
    cb[t] = make_double2(e.x * c0.x + e.y * c0.y, e.y * c0.x - e.x * c0.y);  // e * conj(c0)
  }
  cf.resize(m);
  ab.resize(m);
  for (int t = 0; t < m; ++t) {
    const double2 e = lin(ib, b2, du * x0, N, t, -1.0);
    cf[t] = make_double2(e.x * c0.x - e.y * c0.y, e.x * c0.y + e.y * c0.x);
    ab[t] = lin(ib, b2, du * x0, N, t, +1.0);
  }
}

}  // namespace
}  // namespace hcipy

using namespace hcipy;

struct hcipy_fft_plan_s {
  int ndim = 0, dtype = 0, precision = 0, device = 0;
  int n[2] = {1, 1}, m[2] = {1, 1}, N[2] = {1, 1};  // input dims, output dims, FFT lengths (x, y)
  void* tw[2] = {nullptr, nullptr};                 // twiddles of N[0], N[1] in the compute precision
  double2* ramps = nullptr;  // per axis: a_f[n], c_f[m], a_b[m], c_b[n]
  int64_t off[2][4] = {};    // offsets (elements) of the tables in `ramps`
};

namespace {

bool dfft_pow2(int n) { return n >= 4 && n <= 8192 && (n & (n - 1)) == 0; }

#define HCIPY_DFFT_SWITCH(N, CALL) \
  switch (N) {                     \
    case 4: CALL(4); break;        \
    case 8: CALL(8); break;        \
    case 16: CALL(16); break;      \
    case 32: CALL(32); break;      \
    case 64: CALL(64); break;      \
    case 128: CALL(128); break;    \
    case 256: CALL(256); break;    \
    case 512: CALL(512); break;    \
    case 1024: CALL(1024); break;  \
    case 2048: CALL(2048); break;  \
    case 4096: CALL(4096); break;  \
    case 8192: CALL(8192); break;  \
    default: break;                \
  }

template <typename TI, typename TC, int N>
int launch_drows(const void* in, void* w1, void* out, const void* tw, const double2* ax, const double2* ay,
                 const double2* cx, const double2* cy, double ar, double ai, int nx, int ny, int mx, int sign, int final_,
                 int64_t batch, hipStream_t s) {
  using Cfg = DRowCfg<TC, N>;
  auto kern = dfft_rows_kernel<TI, TC, N>;
  if (Cfg::LDS_BYTES > 65536)
    HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)Cfg::LDS_BYTES));
  const dim3 g((unsigned)ceil_div(ny, Cfg::LINES), (unsigned)batch);
  hipLaunchKernelGGL(kern, g, dim3(Cfg::WG), Cfg::LDS_BYTES, s, static_cast<const typename Cplx<TI>::V*>(in),
                     static_cast<typename Cplx<TC>::V*>(w1), static_cast<typename Cplx<TI>::V*>(out),
                     static_cast<const typename Cplx<TC>::V*>(tw), ax, ay, cx, cy, ar, ai, nx, ny, mx, sign, final_);
  return check_hip(hipGetLastError(), "dfft_rows_kernel launch");
}

template <typename TI, typename TC, int N>
int launch_dcols(const void* w1, void* out, const void* tw, const double2* cx, const double2* cy, double ar, double ai,
                 int ny, int mx, int my, int sign, int64_t batch, hipStream_t s) {
  using Cf = DColCfg<TC, N>;
  auto kern = dfft_cols_kernel<TI, TC, N>;
  if (Cf::LDS_BYTES > 65536)
    HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)Cf::LDS_BYTES));
  const dim3 g((unsigned)ceil_div(mx, Cf::C), (unsigned)batch);
  hipLaunchKernelGGL(kern, g, dim3(Cf::NT), Cf::LDS_BYTES, s, static_cast<const typename Cplx<TC>::V*>(w1),
                     static_cast<typename Cplx<TI>::V*>(out), static_cast<const typename Cplx<TC>::V*>(tw), cx, cy, ar,
                     ai, ny, mx, my, sign);
  return check_hip(hipGetLastError(), "dfft_cols_kernel launch");
}

size_t dfft_compute_bytes(const hcipy_fft_plan_s& p) {
  return (p.dtype == HCIPY_DTYPE_C128 || p.precision == HCIPY_PREC_FP64) ? 16 : 8;
}

// Workspace of `batch` transforms in either direction: the intermediate [rows][round_up(cols, 8)].
size_t dfft_ws_bytes(const hcipy_fft_plan_s& p, int64_t batch) {
  if (p.ndim == 1) return 0;
  const int64_t fwd = (int64_t)p.n[1] * round_up_dev(p.m[0]);
  const int64_t bwd = (int64_t)p.m[1] * round_up_dev(p.n[0]);
  return (size_t)std::max(fwd, bwd) * (size_t)batch * dfft_compute_bytes(p);
}

template <typename TI, typename TC>
int run_dfft(const hcipy_fft_plan_s& p, int dir, const void* in, void* out, int64_t batch, double ar, double ai,
             char* ws, hipStream_t s) {
  // forward: (n_y x n_x) -> (m_y x m_x), tables a_f / c_f, sign -1; backward: the reverse
  const bool fw = dir == HCIPY_FORWARD;
  const int nx = fw ? p.n[0] : p.m[0], ny = fw ? p.n[1] : p.m[1];
  const int mx = fw ? p.m[0] : p.n[0], my = fw ? p.m[1] : p.n[1];
  const int ai_ = fw ? 0 : 2, ci_ = fw ? 1 : 3;
  const double2* axt = p.ramps + p.off[0][ai_];
  const double2* cxt = p.ramps + p.off[0][ci_];
  const double2* ayt = p.ramps + p.off[1][ai_];
  const double2* cyt = p.ramps + p.off[1][ci_];
  const int sign = fw ? -1 : +1;
  const size_t fin = (size_t)nx * ny * 2 * sizeof(TI), fout = (size_t)mx * my * 2 * sizeof(TI);
  const size_t fws = (size_t)ny * round_up_dev(mx) * 2 * sizeof(TC);
  int st = kOk;
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const int64_t nb = std::min<int64_t>(65535, batch - b0);
    const char* ib = static_cast<const char*>(in) + (size_t)b0 * fin;
    char* ob = static_cast<char*>(out) + (size_t)b0 * fout;
    char* wb = ws != nullptr ? ws + (size_t)b0 * fws : nullptr;
#define ROWS(NN)                                                                                                   \
  st = launch_drows<TI, TC, NN>(ib, wb, ob, p.tw[0], axt, ayt, cxt, cyt, ar, ai, nx, ny, mx, sign, p.ndim == 1, nb, \
                                s)
    HCIPY_DFFT_SWITCH(p.N[0], ROWS)
#undef ROWS
    if (st) return st;
    if (p.ndim == 2) {
#define COLS(NN) st = launch_dcols<TI, TC, NN>(wb, ob, p.tw[1], cxt, cyt, ar, ai, ny, mx, my, sign, nb, s)
      HCIPY_DFFT_SWITCH(p.N[1], COLS)
#undef COLS
      if (st) return st;
    }
  }
  return kOk;
}

}  // namespace

extern "C" {

int hcipy_fft_plan_destroy(hcipy_fft_plan plan) {
  if (plan == nullptr) return kOk;
  for (int d = 0; d < 2; ++d)
    if (plan->tw[d]) (void)hipFree(plan->tw[d]);
  if (plan->ramps) (void)hipFree(plan->ramps);
  delete plan;
  return kOk;
}

int hcipy_fft_plan_create(hcipy_fft_plan* plan, int ndim, const int* n_in, const double* delta_in,
                          const double* zero_in, const int* n_out, const double* delta_out, const double* zero_out,
                          const int* fft_n, int dtype, int precision, void* stream) {
  HCIPY_REQUIRE(plan != nullptr, "plan pointer is NULL");
  *plan = nullptr;
  HCIPY_REQUIRE(ndim == 1 || ndim == 2, "the direct FFT supports one- and two-dimensional grids");
  HCIPY_REQUIRE(dtype == HCIPY_DTYPE_C64 || dtype == HCIPY_DTYPE_C128, "FFT dtype must be complex");
  HCIPY_REQUIRE(precision == HCIPY_PREC_FP64 || (precision == HCIPY_PREC_FP32 && dtype == HCIPY_DTYPE_C64),
                "FFT precision must be HCIPY_PREC_FP64, or HCIPY_PREC_FP32 for complex64 fields");
  HCIPY_REQUIRE(n_in && delta_in && zero_in && n_out && delta_out && zero_out && fft_n, "NULL grid argument");
  for (int d = 0; d < ndim; ++d) {
    HCIPY_REQUIRE(dfft_pow2(fft_n[d]), "axis %d: FFT length %d is not a power of two in [4, 8192]", d, fft_n[d]);
    HCIPY_REQUIRE(n_in[d] >= 1 && n_in[d] <= fft_n[d] && n_out[d] >= 1 && n_out[d] <= fft_n[d],
                  "axis %d: %d input and %d output samples for an FFT of %d", d, n_in[d], n_out[d], fft_n[d]);
    const double r = delta_in[d] * delta_out[d] * (double)fft_n[d] / (2.0 * M_PI);
    HCIPY_REQUIRE(std::fabs(r - 1.0) < 1e-9, "axis %d: delta_in * delta_out = %.17g is not 2 pi / %d", d,
                  delta_in[d] * delta_out[d], fft_n[d]);
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (int st = refuse_under_capture(s, "hcipy_fft_plan_create")) return st;
  hcipy_fft_plan_s* h = new (std::nothrow) hcipy_fft_plan_s();
  HCIPY_REQUIRE(h != nullptr, "out of host memory");
  auto fail = [&](int st) {
    hcipy_fft_plan_destroy(h);
    return st;
  };
  h->ndim = ndim;
  h->dtype = dtype;
  h->precision = precision;
  int st = check_hip(hipGetDevice(&h->device), "hipGetDevice");
  if (st) return fail(st);
  std::vector<double2> all;
  for (int d = 0; d < 2; ++d) {
    std::vector<double2> af, cf, ab, cb;
    if (d < ndim) {
      h->n[d] = n_in[d];
      h->m[d] = n_out[d];
      h->N[d] = fft_n[d];
      axis_ramps(n_in[d], delta_in[d], zero_in[d], n_out[d], delta_out[d], zero_out[d], fft_n[d], af, cf, ab, cb);
    } else {
      af = cf = ab = cb = {make_double2(1.0, 0.0)};
    }
    const std::vector<double2>* t[4] = {&af, &cf, &ab, &cb};
    for (int q = 0; q < 4; ++q) {
      h->off[d][q] = (int64_t)all.size();
      all.insert(all.end(), t[q]->begin(), t[q]->end());
    }
  }
  if ((st = check_hip(hipMalloc(&h->ramps, all.size() * sizeof(double2)), "hipMalloc(ramps)"))) return fail(st);
  if ((st = check_hip(hipMemcpy(h->ramps, all.data(), all.size() * sizeof(double2), hipMemcpyHostToDevice),
                      "copy ramps")))
    return fail(st);
  const bool f64 = dfft_compute_bytes(*h) == 16;
  for (int d = 0; d < ndim; ++d) {
    if ((st = check_hip(hipMalloc(&h->tw[d], (size_t)h->N[d] * (f64 ? 16 : 8)), "hipMalloc(twiddles)"))) return fail(st);
    hipLaunchKernelGGL(dfft_twiddle_kernel, dim3((unsigned)ceil_div(h->N[d], 256)), dim3(256), 0, s, h->N[d], f64 ? 1 : 0,
                       h->tw[d]);
  }
  if ((st = check_hip(hipGetLastError(), "dfft_twiddle_kernel"))) return fail(st);
  if ((st = check_hip(hipStreamSynchronize(s), "FFT plan creation"))) return fail(st);
  *plan = h;
  return kOk;
}

int hcipy_fft_workspace_size(hcipy_fft_plan plan, int64_t batch, size_t* bytes) {
  HCIPY_REQUIRE(plan != nullptr && bytes != nullptr && batch >= 0, "bad argument");
  *bytes = dfft_ws_bytes(*plan, std::min<int64_t>(batch, 65535));
  return kOk;
}

int hcipy_fft_execute(hcipy_fft_plan plan, int direction, const void* in, void* out, int64_t batch, double alpha_re,
                      double alpha_im, void* workspace, size_t workspace_bytes, void* stream) {
  HCIPY_REQUIRE(plan != nullptr, "NULL plan");
  HCIPY_REQUIRE(direction == HCIPY_FORWARD || direction == HCIPY_BACKWARD, "bad direction %d", direction);
  HCIPY_REQUIRE(batch >= 0, "negative batch");
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(in != nullptr && out != nullptr && in != out, "NULL or aliased data pointer");
  const size_t need = dfft_ws_bytes(*plan, std::min<int64_t>(batch, 65535));
  HCIPY_REQUIRE(need == 0 || (workspace != nullptr && workspace_bytes >= need), "workspace too small: %zu < %zu bytes",
                workspace_bytes, need);
  int cur = -1;
  HCIPY_HIP(hipGetDevice(&cur));
  HCIPY_REQUIRE(cur == plan->device, "plan was created on device %d but device %d is current", plan->device, cur);
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* ws = static_cast<char*>(workspace);
  if (plan->dtype == HCIPY_DTYPE_C128) return run_dfft<double, double>(*plan, direction, in, out, batch, alpha_re, alpha_im, ws, s);
  if (plan->precision == HCIPY_PREC_FP64) return run_dfft<float, double>(*plan, direction, in, out, batch, alpha_re, alpha_im, ws, s);
  return run_dfft<float, float>(*plan, direction, in, out, batch, alpha_re, alpha_im, ws, s);
}

}  // extern "C"
