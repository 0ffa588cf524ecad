// Zoom FFT: the Fourier transform between two regular Cartesian grids of any spacing, as one
// Bluestein chirp-z transform along each axis (reference: hcipy/fourier/zoom_fast_fourier_transform.py
// and hcipy/fourier/chirp_z_transform.py).
//
// One axis pass maps rows of n samples at p_j = z_in + j d_in to m samples at q_k = z_out + k d_out:
//
//   out_k = alpha * sum_j x_j exp(s i p_j q_k),   s = -1 forward, +1 backward.
//
// With D = d_in d_out and jk = (j^2 + k^2 - (k - j)^2) / 2 this is
//
//   out_k = alpha * C_k * sum_j (x_j A_j) h_(k-j),
//   A_j = exp(s i (j d_in z_out + D j^2 / 2)),  h_d = exp(-s i D d^2 / 2),
//   C_k = exp(s i (D k^2 / 2 + z_in q_k)),
//
// a linear convolution evaluated as a circular one of length L >= n + m - 1 (a power of two):
// h_d sits at index d mod L for d in [-(n - 1), m - 1], zero elsewhere, so no wrapped term reaches
// the outputs k < m.  Per row, in LDS: y = x A (zero-padded to L), FFT, times H = FFT(h) / L,
// inverse FFT, outputs k < m times alpha C_k; the line FFTs are register-resident radix-16
// Stockham stages with one LDS round trip per stage.
//
// A pass writes its output transposed, out[b][k][row], so that the next axis again reads rows:
// the 2-D forward transform is pass x (rows of nx -> [mx][ny]) then pass y (rows of ny ->
// [my][mx], the output grid's layout).  The chirps and H are computed once per plan in FP64 on
// the device (H by an FP64 line FFT) and kept in the plan's arithmetic precision: FP64 for
// complex128 fields and for complex64 fields at HCIPY_PREC_FP64 (FP32 loads and stores, FP64
// line FFTs and an FP64 intermediate), FP32 at HCIPY_PREC_FP32.
#include "common.hpp"
#include "line_fft.hpp"
#include "../../include/hcipy_amd.h"

#include <algorithm>
#include <cmath>
#include <new>

namespace hcipy {

namespace {

constexpr int kCztMaxL = 8192;

template <int L>
struct CztCfg {
  static constexpr int TPL = (L / 4 < 256) ? (L / 4 < 1 ? 1 : L / 4) : 256;  // threads per line
  static constexpr int LINES = 256 / TPL;                                     // lines per workgroup
};

// Chirps of one axis pass in FP64 (formulas in the header comment).  One thread per index of
// the longest array; h is written in FFT order (index d mod L).
// exp(i s pi (t^2 mod 2N) / N) = exp(i s D t^2 / 2) for D = 2 pi / N exactly: the quadratic
// chirp phase of a transform onto the FFT grid of N points, reduced in integers, so that its
// rounding does not grow with t^2.
__device__ inline double2 dft_chirp(int64_t t, int fft_n, double s) {
  const int64_t r = (t * t) % (2 * (int64_t)fft_n);
  double sn, cs;
  sincospi(s * (double)r / (double)fft_n, &sn, &cs);
  return make_double2(cs, sn);
}

__device__ inline double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// fft_n > 0: the grids are an FFT pair, d_in * d_out = 2 pi / fft_n; the quadratic chirp terms
// use that exact value (the sum is then the reference FFT's exact DFT) and are multiplied by
// the linear terms, whose phases stay below pi n m / fft_n.  fft_n = 0: one FP64 phase per term.
__global__ void czt_chirp_kernel(int n, int m, int L, double d_in, double z_in, double d_out, double z_out, int sign,
                                 int fft_n, double2* __restrict__ A, double2* __restrict__ h,
                                 double2* __restrict__ C) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L) return;
  const double s = (double)sign;
  const double D = d_in * d_out;
  double sn, cs;
  if (t < n) {
    const double j = (double)t;
    if (fft_n > 0) {
      sincos(s * (j * d_in * z_out), &sn, &cs);
      A[t] = cmul2(make_double2(cs, sn), dft_chirp(t, fft_n, s));
    } else {
      sincos(s * (j * d_in * z_out + 0.5 * D * (j * j)), &sn, &cs);
      A[t] = make_double2(cs, sn);
    }
  }
  if (t < m) {
    const double k = (double)t;
    if (fft_n > 0) {
      sincos(s * (z_in * (z_out + k * d_out)), &sn, &cs);
      C[t] = cmul2(make_double2(cs, sn), dft_chirp(t, fft_n, s));
    } else {
      sincos(s * (0.5 * D * (k * k) + z_in * (z_out + k * d_out)), &sn, &cs);
      C[t] = make_double2(cs, sn);
    }
  }
  // h at index t: d = t for t < m, d = t - L for t >= L - (n - 1), zero in between
  double2 hv = make_double2(0.0, 0.0);
  int d = 0;
  bool on = false;
  if (t < m) {
    d = t;
    on = true;
  } else if (t >= L - (n - 1)) {
    d = t - L;
    on = true;
  }
  if (on) {
    if (fft_n > 0) {
      hv = dft_chirp(d, fft_n, -s);
    } else {
      const double dd = (double)d;
      sincos(-s * 0.5 * D * (dd * dd), &sn, &cs);
      hv = make_double2(cs, sn);
    }
  }
  h[t] = hv;
}

// H = FFT(h) / L in FP64 by one workgroup, written as FP64 and (for complex64 plans) FP32.
template <int L>
__global__ __launch_bounds__(256) void czt_filter_kernel(const double2* __restrict__ h,
                                                         const double2* __restrict__ tw64, double2* __restrict__ H64,
                                                         float2* __restrict__ H32) {
  using V = typename Cplx<double>::V;
  constexpr int TPL = 256;
  extern __shared__ __align__(16) unsigned char czt_lds_raw[];
  V* line = reinterpret_cast<V*>(czt_lds_raw);
  for (int t = threadIdx.x; t < L; t += TPL) {
    const double2 v = h[t];
    line[t] = V{v.x, v.y};
  }
  __syncthreads();
  // threads at or beyond the line's thread count skip every butterfly but join the barriers
  line_fft<double, L, CztCfg<L>::TPL>(line, (int)threadIdx.x, reinterpret_cast<const V*>(tw64), -1, 0);
  const double inv = 1.0 / (double)L;
  for (int t = threadIdx.x; t < L; t += TPL) {
    const V v = line[t];
    H64[t] = make_double2(v.x * inv, v.y * inv);
    if (H32) H32[t] = make_float2((float)(v.x * inv), (float)(v.y * inv));
  }
}

__global__ void czt_twiddle_kernel(int L, double2* __restrict__ tw64, float2* __restrict__ tw32) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L) return;
  double s, c;
  sincospi(-2.0 * (double)t / (double)L, &s, &c);
  tw64[t] = make_double2(c, s);
  if (tw32) tw32[t] = make_float2((float)c, (float)s);
}

__global__ void cast_c128_to_c64_kernel(const double2* __restrict__ in, float2* __restrict__ out, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = make_float2((float)in[t].x, (float)in[t].y);
}

// Thread layout of the register-resident line FFT: E = 8 points per thread (measured faster than
// 16 on the 2048-point lines of a 1024^2 -> 512^2 transform: twice the waves for the few rows a
// pass has), P = N / E threads per line, workgroups of max(256, P) threads holding LINES lines.
// FP32 lines read their twiddles from an LDS copy (measured faster); FP64 lines read them from
// global memory (the LDS copy lowers their occupancy and was measured slower).
template <typename T, int N>
struct HrCfg {
  static constexpr int E = N >= 8 ? 8 : N;
  static constexpr int P = N / E;
  static constexpr int WG = P > 256 ? P : 256;
  static constexpr int LINES = WG / P;
  static constexpr bool TW_LDS = sizeof(T) == 4;
  static constexpr int LINE = hr_lds_points(N);  // LDS slots per line
  static constexpr size_t LDS_BYTES = ((size_t)LINES * LINE + (TW_LDS ? N : 0)) * 2 * sizeof(T);
};

// One axis pass over `total_rows` rows of n samples (row r = b * rows + y), output transposed to
// out[(b * m + k) * rows + y].  Each row: load x A (zero-padded to L) in registers, FFT, times H,
// inverse FFT, outputs k < m times alpha C.  Workgroups take LINES consecutive rows; the
// workgroup index is remapped so that the consecutive row groups of one output cache line run on
// one XCD (the hardware deals workgroups to the 8 XCDs round robin).
template <typename TI, typename TO, typename T, int L>
__global__ __launch_bounds__((HrCfg<T, L>::WG)) void czt_rows_kernel(const typename Cplx<TI>::V* __restrict__ in,
                                                                typename Cplx<TO>::V* __restrict__ out,
                                                                const typename Cplx<T>::V* __restrict__ A,
                                                                const typename Cplx<T>::V* __restrict__ H,
                                                                const typename Cplx<T>::V* __restrict__ C,
                                                                const typename Cplx<T>::V* __restrict__ tw, int n,
                                                                int m, int rows, int64_t total_rows, int64_t groups,
                                                                T alpha_re, T alpha_im) {
  using V = typename Cplx<T>::V;
  using VI = typename Cplx<TI>::V;
  using VO = typename Cplx<TO>::V;
  using Cfg = HrCfg<T, L>;
  constexpr int E = Cfg::E;
  constexpr int P = Cfg::P;
  constexpr int LINES = Cfg::LINES;
  extern __shared__ __align__(16) unsigned char czt_lds_raw[];
  V* lds = reinterpret_cast<V*>(czt_lds_raw);
  const int64_t per_xcd = (groups + 7) / 8;
  const int64_t g = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (g >= groups) return;  // uniform over the workgroup, before any barrier
  const int line = threadIdx.x / P;
  const int tl = threadIdx.x % P;
  const int64_t r = g * LINES + line;
  const bool valid = r < total_rows;
  V* Lp = lds + (size_t)line * Cfg::LINE;
  const V* twp = tw;
  if constexpr (Cfg::TW_LDS) {
    V* twl = lds + (size_t)LINES * Cfg::LINE;
    for (int t = threadIdx.x; t < L; t += Cfg::WG) twl[t] = tw[t];
    __syncthreads();
    twp = twl;
  }
  const VI* src = in + (valid ? r : 0) * (int64_t)n;
  V v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int j = tl + q * P;
    V x = {0, 0};
    if (valid && j < n) {
      const VI e = src[j];
      x = cmul(V{(T)e.x, (T)e.y}, A[j]);
    }
    v[q] = x;
  }
  hr_fft<T, L, P>(v, Lp, tl, twp, -1);
#pragma unroll
  for (int q = 0; q < E; ++q) v[q] = cmul(v[q], H[tl + q * P]);
  hr_fft<T, L, P>(v, Lp, tl, twp, +1);
  if (valid) {
    const V alpha = {alpha_re, alpha_im};
    const int64_t b = r / rows;
    const int64_t y = r - b * rows;
    VO* dst = out + b * (int64_t)m * rows + y;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int k = tl + q * P;
      if (k < m) {
        const V o = cmul(v[q], cmul(C[k], alpha));
        dst[(int64_t)k * rows] = VO{(TO)o.x, (TO)o.y};
      }
    }
  }
}

struct CztAxis {
  int n = 0, m = 0, L = 0;
  // FP64 chirps and filter, and the FP32 copies used by complex64 plans
  double2* A64 = nullptr;
  double2* H64 = nullptr;
  double2* C64 = nullptr;
  double2* tw64 = nullptr;
  float2* A32 = nullptr;
  float2* H32 = nullptr;
  float2* C32 = nullptr;
  float2* tw32 = nullptr;
};

struct ZoomPlan {
  int ndim = 0, dtype = 0, precision = HCIPY_PREC_FP64, device = 0;
  int n_in[2] = {1, 1}, n_out[2] = {1, 1};
  CztAxis ax[2][2];  // [direction][axis]
  void* scratch = nullptr;  // FP64 h of the largest axis (set-up only)
};

int next_pow2(int v) {
  int L = 4;
  while (L < v) L <<= 1;
  return L;
}

size_t czt_lds_bytes(int L, int dtype) {
  int tpl = L / 4 < 256 ? (L / 4 < 1 ? 1 : L / 4) : 256;
  int lines = 256 / tpl;
  return (size_t)lines * L * (dtype == HCIPY_DTYPE_C64 ? 8 : 16);
}

#define HCIPY_CZT_SWITCH(N, CALL) \
  switch (N) {                    \
    case 4: CALL(4); break;       \
    case 8: CALL(8); break;       \
    case 16: CALL(16); break;     \
    case 32: CALL(32); break;     \
    case 64: CALL(64); break;     \
    case 128: CALL(128); break;   \
    case 256: CALL(256); break;   \
    case 512: CALL(512); break;   \
    case 1024: CALL(1024); break; \
    case 2048: CALL(2048); break; \
    case 4096: CALL(4096); break; \
    case 8192: CALL(8192); break; \
    default: break;               \
  }

template <int L>
int launch_filter(const double2* h, const double2* tw64, double2* H64, float2* H32, hipStream_t s) {
  const size_t lds = (size_t)L * 16;
  auto kern = czt_filter_kernel<L>;
  HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  hipLaunchKernelGGL(kern, dim3(1), dim3(256), lds, s, h, tw64, H64, H32);
  return check_hip(hipGetLastError(), "czt_filter_kernel");
}

int axis_setup(CztAxis& a, int n, int m, double d_in, double z_in, double d_out, double z_out, int sign, int fft_n,
               int dtype, double2* h, hipStream_t s) {
  a.n = n;
  a.m = m;
  a.L = next_pow2(n + m - 1);
  const int L = a.L;
  HCIPY_HIP(hipMalloc(&a.A64, (size_t)n * 16));
  HCIPY_HIP(hipMalloc(&a.H64, (size_t)L * 16));
  HCIPY_HIP(hipMalloc(&a.C64, (size_t)m * 16));
  HCIPY_HIP(hipMalloc(&a.tw64, (size_t)L * 16));
  if (dtype == HCIPY_DTYPE_C64) {
    HCIPY_HIP(hipMalloc(&a.A32, (size_t)n * 8));
    HCIPY_HIP(hipMalloc(&a.H32, (size_t)L * 8));
    HCIPY_HIP(hipMalloc(&a.C32, (size_t)m * 8));
    HCIPY_HIP(hipMalloc(&a.tw32, (size_t)L * 8));
  }
  const unsigned gb = (unsigned)ceil_div(L, 256);
  hipLaunchKernelGGL(czt_chirp_kernel, dim3(gb), dim3(256), 0, s, n, m, L, d_in, z_in, d_out, z_out, sign, fft_n,
                     a.A64, h, a.C64);
  HCIPY_HIP(hipGetLastError());
  hipLaunchKernelGGL(czt_twiddle_kernel, dim3(gb), dim3(256), 0, s, L, a.tw64, a.tw32);
  HCIPY_HIP(hipGetLastError());
  int st = kOk;
#define FILT(NN) st = launch_filter<NN>(h, a.tw64, a.H64, a.H32, s)
  HCIPY_CZT_SWITCH(L, FILT)
#undef FILT
  if (st) return st;
  if (dtype == HCIPY_DTYPE_C64) {
    hipLaunchKernelGGL(cast_c128_to_c64_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, a.A64, a.A32, n);
    hipLaunchKernelGGL(cast_c128_to_c64_kernel, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, a.C64, a.C32, m);
    HCIPY_HIP(hipGetLastError());
  }
  return kOk;
}

void axis_free(CztAxis& a) {
  void* ptrs[] = {a.A64, a.H64, a.C64, a.tw64, a.A32, a.H32, a.C32, a.tw32};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  a = CztAxis();
}

template <typename TI, typename TO, typename T, int L>
int launch_pass(const CztAxis& a, const void* in, void* out, int rows, int64_t total_rows, double are, double aim,
                hipStream_t s) {
  using V = typename Cplx<T>::V;
  using VI = typename Cplx<TI>::V;
  using VO = typename Cplx<TO>::V;
  using Cfg = HrCfg<T, L>;
  constexpr int LINES = Cfg::LINES;
  const size_t lds = Cfg::LDS_BYTES;
  auto kern = czt_rows_kernel<TI, TO, T, L>;
  if (lds > 65536)
    HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
  const int64_t groups = ceil_div(total_rows, LINES);
  const int64_t blocks = ceil_div(groups, 8) * 8;
  if (blocks > 0x7fffffff) return kOutOfRange;
  const bool f32 = sizeof(T) == 4;
  const V* A = f32 ? (const V*)a.A32 : (const V*)a.A64;
  const V* H = f32 ? (const V*)a.H32 : (const V*)a.H64;
  const V* C = f32 ? (const V*)a.C32 : (const V*)a.C64;
  const V* tw = f32 ? (const V*)a.tw32 : (const V*)a.tw64;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(Cfg::WG), lds, s, (const VI*)in, (VO*)out, A, H, C, tw,
                     a.n, a.m,
                     rows, total_rows, groups, (T)are, (T)aim);
  return check_hip(hipGetLastError(), "czt_rows_kernel");
}

template <typename TI, typename TO, typename T>
int run_pass(const CztAxis& a, const void* in, void* out, int rows, int64_t total_rows, double are, double aim,
             hipStream_t s) {
  int st = kUnsupported;
#define PASS(NN) st = launch_pass<TI, TO, T, NN>(a, in, out, rows, total_rows, are, aim, s)
  HCIPY_CZT_SWITCH(a.L, PASS)
#undef PASS
  return st;
}

}  // namespace
}  // namespace hcipy

using namespace hcipy;

struct hcipy_zoom_plan_s {
  ZoomPlan p;
};

extern "C" {

static int zoom_plan_create(hcipy_zoom_plan* plan, int ndim, const int* n_in, const double* delta_in,
                            const double* zero_in, const int* n_out, const double* delta_out, const double* zero_out,
                            const int* fft_n, int dtype, int precision, void* stream) {
  HCIPY_REQUIRE(plan != nullptr, "plan pointer is NULL");
  *plan = nullptr;
  HCIPY_REQUIRE(ndim == 1 || ndim == 2, "the zoom FFT supports one- and two-dimensional grids");
  HCIPY_REQUIRE(dtype == HCIPY_DTYPE_C64 || dtype == HCIPY_DTYPE_C128, "zoom FFT dtype must be complex");
  HCIPY_REQUIRE(precision == HCIPY_PREC_FP64 || (precision == HCIPY_PREC_FP32 && dtype == HCIPY_DTYPE_C64),
                "zoom FFT precision must be HCIPY_PREC_FP64, or HCIPY_PREC_FP32 for complex64 fields");
  HCIPY_REQUIRE(n_in && delta_in && zero_in && n_out && delta_out && zero_out, "NULL grid argument");
  for (int d = 0; d < ndim; ++d) {
    HCIPY_REQUIRE(n_in[d] >= 1 && n_out[d] >= 1, "grid sizes must be positive");
    HCIPY_REQUIRE(n_in[d] + n_out[d] - 1 <= kCztMaxL,
                  "axis %d: n_in + n_out - 1 = %d exceeds the %d-point line FFT of the zoom FFT", d,
                  n_in[d] + n_out[d] - 1, kCztMaxL);
    if (fft_n) {
      HCIPY_REQUIRE(fft_n[d] >= 1, "axis %d: fft_n must be positive", d);
      const double r = delta_in[d] * delta_out[d] * (double)fft_n[d] / (2.0 * M_PI);
      HCIPY_REQUIRE(std::fabs(r - 1.0) < 1e-9, "axis %d: delta_in * delta_out = %.17g is not 2 pi / %d", d,
                    delta_in[d] * delta_out[d], fft_n[d]);
    }
  }
  if (int st0 = refuse_under_capture(static_cast<hipStream_t>(stream), "hcipy_zoom_plan_create")) return st0;
  hcipy_zoom_plan_s* hp = new (std::nothrow) hcipy_zoom_plan_s();
  HCIPY_REQUIRE(hp != nullptr, "out of host memory");
  ZoomPlan& p = hp->p;
  p.ndim = ndim;
  p.dtype = dtype;
  p.precision = precision;
  for (int d = 0; d < ndim; ++d) {
    p.n_in[d] = n_in[d];
    p.n_out[d] = n_out[d];
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  int st = check_hip(hipGetDevice(&p.device), "hipGetDevice");
  if (!st) st = check_hip(hipMalloc(&p.scratch, (size_t)kCztMaxL * 16), "hipMalloc(scratch)");
  for (int d = 0; d < ndim && !st; ++d) {
    // forward: input grid -> output grid, exp(-i p q); backward: output grid -> input grid, exp(+i p q)
    const int tdt = precision == HCIPY_PREC_FP32 ? HCIPY_DTYPE_C64 : HCIPY_DTYPE_C128;  // table precision
    const int fn = fft_n ? fft_n[d] : 0;
    st = axis_setup(p.ax[0][d], n_in[d], n_out[d], delta_in[d], zero_in[d], delta_out[d], zero_out[d], -1, fn, tdt,
                    static_cast<double2*>(p.scratch), s);
    // the FP64 h buffer is reused: stream order keeps each filter kernel before the next chirp kernel
    if (!st)
      st = axis_setup(p.ax[1][d], n_out[d], n_in[d], delta_out[d], zero_out[d], delta_in[d], zero_in[d], +1, fn, tdt,
                      static_cast<double2*>(p.scratch), s);
  }
  if (st) {
    hcipy_zoom_plan_destroy(hp);
    return st;
  }
  *plan = hp;
  return kOk;
}

int hcipy_zoom_plan_create(hcipy_zoom_plan* plan, int ndim, const int* n_in, const double* delta_in,
                           const double* zero_in, const int* n_out, const double* delta_out, const double* zero_out,
                           int dtype, int precision, void* stream) {
  return zoom_plan_create(plan, ndim, n_in, delta_in, zero_in, n_out, delta_out, zero_out, nullptr, dtype, precision,
                          stream);
}

int hcipy_zoom_plan_create_fft(hcipy_zoom_plan* plan, int ndim, const int* n_in, const double* delta_in,
                               const double* zero_in, const int* n_out, const double* delta_out,
                               const double* zero_out, const int* fft_n, int dtype, int precision, void* stream) {
  HCIPY_REQUIRE(fft_n != nullptr, "NULL fft_n");
  return zoom_plan_create(plan, ndim, n_in, delta_in, zero_in, n_out, delta_out, zero_out, fft_n, dtype, precision,
                          stream);
}

int hcipy_zoom_plan_destroy(hcipy_zoom_plan plan) {
  if (plan == nullptr) return kOk;
  for (auto& dir : plan->p.ax)
    for (auto& a : dir) axis_free(a);
  if (plan->p.scratch) (void)hipFree(plan->p.scratch);
  delete plan;
  return kOk;
}

int hcipy_zoom_workspace_size(hcipy_zoom_plan plan, int64_t batch, size_t* bytes) {
  HCIPY_REQUIRE(plan != nullptr && bytes != nullptr, "NULL argument");
  HCIPY_REQUIRE(batch >= 0, "negative batch");
  const ZoomPlan& p = plan->p;
  if (p.ndim == 1) {
    *bytes = 0;
    return kOk;
  }
  const int64_t fwd = (int64_t)p.n_out[0] * p.n_in[1];  // [mx][ny]
  const int64_t bwd = (int64_t)p.n_in[0] * p.n_out[1];  // [nx][my]
  // the intermediate is held in the arithmetic precision
  *bytes = (size_t)(batch * std::max(fwd, bwd) * (p.precision == HCIPY_PREC_FP32 ? 8 : 16));
  return kOk;
}

int hcipy_zoom_execute(hcipy_zoom_plan plan, int direction, const void* in, void* out, int64_t batch, double alpha_re,
                       double alpha_im, void* workspace, size_t workspace_bytes, void* stream) {
  HCIPY_REQUIRE(plan != nullptr, "NULL plan");
  HCIPY_REQUIRE(direction == 0 || direction == 1, "direction must be 0 (forward) or 1 (backward)");
  HCIPY_REQUIRE(batch >= 0, "negative batch");
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(in != nullptr && out != nullptr && in != out, "NULL or aliased field pointers");
  const ZoomPlan& p = plan->p;
  int dev = 0;
  HCIPY_HIP(hipGetDevice(&dev));
  HCIPY_REQUIRE(dev == p.device, "plan was created on device %d, current device is %d", p.device, dev);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // kind 0: complex64 in FP32, 1: complex64 I/O with FP64 arithmetic, 2: complex128
  const int kind = p.dtype == HCIPY_DTYPE_C128 ? 2 : (p.precision == HCIPY_PREC_FP32 ? 0 : 1);
  const CztAxis* ax = p.ax[direction];
  if (p.ndim == 1) {
    if (kind == 0) return run_pass<float, float, float>(ax[0], in, out, 1, batch, alpha_re, alpha_im, s);
    if (kind == 1) return run_pass<float, float, double>(ax[0], in, out, 1, batch, alpha_re, alpha_im, s);
    return run_pass<double, double, double>(ax[0], in, out, 1, batch, alpha_re, alpha_im, s);
  }
  size_t need = 0;
  HCIPY_REQUIRE(hcipy_zoom_workspace_size(plan, batch, &need) == kOk, "workspace size");
  HCIPY_REQUIRE(workspace != nullptr && workspace_bytes >= need, "workspace of %zu bytes needed, %zu given", need,
                workspace_bytes);
  // axis 0 (x / u) over the rows of the field, then axis 1 over the rows of the transposed result
  const int rows0 = ax[1].n;  // rows of the input field: ny forward, my backward
  const int rows1 = ax[0].m;  // rows after the first pass: mx forward, nx backward
  int st;
  if (kind == 0) {
    st = run_pass<float, float, float>(ax[0], in, workspace, rows0, batch * rows0, 1.0, 0.0, s);
    if (!st) st = run_pass<float, float, float>(ax[1], workspace, out, rows1, batch * rows1, alpha_re, alpha_im, s);
  } else if (kind == 1) {
    st = run_pass<float, double, double>(ax[0], in, workspace, rows0, batch * rows0, 1.0, 0.0, s);
    if (!st) st = run_pass<double, float, double>(ax[1], workspace, out, rows1, batch * rows1, alpha_re, alpha_im, s);
  } else {
    st = run_pass<double, double, double>(ax[0], in, workspace, rows0, batch * rows0, 1.0, 0.0, s);
    if (!st) st = run_pass<double, double, double>(ax[1], workspace, out, rows1, batch * rows1, alpha_re, alpha_im, s);
  }
  return st;
}

}  // extern "C"
