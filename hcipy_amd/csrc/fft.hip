// FourierFilter on power-of-two internal grids: zero-pad -> 2-D FFT -> transfer function ->
// 2-D inverse FFT -> crop, in three passes over HBM (reference: hcipy/fourier/fourier_operations.py:89-151).
//
//   pass A  rows   : load one input row into the centre of a zeroed line of length NX, forward FFT,
//                    write the line (ny rows x NX, column-blocked: see blk()).
//   pass B  columns: load C adjacent columns of pass A's output (contiguous in the blocked
//                    layout; zero outside the centred ny rows), forward FFT, multiply by the
//                    transfer function (conjugated for the adjoint), inverse FFT, write back only
//                    the centred ny rows (the crop along y happens here).
//   pass C  rows   : inverse FFT of each of the ny rows, write the centred nx samples scaled by
//                    1 / (NX NY) (numpy.fft.ifftn normalisation) -- the crop along x.
//
// Line FFTs run on the register-resident Stockham engine of line_fft.hpp (radix-16 stages in
// registers, one LDS round trip per stage); twiddles come from an FP64-generated table
// exp(-2 pi i t / N).
#include "common.hpp"
#include "line_fft.hpp"
#include "../../include/hcipy_amd.h"

#include <algorithm>
#include <new>

namespace hcipy {

namespace {

// Column-blocked layout of the intermediates and the transfer function: columns are grouped in
// blocks of CB = min(8, NX) and each block is stored as a contiguous [rows][CB] array, so that
// the column pass reads and writes whole contiguous blocks and a row pass that handles two
// adjacent rows touches full 128-byte lines (2 x 8 complex64, or 8 complex128).
__host__ __device__ inline int64_t blk(int64_t row, int col, int rows, int NX) {
  const int cb = NX < 8 ? NX : 8;
  return ((int64_t)(col / cb) * rows + row) * cb + col % cb;
}

// Line FFTs run on the register-resident engine (hr_fft): E points per thread, P = N / E
// threads per line, one LDS round trip per radix-E stage.  Lines of one workgroup are
// LINE = hr_lds_points(N) + 1 slots apart (the odd slot staggers their banks).
template <typename T, int N>
struct RowCfg {
  static constexpr int E = N >= 16 ? 16 : N;
  static constexpr int P = N / E;
  static constexpr int WG = P > 256 ? P : 256;
  static constexpr int LINES = WG / P;
  static constexpr int LINE = hr_lds_points(N) + 1;
  static constexpr bool TW_LDS = sizeof(T) == 4;
  static constexpr size_t LDS_BYTES = ((size_t)LINES * LINE + (TW_LDS ? N : 0)) * 2 * sizeof(T);
  static_assert(LDS_BYTES <= 160 * 1024, "row pass exceeds the 160 KiB of LDS per CU");
};

template <typename T, int N, int WG>
__device__ inline const typename Cplx<T>::V* stage_twiddles(const typename Cplx<T>::V* __restrict__ tw,
                                                            typename Cplx<T>::V* lds_tw, bool to_lds) {
  if (!to_lds) return tw;
  for (int t = threadIdx.x; t < N; t += WG) lds_tw[t] = tw[t];
  __syncthreads();
  return lds_tw;
}

// pass A: forward row FFTs of the input rows, centred in zeroed lines of NX points.
template <typename T, int NX>
__global__ __launch_bounds__((RowCfg<T, NX>::WG)) void rows_forward_kernel(const typename Cplx<T>::V* __restrict__ in,
                                                                           typename Cplx<T>::V* __restrict__ w1,
                                                                           const typename Cplx<T>::V* __restrict__ twx,
                                                                           int nx, int ny, int c0x) {
  using V = typename Cplx<T>::V;
  using Cfg = RowCfg<T, NX>;
  constexpr int E = Cfg::E;
  constexpr int P = Cfg::P;
  extern __shared__ __align__(16) unsigned char fft_lds_raw[];
  V* lds = reinterpret_cast<V*>(fft_lds_raw);
  const V* tw = stage_twiddles<T, NX, Cfg::WG>(twx, lds + Cfg::LINES * Cfg::LINE, Cfg::TW_LDS);
  const int line = threadIdx.x / P;
  const int tl = threadIdx.x % P;
  const int y = blockIdx.x * Cfg::LINES + line;
  const int64_t b = blockIdx.y;
  const bool valid = y < ny;
  const V* src = in + (b * ny + (valid ? y : 0)) * (int64_t)nx;
  V v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int xi = tl + q * P - c0x;
    v[q] = (valid && xi >= 0 && xi < nx) ? src[xi] : V{0, 0};
  }
  hr_fft<T, NX, P>(v, lds + line * Cfg::LINE, tl, tw, -1);
  if (valid) {
    V* o = w1 + b * ny * (int64_t)NX;
#pragma unroll
    for (int q = 0; q < E; ++q) o[blk(y, tl + q * P, ny, NX)] = v[q];
  }
}

// pass C: inverse row FFTs, crop to the centred nx samples, scale.
template <typename T, int NX>
__global__ __launch_bounds__((RowCfg<T, NX>::WG)) void rows_inverse_kernel(const typename Cplx<T>::V* __restrict__ w2,
                                                                           typename Cplx<T>::V* __restrict__ out,
                                                                           const typename Cplx<T>::V* __restrict__ twx,
                                                                           int nx, int ny, int c0x, T scale) {
  using V = typename Cplx<T>::V;
  using Cfg = RowCfg<T, NX>;
  constexpr int E = Cfg::E;
  constexpr int P = Cfg::P;
  extern __shared__ __align__(16) unsigned char fft_lds_raw[];
  V* lds = reinterpret_cast<V*>(fft_lds_raw);
  const V* tw = stage_twiddles<T, NX, Cfg::WG>(twx, lds + Cfg::LINES * Cfg::LINE, Cfg::TW_LDS);
  const int line = threadIdx.x / P;
  const int tl = threadIdx.x % P;
  const int y = blockIdx.x * Cfg::LINES + line;
  const int64_t b = blockIdx.y;
  const bool valid = y < ny;
  const V* src = w2 + b * ny * (int64_t)NX;
  V v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) v[q] = valid ? src[blk(y, tl + q * P, ny, NX)] : V{0, 0};
  hr_fft<T, NX, P>(v, lds + line * Cfg::LINE, tl, tw, +1);
  if (valid) {
    V* o = out + (b * ny + y) * (int64_t)nx;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int xi = tl + q * P - c0x;
      if (xi >= 0 && xi < nx) o[xi] = V{v[q].x * scale, v[q].y * scale};
    }
  }
}

template <typename T, int NY>
struct ColCfg {
  // 16 points per thread (one LDS round trip fewer than 8), with at most 512 threads per
  // workgroup so that they fit the register budget unspilled (measured faster than 8 points
  // with 1024 threads)
  static constexpr int E = NY >= 16 ? 16 : NY;
  static constexpr int P = NY / E;
  static constexpr int LINE = hr_lds_points(NY) + 1;
  // columns per workgroup: as many lines as fit in 70 KiB of LDS (two workgroups per CU), at most
  // 16, and at most 512 threads.  Twiddles are read from global memory: an LDS copy would take
  // the room of the second workgroup, which was measured faster.
  static constexpr int CF = (int)(71680 / ((size_t)LINE * 2 * sizeof(T)));
  static constexpr int CT = 512 / P > 0 ? 512 / P : 1;
  static constexpr int C0 = CF < CT ? CF : CT;
  static constexpr int C = C0 > 16 ? 16 : (C0 < 1 ? 1 : C0);
  static constexpr int NT = C * P;
  static constexpr bool TW_LDS = false;
  static constexpr size_t LDS_BYTES = (size_t)C * LINE * 2 * sizeof(T);
  static_assert(LDS_BYTES <= 160 * 1024, "column pass exceeds the 160 KiB of LDS per CU");
};

// pass B: column FFT, transfer function, inverse column FFT, crop to the centred ny rows.  C
// adjacent columns per workgroup; thread (c, tl) = (tid % C, tid / C), so that the threads of a
// wave load and store C consecutive samples of each row they touch.
template <typename T, int NY>
__global__ __launch_bounds__((ColCfg<T, NY>::NT)) void columns_kernel(const typename Cplx<T>::V* __restrict__ w1,
                                                                      typename Cplx<T>::V* __restrict__ w2,
                                                                      const typename Cplx<T>::V* __restrict__ tf,
                                                                      const typename Cplx<T>::V* __restrict__ twy,
                                                                      int NX, int ny, int c0y, int conj_tf) {
  using V = typename Cplx<T>::V;
  using Cf = ColCfg<T, NY>;
  constexpr int C = Cf::C;
  constexpr int E = Cf::E;
  constexpr int P = Cf::P;
  extern __shared__ __align__(16) unsigned char fft_lds_raw[];
  V* lds = reinterpret_cast<V*>(fft_lds_raw);
  const V* tw = stage_twiddles<T, NY, Cf::NT>(twy, lds + C * Cf::LINE, Cf::TW_LDS);
  const int c = threadIdx.x % C;
  const int tl = threadIdx.x / C;
  const int kx = blockIdx.x * C + c;
  const int64_t b = blockIdx.y;
  const bool col = kx < NX;
  V v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int yi = tl + q * P - c0y;
    v[q] = (col && yi >= 0 && yi < ny) ? w1[b * ny * (int64_t)NX + blk(yi, kx, ny, NX)] : V{0, 0};
  }
  V* line = lds + c * Cf::LINE;
  hr_fft<T, NY, P>(v, line, tl, tw, -1);
  // transfer function (already in FFT order, column-blocked)
#pragma unroll
  for (int q = 0; q < E; ++q) {
    V t = col ? tf[blk(tl + q * P, kx, NY, NX)] : V{0, 0};
    if (conj_tf) t.y = -t.y;
    v[q] = cmul(v[q], t);
  }
  hr_fft<T, NY, P>(v, line, tl, tw, +1);
  if (col) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int yi = tl + q * P - c0y;
      if (yi >= 0 && yi < ny) w2[b * ny * (int64_t)NX + blk(yi, kx, ny, NX)] = v[q];
    }
  }
}

// pass B for k transfer functions applied to the same input: the column FFT once, then per
// transfer function t: multiply, inverse FFT, crop, write to w2s.p[t] (same layout as w2).
constexpr int kMaxMultiTf = 8;
struct TfList {
  const void* tf[kMaxMultiTf];
  void* w2[kMaxMultiTf];
};

template <typename T, int NY>
__global__ __launch_bounds__((ColCfg<T, NY>::NT)) void columns_multi_kernel(const typename Cplx<T>::V* __restrict__ w1,
                                                                            TfList l, int k,
                                                                            const typename Cplx<T>::V* __restrict__ twy,
                                                                            int NX, int ny, int c0y, int conj_tf) {
  using V = typename Cplx<T>::V;
  using Cf = ColCfg<T, NY>;
  constexpr int C = Cf::C;
  constexpr int E = Cf::E;
  constexpr int P = Cf::P;
  extern __shared__ __align__(16) unsigned char fft_lds_raw[];
  V* lds = reinterpret_cast<V*>(fft_lds_raw);
  const V* tw = stage_twiddles<T, NY, Cf::NT>(twy, lds + C * Cf::LINE, Cf::TW_LDS);
  const int c = threadIdx.x % C;
  const int tl = threadIdx.x / C;
  const int kx = blockIdx.x * C + c;
  const int64_t b = blockIdx.y;
  const bool col = kx < NX;
  V v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int yi = tl + q * P - c0y;
    v[q] = (col && yi >= 0 && yi < ny) ? w1[b * ny * (int64_t)NX + blk(yi, kx, ny, NX)] : V{0, 0};
  }
  V* line = lds + c * Cf::LINE;
  hr_fft<T, NY, P>(v, line, tl, tw, -1);
  for (int t = 0; t < k; ++t) {
    const V* tf = static_cast<const V*>(l.tf[t]);
    V u[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      V tv = col ? tf[blk(tl + q * P, kx, NY, NX)] : V{0, 0};
      if (conj_tf) tv.y = -tv.y;
      u[q] = cmul(v[q], tv);
    }
    hr_fft<T, NY, P>(u, line, tl, tw, +1);
    if (col) {
      V* w2 = static_cast<V*>(l.w2[t]);
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const int yi = tl + q * P - c0y;
        if (yi >= 0 && yi < ny) w2[b * ny * (int64_t)NX + blk(yi, kx, ny, NX)] = u[q];
      }
    }
  }
}

// pass B of a matrix transfer function (p x q <= 4 x 4): out_i = IFFT(sum_j TF_ij FFT(in_j)).
// Inputs are the Q fields j * batch + b of w1, outputs the p fields i * batch + b of w2; the
// q forward column FFTs are held in registers, so each output costs one multiply-add pass and
// one inverse FFT.  tf[i * q + j] (conjugated with conj_tf).
constexpr int kMaxMimo = 4;
struct TfMatrix {
  const void* tf[kMaxMimo * kMaxMimo];
};

template <typename T, int NY, int Q>
struct MimoCfg {
  // the Q held transforms take Q * E * 2 * sizeof(T) / 4 VGPRs: 16 points per thread while that
  // stays within 128, else 8
  static constexpr int EMAX = Q * sizeof(T) <= 16 ? 16 : 8;
  static constexpr int E = NY >= EMAX ? EMAX : NY;
  static constexpr int P = NY / E;
  static constexpr int LINE = hr_lds_points(NY) + 1;
  static constexpr int CF = (int)(71680 / ((size_t)LINE * 2 * sizeof(T)));
  // the Q held transforms take registers: at most 256 threads (512 for one input), except where
  // one line alone needs more (P = 512 or 1024 at 8192 points): then one line per workgroup.  At
  // 8192 points these kernels spill (36 to 952 bytes per lane in FP32, 628 to 1472 in FP64, most
  // for 3-4 inputs): slower, not wrong.
  static constexpr int CT = (Q == 1 ? 512 : 256) / P > 0 ? (Q == 1 ? 512 : 256) / P : 1;
  static constexpr int C0 = CF < CT ? CF : CT;
  static constexpr int C = C0 > 16 ? 16 : (C0 < 1 ? 1 : C0);
  static constexpr int NT = C * P;
  static constexpr size_t LDS_BYTES = (size_t)C * LINE * 2 * sizeof(T);
  static_assert(LDS_BYTES <= 160 * 1024, "column pass exceeds the 160 KiB of LDS per CU");
};

template <typename T, int NY, int Q>
__global__ __launch_bounds__((MimoCfg<T, NY, Q>::NT)) void columns_mimo_kernel(
    const typename Cplx<T>::V* __restrict__ w1, typename Cplx<T>::V* __restrict__ w2, TfMatrix m, int p, int64_t batch,
    const typename Cplx<T>::V* __restrict__ tw, int NX, int ny, int c0y, int conj_tf) {
  using V = typename Cplx<T>::V;
  using Cf = MimoCfg<T, NY, Q>;
  constexpr int C = Cf::C;
  constexpr int E = Cf::E;
  constexpr int P = Cf::P;
  extern __shared__ __align__(16) unsigned char fft_lds_raw[];
  V* lds = reinterpret_cast<V*>(fft_lds_raw);
  const int c = threadIdx.x % C;
  const int tl = threadIdx.x / C;
  const int kx = blockIdx.x * C + c;
  const int64_t b = blockIdx.y;
  const bool col = kx < NX;
  const int64_t field = (int64_t)ny * NX;
  V* line = lds + c * Cf::LINE;
  V F[Q][E];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const V* src = w1 + (j * batch + b) * field;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int yi = tl + q * P - c0y;
      F[j][q] = (col && yi >= 0 && yi < ny) ? src[blk(yi, kx, ny, NX)] : V{0, 0};
    }
    hr_fft<T, NY, P>(F[j], line, tl, tw, -1);
  }
  for (int i = 0; i < p; ++i) {
    V u[E];
#pragma unroll
    for (int q = 0; q < E; ++q) u[q] = V{0, 0};
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      const V* tf = static_cast<const V*>(m.tf[i * Q + j]);
#pragma unroll
      for (int q = 0; q < E; ++q) {
        V t = col ? tf[blk(tl + q * P, kx, NY, NX)] : V{0, 0};
        if (conj_tf) t.y = -t.y;
        u[q] = cadd(u[q], cmul(F[j][q], t));
      }
    }
    hr_fft<T, NY, P>(u, line, tl, tw, +1);
    if (col) {
      V* dst = w2 + (i * batch + b) * field;
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const int yi = tl + q * P - c0y;
        if (yi >= 0 && yi < ny) dst[blk(yi, kx, ny, NX)] = u[q];
      }
    }
  }
}

__global__ void twiddle_kernel(int n, int dtype, void* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double s, c;
  sincospi(-2.0 * (double)t / (double)n, &s, &c);
  if (dtype == HCIPY_DTYPE_C64)
    static_cast<float2*>(out)[t] = make_float2((float)c, (float)s);
  else
    static_cast<double2*>(out)[t] = make_double2(c, s);
}

// tf_out[ky][kx] = tf_centred[(ky + NY/2) % NY][(kx + NX/2) % NX]  (numpy ifftshift), cast, stored
// column-blocked.
__global__ void ifftshift_cast_kernel(const void* in, int in_dtype, void* out, int out_dtype, int NX, int NY) {
  const int64_t total = (int64_t)NX * NY;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int ky = (int)(i / NX);
    const int kx = (int)(i % NX);
    const int64_t src = (int64_t)((ky + NY / 2) % NY) * NX + (kx + NX / 2) % NX;
    double re, im;
    if (in_dtype == HCIPY_DTYPE_C64) {
      const float2 v = static_cast<const float2*>(in)[src];
      re = v.x;
      im = v.y;
    } else {
      const double2 v = static_cast<const double2*>(in)[src];
      re = v.x;
      im = v.y;
    }
    const int64_t o = blk(ky, kx, NY, NX);
    if (out_dtype == HCIPY_DTYPE_C64)
      static_cast<float2*>(out)[o] = make_float2((float)re, (float)im);
    else
      static_cast<double2*>(out)[o] = make_double2(re, im);
  }
}

// Transfer functions of the near-field propagators on the filter's internal (FFT) grid, averaged
// over oversampling^2 dithered sub-grids (hcipy/field/util.py:606-698).  Coordinates follow the
// reference's FP64 arithmetic: separated coords arange(n) * delta + zero, dither offsets from the
// per-sample spacing (x[i+1] - x[i-1]) / 2 (end samples: the one-sided difference).
//   kind 0: angular spectrum  exp(i sqrt(k^2 - |k_perp|^2 + 0j) z)      (angular_spectrum.py:71-78)
//   kind 1: Fresnel           exp(i k z) exp(-i z / (2k) |k_perp|^2)    (fresnel.py:69-80)
__device__ inline double sep_coord(int i, double delta, double zero) { return (double)i * delta + zero; }
__device__ inline double sep_spacing(int i, int n, double delta, double zero) {
  if (i == 0) return sep_coord(1, delta, zero) - sep_coord(0, delta, zero);
  if (i == n - 1) return sep_coord(n - 1, delta, zero) - sep_coord(n - 2, delta, zero);
  return (sep_coord(i + 1, delta, zero) - sep_coord(i - 1, delta, zero)) / 2.;
}

__global__ void nearfield_tf_kernel(int kind, int NX, int NY, double dx, double zx, double dy, double zy, double k,
                                    double z, int ovs, double pf_re, double pf_im, void* out, int out_dtype,
                                    int centred) {
  const int64_t total = (int64_t)NX * NY;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int iy = (int)(i / NX);
    const int ix = (int)(i % NX);
    const double x0 = sep_coord(ix, dx, zx), y0 = sep_coord(iy, dy, zy);
    const double sx = sep_spacing(ix, NX, dx, zx), sy = sep_spacing(iy, NY, dy, zy);
    double acc_re = 0.0, acc_im = 0.0;
    const double dd = 1.0 / ovs;
    for (int a = 0; a < ovs * ovs; ++a) {
      // dithers = make_uniform_grid(ovs, 1).points, x fastest
      const double ox = (double)(a % ovs) * dd + (-0.5 + dd / 2);
      const double oy = (double)(a / ovs) * dd + (-0.5 + dd / 2);
      const double x = x0 + ox * sx;
      const double y = y0 + oy * sy;
      double re, im;
      if (kind == 0) {
        const double r = hypot(x, y);
        const double a2 = k * k - r * r;
        if (a2 >= 0) {
          const double kz = sqrt(a2);
          double s, c;
          sincos(kz * z, &s, &c);
          re = c;
          im = s;
        } else {
          const double kappa = sqrt(-a2);
          re = exp(-(kappa * z));
          im = 0.0;
        }
      } else {
        const double ph = (-0.5 * z / k) * (x * x + y * y);
        double s, c;
        sincos(ph, &s, &c);
        re = pf_re * c - pf_im * s;
        im = pf_re * s + pf_im * c;
      }
      acc_re += re;
      acc_im += im;
    }
    acc_re /= (double)(ovs * ovs);
    acc_im /= (double)(ovs * ovs);
    // filter plans: FFT order (ifftshift of the centred internal grid), column-blocked;
    // centred: grid order
    const int oxi = (ix + NX - NX / 2) % NX;
    const int oyi = (iy + NY - NY / 2) % NY;
    const int64_t o = centred ? i : blk(oyi, oxi, NY, NX);
    if (out_dtype == HCIPY_DTYPE_C64)
      static_cast<float2*>(out)[o] = make_float2((float)acc_re, (float)acc_im);
    else
      static_cast<double2*>(out)[o] = make_double2(acc_re, acc_im);
  }
}

// Impulse responses of the near-field propagators on a spatial grid, FP64, averaged over
// oversampling^2 dithered sub-grids like the transfer functions above; written as complex128 in
// grid order.  With r^2 = x^2 + y^2 + z^2:
//   kind 0: Rayleigh-Sommerfeld  (z / r) / (2 pi) exp(i k r) (1 / r^2 - i k / r)   (angular_spectrum.py:59-64)
//   kind 1: Fresnel              exp(i k z) / (i lambda z) exp(i k (x^2 + y^2) / (2 z))  (fresnel.py:55-62)
__global__ void nearfield_ir_kernel(int kind, int NX, int NY, double dx, double zx, double dy, double zy, double k,
                                    double z, int ovs, double pf_re, double pf_im, double2* __restrict__ out) {
  const int64_t total = (int64_t)NX * NY;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int iy = (int)(i / NX);
    const int ix = (int)(i % NX);
    const double x0 = sep_coord(ix, dx, zx), y0 = sep_coord(iy, dy, zy);
    const double sx = sep_spacing(ix, NX, dx, zx), sy = sep_spacing(iy, NY, dy, zy);
    double acc_re = 0.0, acc_im = 0.0;
    const double dd = 1.0 / ovs;
    for (int a = 0; a < ovs * ovs; ++a) {
      const double ox = (double)(a % ovs) * dd + (-0.5 + dd / 2);
      const double oy = (double)(a / ovs) * dd + (-0.5 + dd / 2);
      const double x = x0 + ox * sx;
      const double y = y0 + oy * sy;
      double re, im;
      if (kind == 0) {
        const double r2 = x * x + y * y + z * z;
        const double r = sqrt(r2);
        const double c0 = (z / r) / (2.0 * M_PI);
        double s, c;
        sincos(k * r, &s, &c);
        // (c + i s) (1 / r^2 - i k / r)
        const double a_re = 1.0 / r2, a_im = -k / r;
        re = c0 * (c * a_re - s * a_im);
        im = c0 * (c * a_im + s * a_re);
      } else {
        const double ph = (0.5 * k / z) * (x * x + y * y);
        double s, c;
        sincos(ph, &s, &c);
        re = pf_re * c - pf_im * s;
        im = pf_re * s + pf_im * c;
      }
      acc_re += re;
      acc_im += im;
    }
    out[i] = make_double2(acc_re / (double)(ovs * ovs), acc_im / (double)(ovs * ovs));
  }
}

struct FilterPlan {
  int nx, ny, NX, NY, dtype, device;
  int c0x, c0y;
  void* tf = nullptr;
  void* twx = nullptr;
  void* twy = nullptr;
  bool tf_set = false;  // the transfer function is undefined until set (no memset at creation)
};

size_t cbytes(int dtype) { return dtype == HCIPY_DTYPE_C64 ? 8 : 16; }
bool pow2(int n) { return n >= 4 && n <= 8192 && (n & (n - 1)) == 0; }

template <typename T, int N>
int launch_rows_fwd(const FilterPlan& p, const void* in, void* w1, int64_t batch, hipStream_t s) {
  using V = typename Cplx<T>::V;
  using Cfg = RowCfg<T, N>;
  auto kern = rows_forward_kernel<T, N>;
  if (Cfg::LDS_BYTES > 65536)
    HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)Cfg::LDS_BYTES));
  dim3 g((unsigned)ceil_div(p.ny, Cfg::LINES), (unsigned)batch);
  hipLaunchKernelGGL(kern, g, dim3(Cfg::WG), Cfg::LDS_BYTES, s, (const V*)in, (V*)w1, (const V*)p.twx, p.nx, p.ny,
                     p.c0x);
  return kOk;
}
template <typename T, int N>
int launch_rows_inv(const FilterPlan& p, const void* w2, void* out, int64_t batch, hipStream_t s) {
  using V = typename Cplx<T>::V;
  using Cfg = RowCfg<T, N>;
  auto kern = rows_inverse_kernel<T, N>;
  if (Cfg::LDS_BYTES > 65536)
    HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)Cfg::LDS_BYTES));
  dim3 g((unsigned)ceil_div(p.ny, Cfg::LINES), (unsigned)batch);
  const T scale = (T)(1.0 / ((double)p.NX * (double)p.NY));
  hipLaunchKernelGGL(kern, g, dim3(Cfg::WG), Cfg::LDS_BYTES, s, (const V*)w2, (V*)out, (const V*)p.twx, p.nx, p.ny,
                     p.c0x, scale);
  return kOk;
}
template <typename T, int N>
int launch_cols(const FilterPlan& p, const void* w1, void* w2, int conj, int64_t batch, hipStream_t s) {
  using V = typename Cplx<T>::V;
  using Cf = ColCfg<T, N>;
  auto kern = columns_kernel<T, N>;
  if (Cf::LDS_BYTES > 65536)
    HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)Cf::LDS_BYTES));
  dim3 g((unsigned)ceil_div(p.NX, Cf::C), (unsigned)batch);
  hipLaunchKernelGGL(kern, g, dim3(Cf::NT), Cf::LDS_BYTES, s, (const V*)w1, (V*)w2, (const V*)p.tf, (const V*)p.twy,
                     p.NX, p.ny, p.c0y, conj);
  return kOk;
}

template <typename T, int N>
int launch_cols_multi(const FilterPlan& p, const void* w1, const TfList& l, int k, int conj, int64_t batch,
                      hipStream_t s) {
  using V = typename Cplx<T>::V;
  using Cf = ColCfg<T, N>;
  auto kern = columns_multi_kernel<T, N>;
  if (Cf::LDS_BYTES > 65536)
    HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)Cf::LDS_BYTES));
  dim3 g((unsigned)ceil_div(p.NX, Cf::C), (unsigned)batch);
  hipLaunchKernelGGL(kern, g, dim3(Cf::NT), Cf::LDS_BYTES, s, (const V*)w1, l, k, (const V*)p.twy, p.NX, p.ny, p.c0y,
                     conj);
  return kOk;
}

template <typename T, int N, int Q>
int launch_cols_mimo(const FilterPlan& p, const void* w1, void* w2, const TfMatrix& m, int n_out, int64_t batch,
                     int conj, hipStream_t s) {
  using V = typename Cplx<T>::V;
  using Cf = MimoCfg<T, N, Q>;
  auto kern = columns_mimo_kernel<T, N, Q>;
  if (Cf::LDS_BYTES > 65536)
    HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)Cf::LDS_BYTES));
  dim3 g((unsigned)ceil_div(p.NX, Cf::C), (unsigned)batch);
  hipLaunchKernelGGL(kern, g, dim3(Cf::NT), Cf::LDS_BYTES, s, (const V*)w1, (V*)w2, m, n_out, batch, (const V*)p.twy,
                     p.NX, p.ny, p.c0y, conj);
  return kOk;
}

#define HCIPY_POW2_SWITCH(N, CALL) \
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

template <typename T>
int run_filter(const FilterPlan& p, int adjoint, const void* in, void* out, int64_t batch, char* ws, hipStream_t s) {
  const size_t lbytes = (size_t)batch * p.ny * p.NX * cbytes(p.dtype);
  void* w1 = ws;
  void* w2 = ws + round_up((int64_t)lbytes, 256);
  int st = kOk;
#define ROWS_F(n) st = launch_rows_fwd<T, n>(p, in, w1, batch, s)
  HCIPY_POW2_SWITCH(p.NX, ROWS_F)
#undef ROWS_F
  if (st) return st;
  HCIPY_HIP(hipGetLastError());
#define COLS(n) st = launch_cols<T, n>(p, w1, w2, adjoint, batch, s)
  HCIPY_POW2_SWITCH(p.NY, COLS)
#undef COLS
  if (st) return st;
  HCIPY_HIP(hipGetLastError());
#define ROWS_I(n) st = launch_rows_inv<T, n>(p, w2, out, batch, s)
  HCIPY_POW2_SWITCH(p.NX, ROWS_I)
#undef ROWS_I
  if (st) return st;
  return check_hip(hipGetLastError(), "fourier filter launch");
}

// k filters of one input: pass A once into w1, pass B (multi) into the k slices of w2s, pass C
// per output.
template <typename T>
int run_filter_multi(const FilterPlan* const* ps, int k, int adjoint, const void* in, void* const* outs,
                     int64_t batch, char* ws, hipStream_t s) {
  const FilterPlan& p = *ps[0];
  const size_t lbytes = (size_t)round_up((int64_t)((size_t)batch * p.ny * p.NX * cbytes(p.dtype)), 256);
  void* w1 = ws;
  TfList l{};
  for (int t = 0; t < k; ++t) {
    l.tf[t] = ps[t]->tf;
    l.w2[t] = ws + (size_t)(t + 1) * lbytes;
  }
  int st = kOk;
#define ROWS_F(n) st = launch_rows_fwd<T, n>(p, in, w1, batch, s)
  HCIPY_POW2_SWITCH(p.NX, ROWS_F)
#undef ROWS_F
  if (st) return st;
  HCIPY_HIP(hipGetLastError());
#define COLS(n) st = launch_cols_multi<T, n>(p, w1, l, k, adjoint, batch, s)
  HCIPY_POW2_SWITCH(p.NY, COLS)
#undef COLS
  if (st) return st;
  HCIPY_HIP(hipGetLastError());
  for (int t = 0; t < k; ++t) {
#define ROWS_I(n) st = launch_rows_inv<T, n>(p, l.w2[t], outs[t], batch, s)
    HCIPY_POW2_SWITCH(p.NX, ROWS_I)
#undef ROWS_I
    if (st) return st;
  }
  return check_hip(hipGetLastError(), "fourier filter (multi) launch");
}

// matrix filter: pass A over the n_in input fields of each batch item (inputs [n_in][batch]),
// pass B (MIMO) into n_out fields, pass C per output field (outputs [n_out][batch]).
template <typename T, int Q>
int run_filter_mimo_q(const FilterPlan& p, const TfMatrix& m, int n_out, int conj, const void* in, void* out,
                      int64_t batch, char* ws, hipStream_t s) {
  const size_t lbytes = (size_t)round_up((int64_t)((size_t)Q * batch * p.ny * p.NX * cbytes(p.dtype)), 256);
  void* w1 = ws;
  void* w2 = ws + lbytes;
  const size_t fin = (size_t)p.ny * p.nx * cbytes(p.dtype);   // one input/output field
  const size_t fint = (size_t)p.ny * p.NX * cbytes(p.dtype);  // one intermediate field
  int st = kOk;
  // the row passes cover Q * batch (n_out * batch) fields, in grid chunks of 65535
  for (int64_t f0 = 0; f0 < (int64_t)Q * batch; f0 += 65535) {
    const int64_t nf = std::min<int64_t>(65535, (int64_t)Q * batch - f0);
    const void* src = static_cast<const char*>(in) + f0 * fin;
    void* dst = static_cast<char*>(w1) + f0 * fint;
#define ROWS_F(n) st = launch_rows_fwd<T, n>(p, src, dst, nf, s)
    HCIPY_POW2_SWITCH(p.NX, ROWS_F)
#undef ROWS_F
    if (st) return st;
    HCIPY_HIP(hipGetLastError());
  }
#define COLS(n) st = launch_cols_mimo<T, n, Q>(p, w1, w2, m, n_out, batch, conj, s)
  HCIPY_POW2_SWITCH(p.NY, COLS)
#undef COLS
  if (st) return st;
  HCIPY_HIP(hipGetLastError());
  for (int64_t f0 = 0; f0 < (int64_t)n_out * batch; f0 += 65535) {
    const int64_t nf = std::min<int64_t>(65535, (int64_t)n_out * batch - f0);
    const void* src = static_cast<const char*>(w2) + f0 * fint;
    void* dst = static_cast<char*>(out) + f0 * fin;
#define ROWS_I(n) st = launch_rows_inv<T, n>(p, src, dst, nf, s)
    HCIPY_POW2_SWITCH(p.NX, ROWS_I)
#undef ROWS_I
    if (st) return st;
    HCIPY_HIP(hipGetLastError());
  }
  return kOk;
}

template <typename T>
int run_filter_mimo(const FilterPlan& p, const TfMatrix& m, int n_out, int n_in, int conj, const void* in, void* out,
                    int64_t batch, char* ws, hipStream_t s) {
  switch (n_in) {
    case 1: return run_filter_mimo_q<T, 1>(p, m, n_out, conj, in, out, batch, ws, s);
    case 2: return run_filter_mimo_q<T, 2>(p, m, n_out, conj, in, out, batch, ws, s);
    case 3: return run_filter_mimo_q<T, 3>(p, m, n_out, conj, in, out, batch, ws, s);
    default: return run_filter_mimo_q<T, 4>(p, m, n_out, conj, in, out, batch, ws, s);
  }
}

}  // namespace
}  // namespace hcipy

using namespace hcipy;

struct hcipy_filter_plan_s {
  FilterPlan p;
};

extern "C" {

int hcipy_filter_plan_create(hcipy_filter_plan* plan, int nx, int ny, int NX, int NY, int dtype, void* stream) {
  HCIPY_REQUIRE(plan != nullptr, "plan pointer is NULL");
  *plan = nullptr;
  HCIPY_REQUIRE(dtype == HCIPY_DTYPE_C64 || dtype == HCIPY_DTYPE_C128, "filter dtype must be complex");
  HCIPY_REQUIRE(pow2(NX) && pow2(NY), "internal grid %d x %d: FFT kernels need powers of two in [4, 8192]", NX, NY);
  HCIPY_REQUIRE(nx >= 1 && ny >= 1 && nx <= NX && ny <= NY, "input %d x %d does not fit the internal grid", nx, ny);
  if (int st0 = refuse_under_capture(static_cast<hipStream_t>(stream), "hcipy_filter_plan_create")) return st0;
  hcipy_filter_plan_s* h = new (std::nothrow) hcipy_filter_plan_s();
  HCIPY_REQUIRE(h != nullptr, "out of host memory");
  FilterPlan& p = h->p;
  p.nx = nx;
  p.ny = ny;
  p.NX = NX;
  p.NY = NY;
  p.dtype = dtype;
  // FastFourierTransform.cutout_input: int(internal/2) - int(input/2) (fast_fourier_transform.py:262-266)
  p.c0x = NX / 2 - nx / 2;
  p.c0y = NY / 2 - ny / 2;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int st = check_hip(hipGetDevice(&p.device), "hipGetDevice");
  if (!st) st = check_hip(hipMalloc(&p.tf, (size_t)NX * NY * cbytes(dtype)), "hipMalloc(tf)");
  if (!st) st = check_hip(hipMalloc(&p.twx, (size_t)NX * cbytes(dtype)), "hipMalloc(twiddles)");
  if (!st) st = check_hip(hipMalloc(&p.twy, (size_t)NY * cbytes(dtype)), "hipMalloc(twiddles)");
  if (!st) {
    hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)ceil_div(NX, 256)), dim3(256), 0, s, NX, dtype, p.twx);
    hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)ceil_div(NY, 256)), dim3(256), 0, s, NY, dtype, p.twy);
    st = check_hip(hipGetLastError(), "twiddle_kernel");
  }
  if (st) {
    hcipy_filter_plan_destroy(h);
    return st;
  }
  *plan = h;
  return kOk;
}

int hcipy_filter_plan_destroy(hcipy_filter_plan plan) {
  if (plan == nullptr) return kOk;
  if (plan->p.tf) (void)hipFree(plan->p.tf);
  if (plan->p.twx) (void)hipFree(plan->p.twx);
  if (plan->p.twy) (void)hipFree(plan->p.twy);
  delete plan;
  return kOk;
}

int hcipy_filter_set_transfer_function(hcipy_filter_plan plan, const void* tf_centred, int tf_dtype, void* stream) {
  HCIPY_REQUIRE(plan != nullptr && tf_centred != nullptr, "NULL argument");
  HCIPY_REQUIRE(tf_dtype == HCIPY_DTYPE_C64 || tf_dtype == HCIPY_DTYPE_C128, "transfer function must be complex");
  FilterPlan& p = plan->p;
  const int64_t total = (int64_t)p.NX * p.NY;
  hipLaunchKernelGGL(ifftshift_cast_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(total, 256), 65536)), dim3(256),
                     0, static_cast<hipStream_t>(stream), tf_centred, tf_dtype, p.tf, p.dtype, p.NX, p.NY);
  HCIPY_HIP(hipGetLastError());
  p.tf_set = true;
  return kOk;
}

int hcipy_filter_nearfield_transfer_function(hcipy_filter_plan plan, int kind, double delta_x, double zero_x,
                                             double delta_y, double zero_y, double wavenumber, double distance,
                                             int oversampling, void* stream) {
  HCIPY_REQUIRE(plan != nullptr, "NULL plan");
  HCIPY_REQUIRE(kind == 0 || kind == 1, "kind must be 0 (angular spectrum) or 1 (Fresnel)");
  HCIPY_REQUIRE(oversampling >= 1 && oversampling <= 16, "oversampling out of range");
  FilterPlan& p = plan->p;
  // Fresnel phase factor exp(i k z), computed on the host in FP64 as the reference does.
  const double pf_re = cos(wavenumber * distance), pf_im = sin(wavenumber * distance);
  const int64_t total = (int64_t)p.NX * p.NY;
  hipLaunchKernelGGL(nearfield_tf_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(total, 256), 65536)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), kind, p.NX, p.NY, delta_x, zero_x, delta_y, zero_y, wavenumber,
                     distance, oversampling, pf_re, pf_im, p.tf, p.dtype, 0);
  HCIPY_HIP(hipGetLastError());
  p.tf_set = true;
  return kOk;
}

int hcipy_nearfield_transfer_function(int kind, int nx, int ny, double delta_x, double zero_x, double delta_y,
                                      double zero_y, double wavenumber, double distance, int oversampling, void* out,
                                      void* stream) {
  HCIPY_REQUIRE(kind == 0 || kind == 1, "kind must be 0 (angular spectrum) or 1 (Fresnel)");
  HCIPY_REQUIRE(oversampling >= 1 && oversampling <= 16, "oversampling out of range");
  HCIPY_REQUIRE(nx >= 1 && ny >= 1, "grid sizes must be positive");
  HCIPY_REQUIRE(out != nullptr, "NULL output");
  const double pf_re = cos(wavenumber * distance), pf_im = sin(wavenumber * distance);
  const int64_t total = (int64_t)nx * ny;
  hipLaunchKernelGGL(nearfield_tf_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(total, 256), 65536)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), kind, nx, ny, delta_x, zero_x, delta_y, zero_y, wavenumber,
                     distance, oversampling, pf_re, pf_im, out, (int)HCIPY_DTYPE_C128, 1);
  return check_hip(hipGetLastError(), "nearfield_tf_kernel");
}

int hcipy_nearfield_impulse_response(int kind, int nx, int ny, double delta_x, double zero_x, double delta_y,
                                     double zero_y, double wavenumber, double wavelength, double distance,
                                     int oversampling, void* out, void* stream) {
  HCIPY_REQUIRE(kind == 0 || kind == 1, "kind must be 0 (angular spectrum) or 1 (Fresnel)");
  HCIPY_REQUIRE(nx >= 2 && ny >= 2, "grid too small");
  HCIPY_REQUIRE(oversampling >= 1 && oversampling <= 16, "oversampling out of range");
  HCIPY_REQUIRE(distance != 0.0 && wavelength > 0.0, "distance must be non-zero and the wavelength positive");
  HCIPY_REQUIRE(out != nullptr, "NULL output");
  // Fresnel prefactor exp(i k z) / (i lambda z), in FP64 on the host as the reference computes it
  const double e_re = cos(wavenumber * distance), e_im = sin(wavenumber * distance);
  const double d = wavelength * distance;  // exp(i k z) / (i d) = (e_im - i e_re) / d
  const double pf_re = e_im / d, pf_im = -e_re / d;
  const int64_t total = (int64_t)nx * ny;
  hipLaunchKernelGGL(nearfield_ir_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(total, 256), 65536)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), kind, nx, ny, delta_x, zero_x, delta_y, zero_y, wavenumber,
                     distance, oversampling, pf_re, pf_im, static_cast<double2*>(out));
  return check_hip(hipGetLastError(), "nearfield_ir_kernel");
}

int hcipy_filter_workspace_size(hcipy_filter_plan plan, int64_t batch, size_t* bytes) {
  HCIPY_REQUIRE(plan != nullptr && bytes != nullptr && batch >= 0, "bad argument");
  const FilterPlan& p = plan->p;
  const size_t lbytes = (size_t)batch * p.ny * p.NX * cbytes(p.dtype);
  *bytes = 2 * (size_t)round_up((int64_t)lbytes, 256);
  return kOk;
}

int hcipy_filter_execute(hcipy_filter_plan plan, int adjoint, const void* in, void* out, int64_t batch,
                         void* workspace, size_t workspace_bytes, void* stream) {
  HCIPY_REQUIRE(plan != nullptr, "NULL plan");
  HCIPY_REQUIRE(batch >= 0, "batch out of range");
  HCIPY_REQUIRE(plan->p.tf_set, "the filter plan has no transfer function yet");
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(in != nullptr && out != nullptr && workspace != nullptr, "NULL pointer");
  size_t need = 0;
  hcipy_filter_workspace_size(plan, batch, &need);
  HCIPY_REQUIRE(workspace_bytes >= need, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
  int cur = -1;
  HCIPY_HIP(hipGetDevice(&cur));
  HCIPY_REQUIRE(cur == plan->p.device, "plan was created on device %d but device %d is current", plan->p.device, cur);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // fields beyond the 65 535 slices of gridDim.y run as further launches on the same workspace
  const FilterPlan& p = plan->p;
  const size_t field_bytes = (size_t)p.nx * p.ny * cbytes(p.dtype);
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const int64_t nb = std::min<int64_t>(65535, batch - b0);
    const void* ib = static_cast<const char*>(in) + (size_t)b0 * field_bytes;
    void* ob = static_cast<char*>(out) + (size_t)b0 * field_bytes;
    const int st = p.dtype == HCIPY_DTYPE_C64
                       ? run_filter<float>(p, adjoint ? 1 : 0, ib, ob, nb, static_cast<char*>(workspace), s)
                       : run_filter<double>(p, adjoint ? 1 : 0, ib, ob, nb, static_cast<char*>(workspace), s);
    if (st != kOk) return st;
  }
  return kOk;
}

int hcipy_filter_workspace_size_multi(const hcipy_filter_plan* plans, int k, int64_t batch, size_t* bytes) {
  HCIPY_REQUIRE(plans != nullptr && bytes != nullptr && batch >= 0, "bad argument");
  HCIPY_REQUIRE(k >= 1 && k <= kMaxMultiTf, "between 1 and %d filters", kMaxMultiTf);
  HCIPY_REQUIRE(plans[0] != nullptr, "NULL plan");
  const FilterPlan& p = plans[0]->p;
  const size_t lbytes = (size_t)batch * p.ny * p.NX * cbytes(p.dtype);
  *bytes = (size_t)(k + 1) * (size_t)round_up((int64_t)lbytes, 256);
  return kOk;
}

int hcipy_filter_execute_multi(const hcipy_filter_plan* plans, int k, int adjoint, const void* in, void* const* outs,
                               int64_t batch, void* workspace, size_t workspace_bytes, void* stream) {
  HCIPY_REQUIRE(plans != nullptr && outs != nullptr, "NULL argument");
  HCIPY_REQUIRE(k >= 1 && k <= kMaxMultiTf, "between 1 and %d filters", kMaxMultiTf);
  HCIPY_REQUIRE(batch >= 0, "batch out of range");
  const FilterPlan* ps[kMaxMultiTf];
  for (int t = 0; t < k; ++t) {
    HCIPY_REQUIRE(plans[t] != nullptr && outs[t] != nullptr, "NULL plan or output %d", t);
    HCIPY_REQUIRE(plans[t]->p.tf_set, "filter plan %d has no transfer function yet", t);
    ps[t] = &plans[t]->p;
    const FilterPlan& a = *ps[0];
    const FilterPlan& b = *ps[t];
    HCIPY_REQUIRE(a.nx == b.nx && a.ny == b.ny && a.NX == b.NX && a.NY == b.NY && a.dtype == b.dtype &&
                      a.device == b.device,
                  "the plans of one multi-filter call must share grids, dtype and device");
  }
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(in != nullptr && workspace != nullptr, "NULL pointer");
  for (int t = 0; t < k; ++t) HCIPY_REQUIRE(outs[t] != in, "output %d aliases the input", t);
  size_t need = 0;
  hcipy_filter_workspace_size_multi(plans, k, std::min<int64_t>(batch, 65535), &need);
  HCIPY_REQUIRE(workspace_bytes >= need, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
  int cur = -1;
  HCIPY_HIP(hipGetDevice(&cur));
  HCIPY_REQUIRE(cur == ps[0]->device, "plans were created on device %d but device %d is current", ps[0]->device, cur);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const FilterPlan& p = *ps[0];
  const size_t field_bytes = (size_t)p.nx * p.ny * cbytes(p.dtype);
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const int64_t nb = std::min<int64_t>(65535, batch - b0);
    const void* ib = static_cast<const char*>(in) + (size_t)b0 * field_bytes;
    void* ob[kMaxMultiTf];
    for (int t = 0; t < k; ++t) ob[t] = static_cast<char*>(outs[t]) + (size_t)b0 * field_bytes;
    const int st = p.dtype == HCIPY_DTYPE_C64
                       ? run_filter_multi<float>(ps, k, adjoint ? 1 : 0, ib, ob, nb, static_cast<char*>(workspace), s)
                       : run_filter_multi<double>(ps, k, adjoint ? 1 : 0, ib, ob, nb, static_cast<char*>(workspace), s);
    if (st != kOk) return st;
  }
  return kOk;
}

int hcipy_filter_workspace_size_matrix(hcipy_filter_plan plan, int n_out, int n_in, int64_t batch, size_t* bytes) {
  HCIPY_REQUIRE(plan != nullptr && bytes != nullptr && batch >= 0, "bad argument");
  HCIPY_REQUIRE(n_out >= 1 && n_out <= kMaxMimo && n_in >= 1 && n_in <= kMaxMimo, "matrix of at most %d x %d",
                kMaxMimo, kMaxMimo);
  const FilterPlan& p = plan->p;
  const int64_t b = std::min<int64_t>(batch, 65535);
  const size_t fb = (size_t)p.ny * p.NX * cbytes(p.dtype);
  *bytes = (size_t)round_up((int64_t)((size_t)n_in * b * fb), 256) + (size_t)round_up((int64_t)((size_t)n_out * b * fb), 256);
  return kOk;
}

int hcipy_filter_execute_matrix(const hcipy_filter_plan* plans, int n_out, int n_in, int conj, const void* in, void* out,
                                int64_t batch, void* workspace, size_t workspace_bytes, void* stream) {
  HCIPY_REQUIRE(plans != nullptr, "NULL plans");
  HCIPY_REQUIRE(n_out >= 1 && n_out <= kMaxMimo && n_in >= 1 && n_in <= kMaxMimo, "matrix of at most %d x %d",
                kMaxMimo, kMaxMimo);
  HCIPY_REQUIRE(batch >= 0, "batch out of range");
  TfMatrix m{};
  for (int t = 0; t < n_out * n_in; ++t) {
    HCIPY_REQUIRE(plans[t] != nullptr, "NULL plan %d", t);
    HCIPY_REQUIRE(plans[t]->p.tf_set, "filter plan %d has no transfer function yet", t);
    const FilterPlan& a = plans[0]->p;
    const FilterPlan& b = plans[t]->p;
    HCIPY_REQUIRE(a.nx == b.nx && a.ny == b.ny && a.NX == b.NX && a.NY == b.NY && a.dtype == b.dtype &&
                      a.device == b.device,
                  "the plans of one matrix filter must share grids, dtype and device");
    m.tf[t] = b.tf;
  }
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(in != nullptr && out != nullptr && workspace != nullptr && in != out, "NULL or aliased pointer");
  size_t need = 0;
  hcipy_filter_workspace_size_matrix(plans[0], n_out, n_in, batch, &need);
  HCIPY_REQUIRE(workspace_bytes >= need, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
  int cur = -1;
  HCIPY_HIP(hipGetDevice(&cur));
  const FilterPlan& p = plans[0]->p;
  HCIPY_REQUIRE(cur == p.device, "plans were created on device %d but device %d is current", p.device, cur);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // the column pass takes one batch item per grid row; inputs are [n_in][batch] and outputs
  // [n_out][batch], so a larger batch is split by the caller (a chunk is a strided subset)
  HCIPY_REQUIRE(batch <= 65535, "matrix filters take at most 65535 fields per component per call");
  return p.dtype == HCIPY_DTYPE_C64
             ? run_filter_mimo<float>(p, m, n_out, n_in, conj ? 1 : 0, in, out, batch, static_cast<char*>(workspace), s)
             : run_filter_mimo<double>(p, m, n_out, n_in, conj ? 1 : 0, in, out, batch, static_cast<char*>(workspace), s);
}

}  // extern "C"
