// Pointwise and reduction kernels of the propagation chain: apodizers, phase screens, DM phase,
// Babinet subtraction, intensity and FP64 (weighted, masked) power sums.
//
// All kernels are HBM-bound streaming kernels: one pass, coalesced along the flattened field,
// a batch of fields on gridDim.y with the modulating array broadcast (stride 0) or per field.
// Batches beyond the 65 535 slices of gridDim.y run as several launches (field offset b_off).
#include "common.hpp"
#include "../../include/hcipy_amd.h"

namespace hcipy {

namespace {

struct C64 {
  float re, im;
};
struct C128 {
  double re, im;
};

template <typename C>
__device__ inline void load_c(const void* p, int64_t i, double& re, double& im) {
  const C v = static_cast<const C*>(p)[i];
  re = v.re;
  im = v.im;
}

__device__ inline void load_any(const void* p, int dtype, int64_t i, double& re, double& im) {
  switch (dtype) {
    case HCIPY_DTYPE_C64: load_c<C64>(p, i, re, im); break;
    case HCIPY_DTYPE_C128: load_c<C128>(p, i, re, im); break;
    case HCIPY_DTYPE_F32: re = static_cast<const float*>(p)[i]; im = 0.0; break;
    default: re = static_cast<const double*>(p)[i]; im = 0.0; break;
  }
}

template <typename C>
__device__ inline void store_c(void* p, int64_t i, double re, double im);
template <>
__device__ inline void store_c<C64>(void* p, int64_t i, double re, double im) {
  static_cast<C64*>(p)[i] = C64{(float)re, (float)im};
}
template <>
__device__ inline void store_c<C128>(void* p, int64_t i, double re, double im) {
  static_cast<C128*>(p)[i] = C128{re, im};
}

// out = (x - y_opt) * a_opt, with conj(a) if requested.  y and a may be null.
// Arithmetic in FP32 for complex64 fields (the reference's numpy complex64 ufuncs) and FP64
// for complex128 fields.
template <typename C, typename R>
__global__ void sub_mul_kernel(const C* __restrict__ x, const C* __restrict__ y, const void* __restrict__ a,
                               int a_dtype, int64_t a_stride, int conj, C* __restrict__ out, int64_t n, int64_t b_off) {
  const int64_t b = blockIdx.y + b_off;
  const C* xb = x ? x + b * n : nullptr;  // absent: zero
  const C* yb = y ? y + b * n : nullptr;
  C* ob = out + b * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    R re = 0, im = 0;
    if (xb) {
      re = xb[i].re;
      im = xb[i].im;
    }
    if (yb) {
      re -= (R)yb[i].re;
      im -= (R)yb[i].im;
    }
    if (a) {
      double ar, ai;
      load_any(a, a_dtype, b * a_stride + i, ar, ai);
      if (conj) ai = -ai;
      const R r2 = re * (R)ar - im * (R)ai;
      const R i2 = re * (R)ai + im * (R)ar;
      re = r2;
      im = i2;
    }
    ob[i] = C{re, im};
  }
}

// out = x * exp(i * scale * phi), phase evaluated in FP64.
template <typename C, typename R>
__global__ void phase_kernel(const C* __restrict__ x, const void* __restrict__ phi, int phi_dtype, int64_t phi_stride,
                             double scale, const double* __restrict__ row_scale, C* __restrict__ out, int64_t n,
                             int64_t b_off) {
  const int64_t b = blockIdx.y + b_off;
  const C* xb = x + b * n;
  C* ob = out + b * n;
  const double sb = row_scale != nullptr ? scale * row_scale[b] : scale;  // this field's scale
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double p, dummy;
    load_any(phi, phi_dtype, b * phi_stride + i, p, dummy);
    double s, c;
    sincos(sb * p, &s, &c);
    const R re = xb[i].re, im = xb[i].im;
    ob[i] = C{(R)(re * (R)c - im * (R)s), (R)(re * (R)s + im * (R)c)};
  }
}

template <typename C, typename R>
__global__ void intensity_kernel(const C* __restrict__ x, R* __restrict__ out, int64_t total) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const R re = x[i].re, im = x[i].im;
    out[i] = re * re + im * im;
  }
}

// partial[b * nblk + blockIdx.x] = sum_i |x_b[i]|^2 * w_i  (FP64), w from mask (optional) times weight.
template <typename C>
__global__ __launch_bounds__(256) void power_partial_kernel(const C* __restrict__ x, const void* __restrict__ mask,
                                                            int mask_dtype, int64_t mask_stride, double weight,
                                                            double* __restrict__ partial, int64_t n, int64_t b_off) {
  __shared__ double red[4];
  const int64_t b = blockIdx.y + b_off;
  const C* xb = x + b * n;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double re = xb[i].re, im = xb[i].im;
    double w = 1.0, dummy;
    if (mask) load_any(mask, mask_dtype, b * mask_stride + i, w, dummy);
    acc += (re * re + im * im) * w;
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[b * gridDim.x + blockIdx.x] = (red[0] + red[1] + red[2] + red[3]) * weight;
}

__global__ __launch_bounds__(256) void sum_partials_kernel(const double* __restrict__ partial, int nblk,
                                                           double* __restrict__ out) {
  __shared__ double red[4];
  const int64_t b = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 256) acc += partial[b * nblk + i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[b] = red[0] + red[1] + red[2] + red[3];
}

// partial[2 (b * nblk + blockIdx.x) + {0, 1}] = (sum_i x_b[i] [m_i != 0], count of m_i != 0), FP64.
template <typename R>
__global__ __launch_bounds__(256) void masked_partial_kernel(const R* __restrict__ x, const void* __restrict__ mask,
                                                             int mask_dtype, int64_t mask_stride,
                                                             double* __restrict__ partial, int64_t n, int64_t b_off) {
  __shared__ double red[2][4];
  const int64_t b = blockIdx.y + b_off;
  const R* xb = x + b * n;
  double acc = 0.0, cnt = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double m = 1.0, dummy;
    if (mask) load_any(mask, mask_dtype, b * mask_stride + i, m, dummy);
    if (m != 0.0) {
      acc += (double)xb[i];
      cnt += 1.0;
    }
  }
  acc = wave_sum(acc);
  cnt = wave_sum(cnt);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = acc;
    red[1][threadIdx.x >> 6] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int64_t k = 2 * (b * gridDim.x + blockIdx.x);
    partial[k] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    partial[k + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

__global__ __launch_bounds__(256) void mean_partials_kernel(const double* __restrict__ partial, int nblk,
                                                            double* __restrict__ out, double* __restrict__ count) {
  __shared__ double red[2][4];
  const int64_t b = blockIdx.x;
  double acc = 0.0, cnt = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 256) {
    acc += partial[2 * (b * nblk + i)];
    cnt += partial[2 * (b * nblk + i) + 1];
  }
  acc = wave_sum(acc);
  cnt = wave_sum(cnt);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = acc;
    red[1][threadIdx.x >> 6] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double c = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    out[b] = (red[0][0] + red[0][1] + red[0][2] + red[0][3]) / c;  // NaN for an empty region, like numpy
    if (count != nullptr) count[b] = c;
  }
}

// Adjoint of out = x exp(i s phi) with respect to phi: d[i] = s Re(conj(g) i y) = -s Im(conj(g) y),
// summed over the batch when phi is broadcast (phi_stride 0), FP64.
template <typename C>
__global__ void phase_adjoint_kernel(const C* __restrict__ y, const C* __restrict__ g, double scale,
                                     const double* __restrict__ row_scale, double* __restrict__ out,
                                     int64_t phi_stride, int64_t n, int64_t batch) {
  const int64_t nb = (phi_stride == 0) ? 1 : batch;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n * nb; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = k % n;
    const int64_t b0 = (phi_stride == 0) ? 0 : k / n;
    const int64_t b1 = (phi_stride == 0) ? batch : b0 + 1;
    double acc = 0.0;
    for (int64_t b = b0; b < b1; ++b) {
      const C gv = g[b * n + i], yv = y[b * n + i];
      const double im = (double)gv.re * (double)yv.im - (double)gv.im * (double)yv.re;  // Im(conj(g) y)
      acc += row_scale != nullptr ? row_scale[b] * im : im;
    }
    out[k] = -scale * acc;
  }
}

// Adjoint of I = |x|^2: out = 2 g x (g real, broadcast when g_stride is 0).
template <typename C, typename R>
__global__ void intensity_adjoint_kernel(const C* __restrict__ x, const void* __restrict__ g, int g_dtype,
                                         int64_t g_stride, C* __restrict__ out, int64_t n, int64_t b_off) {
  const int64_t b = blockIdx.y + b_off;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double gv, dummy;
    load_any(g, g_dtype, b * g_stride + i, gv, dummy);
    const C v = x[b * n + i];
    out[b * n + i] = C{(R)(2.0 * gv * v.re), (R)(2.0 * gv * v.im)};
  }
}

// Adjoint of the region-of-interest mean: out[b][i] = g[b] / count[b] where mask != 0, else 0.
template <typename R>
__global__ void masked_mean_adjoint_kernel(const void* __restrict__ mask, int mask_dtype, int64_t mask_stride,
                                           const double* __restrict__ g, const double* __restrict__ count,
                                           R* __restrict__ out, int64_t n, int64_t b_off) {
  const int64_t b = blockIdx.y + b_off;
  const double v = g[b] / count[b];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double m = 1.0, dummy;
    if (mask) load_any(mask, mask_dtype, b * mask_stride + i, m, dummy);
    out[b * n + i] = (R)((m != 0.0) ? v : 0.0);
  }
}

// Per-pixel matrix product out[i][k] = sum_j J[i][j] E[j][k] (field_dot of a (p, q) matrix field
// with a (q, K) field; J^H with adjoint: out[i][k] = sum_j conj(J[j][i]) E[j][k]).  Arithmetic
// in FP32 for complex64 fields and FP64 for complex128 fields, J of either complex dtype.
template <typename C, typename R>
__global__ void field_dot_kernel(const void* __restrict__ J, int j_dtype, int p, int q, int adjoint,
                                 const C* __restrict__ E, int64_t K, C* __restrict__ out, int64_t n) {
  const int64_t total = (int64_t)p * K * n;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t px = t % n;
    const int64_t k = (t / n) % K;
    const int i = (int)(t / (n * K));
    R re = 0, im = 0;
    for (int j = 0; j < q; ++j) {
      double jr, ji;
      // adjoint: element (i, j) of J^H is conj(J[j][i]); J is (q x p) then
      const int64_t ji_idx = adjoint ? ((int64_t)j * p + i) : ((int64_t)i * q + j);
      load_any(J, j_dtype, ji_idx * n + px, jr, ji);
      if (adjoint) ji = -ji;
      const C e = E[((int64_t)j * K + k) * n + px];
      re += (R)jr * e.re - (R)ji * e.im;
      im += (R)jr * e.im + (R)ji * e.re;
    }
    out[((int64_t)i * K + k) * n + px] = C{re, im};
  }
}

// Stokes parameters per pixel from the output coherency matrix C = <e e^H> (2 x 2, Hermitian):
//   I = C00 + C11, Q = C00 - C11, U = 2 Re C01, V = -2 Im C01.
// A Jones-vector field e (vector_field != 0, layout [2][n]) has C = e e^H.  A Jones-matrix field
// J = (x, y; z, w) (layout [2][2][n]) acting on the input Stokes vector (s0, s1, s2, s3) has
// C = J S J^H with S = 0.5 (s0 + s1, s2 - i s3; s2 + i s3, s0 - s1), formed as T = J S and
// C_ab = sum_k T_ak conj(J_bk).  component 0..3 writes one of (I, Q, U, V) to out[px];
// component -1 writes all four to out[c * n + px].  FP64 arithmetic.
template <typename C, typename R>
__global__ void stokes_kernel(const C* __restrict__ E, int vector_field, double s0, double s1, double s2, double s3,
                              int component, R* __restrict__ out, int64_t n) {
  const double sp = 0.5 * (s0 + s1), sm = 0.5 * (s0 - s1), su = 0.5 * s2, sv = 0.5 * s3;
  for (int64_t px = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; px < n; px += (int64_t)gridDim.x * blockDim.x) {
    double c00, c11, c01r, c01i;
    if (vector_field) {
      const C ex = E[px], ey = E[n + px];
      c00 = (double)ex.re * ex.re + (double)ex.im * ex.im;
      c11 = (double)ey.re * ey.re + (double)ey.im * ey.im;
      c01r = (double)ex.re * ey.re + (double)ex.im * ey.im;
      c01i = (double)ex.im * ey.re - (double)ex.re * ey.im;
    } else {
      const C x = E[px], y = E[n + px], z = E[2 * n + px], w = E[3 * n + px];
      const double xr = x.re, xi = x.im, yr = y.re, yi = y.im, zr = z.re, zi = z.im, wr = w.re, wi = w.im;
      // T = J S:  T_a0 = J_a0 sp + J_a1 (su + i sv),  T_a1 = J_a0 (su - i sv) + J_a1 sm
      const double t00r = xr * sp + yr * su - yi * sv, t00i = xi * sp + yi * su + yr * sv;
      const double t01r = xr * su + xi * sv + yr * sm, t01i = xi * su - xr * sv + yi * sm;
      const double t10r = zr * sp + wr * su - wi * sv, t10i = zi * sp + wi * su + wr * sv;
      const double t11r = zr * su + zi * sv + wr * sm, t11i = zi * su - zr * sv + wi * sm;
      // C00 = T00 conj(x) + T01 conj(y),  C11 = T10 conj(z) + T11 conj(w),  C01 = T00 conj(z) + T01 conj(w)
      c00 = t00r * xr + t00i * xi + t01r * yr + t01i * yi;
      c11 = t10r * zr + t10i * zi + t11r * wr + t11i * wi;
      c01r = t00r * zr + t00i * zi + t01r * wr + t01i * wi;
      c01i = t00i * zr - t00r * zi + t01i * wr - t01r * wi;
    }
    const double st[4] = {c00 + c11, c00 - c11, 2.0 * c01r, -2.0 * c01i};
    if (component >= 0) {
      out[px] = (R)st[component];
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) out[c * n + px] = (R)st[c];
    }
  }
}

// out = x exp(i (a_x u + a_y v)) on a regular grid, u = zero_x + ix delta_x (x fastest),
// v = zero_y + iy delta_y; the phase in FP64.  Batch b0 + blockIdx.y; out may alias x.
template <typename C, typename R>
__global__ void phase_ramp_kernel(const C* x, int nx, int64_t n, double zero_x, double delta_x, double zero_y,
                                  double delta_y, double ax, double ay, C* out, int64_t b0) {
  const int64_t off = (b0 + blockIdx.y) * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t iy = i / nx;
    const int64_t ix = i - iy * nx;
    const double ph = ax * (zero_x + (double)ix * delta_x) + ay * (zero_y + (double)iy * delta_y);
    double sn, cs;
    sincos(ph, &sn, &cs);
    const C v = x[off + i];
    const double re = (double)v.re * cs - (double)v.im * sn;
    const double im = (double)v.re * sn + (double)v.im * cs;
    out[off + i] = C{(R)re, (R)im};
  }
}

// out = x - sum_j y_j over k <= 16 arrays of n complex values, subtracted in order j = 0, 1, ...
// (the rounding of k chained x - y passes); out may alias x.
constexpr int kMaxSubMany = 16;
struct PtrList {
  const void* p[kMaxSubMany];
};

template <typename C, typename R>
__global__ void sub_many_kernel(const C* x, PtrList ys, int k, C* out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    C a = x[i];
    for (int j = 0; j < k; ++j) {
      const C y = static_cast<const C*>(ys.p[j])[i];
      a.re -= y.re;
      a.im -= y.im;
    }
    out[i] = a;
  }
}

constexpr int kThreads = 256;
constexpr int kPowerBlocks = 256;

constexpr int64_t kMaxGridY = 65535;

dim3 grid_for(int64_t n, int64_t batch) {
  const int64_t bx = std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kThreads), 4096));
  return dim3((unsigned)bx, (unsigned)batch);
}

// Number of fields of the launch starting at field b0.
int64_t chunk(int64_t batch, int64_t b0) { return std::min<int64_t>(kMaxGridY, batch - b0); }

bool is_complex(int d) { return d == HCIPY_DTYPE_C64 || d == HCIPY_DTYPE_C128; }
bool is_valid(int d) { return d >= HCIPY_DTYPE_C64 && d <= HCIPY_DTYPE_F64; }

}  // namespace
}  // namespace hcipy

using namespace hcipy;

extern "C" {

int hcipy_pw_sub_mul(const void* x, const void* y, const void* a, int a_dtype, int64_t a_stride, int conjugate_a,
                     void* out, int field_dtype, int64_t n, int64_t batch, void* stream) {
  HCIPY_REQUIRE(field_dtype == HCIPY_DTYPE_C64 || field_dtype == HCIPY_DTYPE_C128, "field must be complex");
  HCIPY_REQUIRE(a == nullptr || is_valid(a_dtype), "bad dtype of the multiplier");
  HCIPY_REQUIRE(n >= 0 && batch >= 0, "bad size");
  if (n == 0 || batch == 0) return kOk;
  HCIPY_REQUIRE((x != nullptr || y != nullptr) && out != nullptr, "NULL pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int64_t b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const dim3 g = grid_for(n, chunk(batch, b0));
    if (field_dtype == HCIPY_DTYPE_C64)
      hipLaunchKernelGGL((sub_mul_kernel<C64, float>), g, dim3(kThreads), 0, s, (const C64*)x, (const C64*)y, a,
                         a_dtype, a_stride, conjugate_a, (C64*)out, n, b0);
    else
      hipLaunchKernelGGL((sub_mul_kernel<C128, double>), g, dim3(kThreads), 0, s, (const C128*)x, (const C128*)y, a,
                         a_dtype, a_stride, conjugate_a, (C128*)out, n, b0);
    HCIPY_HIP(hipGetLastError());
  }
  return kOk;
}

int hcipy_pw_phase(const void* x, const void* phi, int phi_dtype, int64_t phi_stride, double scale, void* out,
                   int field_dtype, int64_t n, int64_t batch, void* stream) {
  return hcipy_pw_phase_rows(x, phi, phi_dtype, phi_stride, nullptr, scale, out, field_dtype, n, batch, stream);
}

int hcipy_pw_phase_rows(const void* x, const void* phi, int phi_dtype, int64_t phi_stride, const double* row_scale,
                        double scale, void* out, int field_dtype, int64_t n, int64_t batch, void* stream) {
  HCIPY_REQUIRE(field_dtype == HCIPY_DTYPE_C64 || field_dtype == HCIPY_DTYPE_C128, "field must be complex");
  HCIPY_REQUIRE(phi_dtype == HCIPY_DTYPE_F32 || phi_dtype == HCIPY_DTYPE_F64, "phase must be real");
  HCIPY_REQUIRE(n >= 0 && batch >= 0, "bad size");
  if (n == 0 || batch == 0) return kOk;
  HCIPY_REQUIRE(x != nullptr && out != nullptr && phi != nullptr, "NULL pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int64_t b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const dim3 g = grid_for(n, chunk(batch, b0));
    if (field_dtype == HCIPY_DTYPE_C64)
      hipLaunchKernelGGL((phase_kernel<C64, float>), g, dim3(kThreads), 0, s, (const C64*)x, phi, phi_dtype,
                         phi_stride, scale, row_scale, (C64*)out, n, b0);
    else
      hipLaunchKernelGGL((phase_kernel<C128, double>), g, dim3(kThreads), 0, s, (const C128*)x, phi, phi_dtype,
                         phi_stride, scale, row_scale, (C128*)out, n, b0);
    HCIPY_HIP(hipGetLastError());
  }
  return kOk;
}

int hcipy_pw_intensity(const void* x, int field_dtype, void* out, int64_t total, void* stream) {
  HCIPY_REQUIRE(is_complex(field_dtype), "field must be complex");
  HCIPY_REQUIRE(total >= 0, "bad size");
  if (total == 0) return kOk;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 g = grid_for(total, 1);
  if (field_dtype == HCIPY_DTYPE_C64)
    hipLaunchKernelGGL((intensity_kernel<C64, float>), g, dim3(kThreads), 0, s, (const C64*)x, (float*)out, total);
  else
    hipLaunchKernelGGL((intensity_kernel<C128, double>), g, dim3(kThreads), 0, s, (const C128*)x, (double*)out, total);
  return check_hip(hipGetLastError(), "intensity_kernel");
}

int hcipy_pw_power_workspace(int64_t batch, size_t* bytes) {
  HCIPY_REQUIRE(bytes != nullptr && batch >= 0, "bad argument");
  *bytes = (size_t)batch * kPowerBlocks * sizeof(double);
  return kOk;
}

int hcipy_pw_power(const void* x, int field_dtype, const void* mask, int mask_dtype, int64_t mask_stride,
                   double weight, int64_t n, int64_t batch, double* out, void* workspace, size_t workspace_bytes,
                   void* stream) {
  HCIPY_REQUIRE(is_complex(field_dtype), "field must be complex");
  HCIPY_REQUIRE(mask == nullptr || is_valid(mask_dtype), "bad mask dtype");
  HCIPY_REQUIRE(n >= 0 && batch >= 0, "bad size");
  HCIPY_REQUIRE(workspace_bytes >= (size_t)batch * kPowerBlocks * sizeof(double), "workspace too small");
  if (batch == 0) return kOk;
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* partial = static_cast<double*>(workspace);
  for (int64_t b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const dim3 g(kPowerBlocks, (unsigned)chunk(batch, b0));
    if (field_dtype == HCIPY_DTYPE_C64)
      hipLaunchKernelGGL((power_partial_kernel<C64>), g, dim3(kThreads), 0, s, (const C64*)x, mask, mask_dtype,
                         mask_stride, weight, partial, n, b0);
    else
      hipLaunchKernelGGL((power_partial_kernel<C128>), g, dim3(kThreads), 0, s, (const C128*)x, mask, mask_dtype,
                         mask_stride, weight, partial, n, b0);
    HCIPY_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(sum_partials_kernel, dim3((unsigned)batch), dim3(kThreads), 0, s, partial, kPowerBlocks, out);
  return check_hip(hipGetLastError(), "sum_partials_kernel");
}

int hcipy_pw_masked_mean_workspace(int64_t batch, size_t* bytes) {
  HCIPY_REQUIRE(bytes != nullptr && batch >= 0, "bad argument");
  *bytes = (size_t)batch * kPowerBlocks * 2 * sizeof(double);
  return kOk;
}

int hcipy_pw_masked_mean(const void* x, int dtype, const void* mask, int mask_dtype, int64_t mask_stride, int64_t n,
                         int64_t batch, double* out, double* count, void* workspace, size_t workspace_bytes,
                         void* stream) {
  HCIPY_REQUIRE(dtype == HCIPY_DTYPE_F32 || dtype == HCIPY_DTYPE_F64, "values must be real");
  HCIPY_REQUIRE(mask == nullptr || is_valid(mask_dtype), "bad mask dtype");
  HCIPY_REQUIRE(n >= 0 && batch >= 0, "bad size");
  HCIPY_REQUIRE(workspace_bytes >= (size_t)batch * kPowerBlocks * 2 * sizeof(double), "workspace too small");
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(x != nullptr && out != nullptr && workspace != nullptr, "NULL pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* partial = static_cast<double*>(workspace);
  for (int64_t b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const dim3 g(kPowerBlocks, (unsigned)chunk(batch, b0));
    if (dtype == HCIPY_DTYPE_F32)
      hipLaunchKernelGGL(masked_partial_kernel<float>, g, dim3(kThreads), 0, s, (const float*)x, mask, mask_dtype,
                         mask_stride, partial, n, b0);
    else
      hipLaunchKernelGGL(masked_partial_kernel<double>, g, dim3(kThreads), 0, s, (const double*)x, mask, mask_dtype,
                         mask_stride, partial, n, b0);
    HCIPY_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(mean_partials_kernel, dim3((unsigned)batch), dim3(kThreads), 0, s, partial, kPowerBlocks, out,
                     count);
  return check_hip(hipGetLastError(), "mean_partials_kernel");
}

int hcipy_pw_masked_mean_adjoint(const void* mask, int mask_dtype, int64_t mask_stride, const double* g,
                                 const double* count, void* out, int out_dtype, int64_t n, int64_t batch,
                                 void* stream) {
  HCIPY_REQUIRE(out_dtype == HCIPY_DTYPE_F32 || out_dtype == HCIPY_DTYPE_F64, "gradient must be real");
  HCIPY_REQUIRE(mask == nullptr || is_valid(mask_dtype), "bad mask dtype");
  HCIPY_REQUIRE(n >= 0 && batch >= 0, "bad size");
  if (n == 0 || batch == 0) return kOk;
  HCIPY_REQUIRE(g != nullptr && count != nullptr && out != nullptr, "NULL pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int64_t b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const dim3 grid = grid_for(n, chunk(batch, b0));
    if (out_dtype == HCIPY_DTYPE_F32)
      hipLaunchKernelGGL(masked_mean_adjoint_kernel<float>, grid, dim3(kThreads), 0, s, mask, mask_dtype, mask_stride,
                         g, count, (float*)out, n, b0);
    else
      hipLaunchKernelGGL(masked_mean_adjoint_kernel<double>, grid, dim3(kThreads), 0, s, mask, mask_dtype,
                         mask_stride, g, count, (double*)out, n, b0);
    HCIPY_HIP(hipGetLastError());
  }
  return kOk;
}

int hcipy_pw_phase_adjoint(const void* y, const void* g, int field_dtype, double scale, double* out,
                           int64_t phi_stride, int64_t n, int64_t batch, void* stream) {
  return hcipy_pw_phase_rows_adjoint(y, g, field_dtype, nullptr, scale, out, phi_stride, n, batch, stream);
}

int hcipy_pw_phase_rows_adjoint(const void* y, const void* g, int field_dtype, const double* row_scale, double scale,
                                double* out, int64_t phi_stride, int64_t n, int64_t batch, void* stream) {
  HCIPY_REQUIRE(is_complex(field_dtype), "field must be complex");
  HCIPY_REQUIRE(phi_stride == 0 || phi_stride == n, "phase stride must be 0 or n");
  HCIPY_REQUIRE(n >= 0 && batch >= 0, "bad size");
  if (n == 0 || batch == 0) return kOk;
  HCIPY_REQUIRE(y != nullptr && g != nullptr && out != nullptr, "NULL pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t total = n * (phi_stride == 0 ? 1 : batch);
  const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, kThreads), 8192)));
  if (field_dtype == HCIPY_DTYPE_C64)
    hipLaunchKernelGGL(phase_adjoint_kernel<C64>, grid, dim3(kThreads), 0, s, (const C64*)y, (const C64*)g, scale,
                       row_scale, out, phi_stride, n, batch);
  else
    hipLaunchKernelGGL(phase_adjoint_kernel<C128>, grid, dim3(kThreads), 0, s, (const C128*)y, (const C128*)g, scale,
                       row_scale, out, phi_stride, n, batch);
  return check_hip(hipGetLastError(), "phase_adjoint_kernel");
}

int hcipy_pw_intensity_adjoint(const void* x, int field_dtype, const void* g, int g_dtype, int64_t g_stride, void* out,
                               int64_t n, int64_t batch, void* stream) {
  HCIPY_REQUIRE(is_complex(field_dtype), "field must be complex");
  HCIPY_REQUIRE(g_dtype == HCIPY_DTYPE_F32 || g_dtype == HCIPY_DTYPE_F64, "gradient of the intensity must be real");
  HCIPY_REQUIRE(n >= 0 && batch >= 0, "bad size");
  if (n == 0 || batch == 0) return kOk;
  HCIPY_REQUIRE(x != nullptr && g != nullptr && out != nullptr, "NULL pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int64_t b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const dim3 grid = grid_for(n, chunk(batch, b0));
    if (field_dtype == HCIPY_DTYPE_C64)
      hipLaunchKernelGGL((intensity_adjoint_kernel<C64, float>), grid, dim3(kThreads), 0, s, (const C64*)x, g,
                         g_dtype, g_stride, (C64*)out, n, b0);
    else
      hipLaunchKernelGGL((intensity_adjoint_kernel<C128, double>), grid, dim3(kThreads), 0, s, (const C128*)x, g,
                         g_dtype, g_stride, (C128*)out, n, b0);
    HCIPY_HIP(hipGetLastError());
  }
  return kOk;
}

int hcipy_pw_field_dot(const void* J, int j_dtype, int p, int q, int adjoint, const void* E, int field_dtype, int64_t K,
                       void* out, int64_t n, void* stream) {
  HCIPY_REQUIRE(is_complex(field_dtype) && is_complex(j_dtype), "complex fields and matrices required");
  HCIPY_REQUIRE(p >= 1 && q >= 1 && p <= 16 && q <= 16 && K >= 1 && n >= 0, "bad size");
  if (n == 0) return kOk;
  HCIPY_REQUIRE(J != nullptr && E != nullptr && out != nullptr && out != E, "NULL or aliased pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t total = (int64_t)(adjoint ? q : p) * K * n;
  const dim3 g((unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, kThreads), 16384)));
  // the kernel's (p, q) are the output and contracted sizes of the product it forms
  const int po = adjoint ? q : p, qo = adjoint ? p : q;
  if (field_dtype == HCIPY_DTYPE_C64)
    hipLaunchKernelGGL((field_dot_kernel<C64, float>), g, dim3(kThreads), 0, s, J, j_dtype, po, qo, adjoint,
                       (const C64*)E, K, (C64*)out, n);
  else
    hipLaunchKernelGGL((field_dot_kernel<C128, double>), g, dim3(kThreads), 0, s, J, j_dtype, po, qo, adjoint,
                       (const C128*)E, K, (C128*)out, n);
  return check_hip(hipGetLastError(), "field_dot_kernel");
}

int hcipy_pw_stokes(const void* E, int field_dtype, int vector_field, const double* stokes, int component,
                    void* out, int64_t n, void* stream) {
  HCIPY_REQUIRE(is_complex(field_dtype), "field must be complex");
  HCIPY_REQUIRE(vector_field || stokes != nullptr, "a Jones-matrix field needs the input Stokes vector");
  HCIPY_REQUIRE(component >= -1 && component <= 3, "component must be -1 (all) or 0..3 (I, Q, U, V)");
  HCIPY_REQUIRE(n >= 0, "bad size");
  if (n == 0) return kOk;
  HCIPY_REQUIRE(E != nullptr && out != nullptr, "NULL pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double z[4] = {0, 0, 0, 0};
  const double* sv = vector_field ? z : stokes;
  const dim3 g = grid_for(n, 1);
  if (field_dtype == HCIPY_DTYPE_C64)
    hipLaunchKernelGGL((stokes_kernel<C64, float>), g, dim3(kThreads), 0, s, (const C64*)E, vector_field, sv[0], sv[1],
                       sv[2], sv[3], component, (float*)out, n);
  else
    hipLaunchKernelGGL((stokes_kernel<C128, double>), g, dim3(kThreads), 0, s, (const C128*)E, vector_field, sv[0],
                       sv[1], sv[2], sv[3], component, (double*)out, n);
  return check_hip(hipGetLastError(), "stokes_kernel");
}

int hcipy_pw_phase_ramp(const void* x, int field_dtype, int nx, int ny, double zero_x, double delta_x, double zero_y,
                        double delta_y, double ax, double ay, void* out, int64_t batch, void* stream) {
  HCIPY_REQUIRE(is_complex(field_dtype), "field must be complex");
  HCIPY_REQUIRE(nx >= 1 && ny >= 1 && batch >= 0, "bad size");
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(x != nullptr && out != nullptr, "NULL pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t n = (int64_t)nx * ny;
  for (int64_t b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const dim3 g = grid_for(n, chunk(batch, b0));
    if (field_dtype == HCIPY_DTYPE_C64)
      hipLaunchKernelGGL((phase_ramp_kernel<C64, float>), g, dim3(kThreads), 0, s, (const C64*)x, nx, n, zero_x,
                         delta_x, zero_y, delta_y, ax, ay, (C64*)out, b0);
    else
      hipLaunchKernelGGL((phase_ramp_kernel<C128, double>), g, dim3(kThreads), 0, s, (const C128*)x, nx, n, zero_x,
                         delta_x, zero_y, delta_y, ax, ay, (C128*)out, b0);
    HCIPY_HIP(hipGetLastError());
  }
  return kOk;
}

int hcipy_pw_sub_many(const void* x, const void* const* ys, int k, void* out, int field_dtype, int64_t n,
                      void* stream) {
  HCIPY_REQUIRE(is_complex(field_dtype), "field must be complex");
  HCIPY_REQUIRE(k >= 0 && k <= kMaxSubMany, "between 0 and %d subtrahends", kMaxSubMany);
  HCIPY_REQUIRE(n >= 0, "bad size");
  if (n == 0) return kOk;
  HCIPY_REQUIRE(x != nullptr && out != nullptr && (k == 0 || ys != nullptr), "NULL pointer");
  PtrList l{};
  for (int j = 0; j < k; ++j) {
    HCIPY_REQUIRE(ys[j] != nullptr, "NULL subtrahend %d", j);
    l.p[j] = ys[j];
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 g = grid_for(n, 1);
  if (field_dtype == HCIPY_DTYPE_C64)
    hipLaunchKernelGGL((sub_many_kernel<C64, float>), g, dim3(kThreads), 0, s, (const C64*)x, l, k, (C64*)out, n);
  else
    hipLaunchKernelGGL((sub_many_kernel<C128, double>), g, dim3(kThreads), 0, s, (const C128*)x, l, k, (C128*)out, n);
  return check_hip(hipGetLastError(), "sub_many_kernel");
}

}  // extern "C"
