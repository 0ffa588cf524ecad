// Dense FP64 linear algebra of the wavefront-control step: the Tikhonov-regularised inverse of a
// real matrix (hcipy/util/matrix_inversion.py:71-100),
//   M+ = V diag(s / (s^2 + lam^2)) U^T = (M^T M + lam^2 I)^-1 M^T,   lam = rcond * max(s),
// computed without an SVD: the Gram matrix M^T M on the FP64 MFMA stage kernel of the Fourier
// transforms, its largest eigenvalue (max(s)^2) by power iteration, a blocked Cholesky factorisation
// of M^T M + lam^2 I and two blocked triangular solves with M^T as right-hand sides.
//
// Gram matrix on the complex stage kernel: with the columns of M as rows of M^T, pairs of
// consecutive entries of a row read as one complex number, z_a = M[2p][a] + i M[2p+1][a], so
//   Re(sum_p conj(z_a[p]) z_b[p]) = sum_p M[2p][a] M[2p][b] + M[2p+1][a] M[2p+1][b] = (M^T M)[a][b].
// The stage kernel contracts compact complex rows A with data rows B (C = A . B^T in complex
// arithmetic), so A holds conj(z) and B holds z, both zero-padded to whole k-tiles.
//
// All kernels here are plain FP64 (vector FMA, LDS tiles): the factorisation and the solves are
// O(n^3 + n^2 m) with n the actuator count (about 10^3), small next to the Jacobian that feeds them.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "mft_gemm.hpp"
#include "../../include/hcipy_amd.h"

namespace hcipy {

int launch_stage(int prec, const StageArgs& a, int nz, int b_mode, hipStream_t s);

namespace {

constexpr int kNb = 64;  // block size of the factorisation and the solves

// Mt[a][k] (n rows, m reals) -> A[a][k] = conj of the complex pairs, B[a][k] = the pairs,
// both with rows of ld reals (zero beyond m) and A padded with zero rows up to rows_a.
__global__ __launch_bounds__(256) void gram_operands_kernel(const double* __restrict__ M, int64_t m, int64_t n,
                                                            double* __restrict__ A, double* __restrict__ B, int64_t ld,
                                                            int64_t rows_a) {
  const int64_t total = rows_a * ld;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = idx / ld, k = idx - a * ld;
    // M is row-major m x n: column a, row k
    const double v = (a < n && k < m) ? M[k * n + a] : 0.0;
    A[idx] = (k & 1) ? -v : v;
    if (a < n) B[idx] = v;
  }
}

// S[i][j] = Re(C[i][j]) (+ shift on the diagonal)
__global__ __launch_bounds__(256) void real_part_kernel(const double2* __restrict__ C, double* __restrict__ S,
                                                        int64_t n, double shift) {
  const int64_t total = n * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / n, j = idx - i * n;
    S[idx] = C[idx].x + (i == j ? shift : 0.0);
  }
}

// y = S x, one wave per row (fixed summation order)
__global__ __launch_bounds__(256) void symv_kernel(const double* __restrict__ S, const double* __restrict__ x,
                                                   double* __restrict__ y, int64_t n) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const double* r = S + row * n;
  double acc = 0.0;
  for (int64_t k = lane; k < n; k += 64) acc += r[k] * x[k];
  acc = wave_sum(acc);
  if (lane == 0) y[row] = acc;
}

// q = x . y, x <- y / |y|, history[it] = q (one workgroup)
__global__ __launch_bounds__(1024) void power_step_kernel(double* __restrict__ x, const double* __restrict__ y,
                                                          int64_t n, double* __restrict__ history, int it) {
  __shared__ double red[2][16];
  __shared__ double inv_norm;
  double q = 0.0, nn = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += blockDim.x) {
    q += x[k] * y[k];
    nn += y[k] * y[k];
  }
  q = wave_sum(q);
  nn = wave_sum(nn);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = q;
    red[1][w] = nn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double qs = 0.0, ns = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      qs += red[0][i];
      ns += red[1][i];
    }
    history[it] = qs;
    inv_norm = ns > 0.0 ? 1.0 / sqrt(ns) : 0.0;
  }
  __syncthreads();
  for (int64_t k = threadIdx.x; k < n; k += blockDim.x) x[k] = y[k] * inv_norm;
}

__global__ void fill_kernel(double* __restrict__ x, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    x[k] = 1.0 + 0.25 * sin(0.7 * (double)k);  // not orthogonal to any particular eigenvector
}

// Unblocked Cholesky of the kb x kb diagonal block at (k0, k0) of A (lower triangle, ld n), in LDS.
// A non-positive pivot sets *info to its global index + 1.
__global__ __launch_bounds__(256) void potrf_diag_kernel(double* __restrict__ A, int64_t n, int64_t k0, int kb,
                                                         int* __restrict__ info) {
  __shared__ double L[kNb][kNb + 1];
  for (int t = threadIdx.x; t < kNb * kNb; t += blockDim.x) {
    const int i = t / kNb, j = t % kNb;
    L[i][j] = (i < kb && j < kb) ? A[(k0 + i) * n + k0 + j] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    // every thread reads the pivot after the previous column's updates
    const double d = L[j][j];
    const double sd = d > 0.0 ? sqrt(d) : 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
      if (!(d > 0.0) && *info == 0) *info = (int)(k0 + j + 1);
      L[j][j] = sd;
    }
    for (int i = j + 1 + threadIdx.x; i < kb; i += blockDim.x) L[i][j] = sd > 0.0 ? L[i][j] / sd : 0.0;
    __syncthreads();
    // trailing update of the block: L[i][k] -= L[i][j] L[k][j] for j < k <= i < kb
    const int rem = kb - j - 1;
    for (int t = threadIdx.x; t < rem * rem; t += blockDim.x) {
      const int i = j + 1 + t / rem, k = j + 1 + t % rem;
      if (k <= i) L[i][k] -= L[i][j] * L[k][j];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < kb * kb; t += blockDim.x) {
    const int i = t / kb, j = t % kb;
    A[(k0 + i) * n + k0 + j] = j <= i ? L[i][j] : 0.0;
  }
}

// Inverse of the kb x kb lower-triangular diagonal block at (k0, k0) of L (ld n), written as a
// dense 64 x 64 array Linv (identity beyond kb).  Thread c solves L x = e_c by forward
// substitution in LDS.
__global__ __launch_bounds__(64) void trtri_kernel(const double* __restrict__ Lm, int64_t n, int64_t k0, int kb,
                                                   double* __restrict__ Linv) {
  __shared__ double L[kNb][kNb + 1];
  __shared__ double X[kNb][kNb + 1];
  for (int t = threadIdx.x; t < kNb * kNb; t += blockDim.x) {
    const int i = t / kNb, j = t % kNb;
    L[i][j] = (i < kb && j < kb) ? Lm[(k0 + i) * n + k0 + j] : (i == j ? 1.0 : 0.0);
  }
  __syncthreads();
  const int c = threadIdx.x;
  for (int j = 0; j < kNb; ++j) {
    double v = (j == c) ? 1.0 : 0.0;
    for (int i = c; i < j; ++i) v -= L[j][i] * X[i][c];  // X[i][c] = 0 for i < c
    X[j][c] = (j < c) ? 0.0 : v / L[j][j];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kNb * kNb; t += blockDim.x) Linv[t] = X[t / kNb][t % kNb];
}

// C = beta C + alpha opA opB over 64 x 64 output tiles, K in chunks of 16 through LDS; opA = A
// (lda) or A^T, opB = B (ldb) or B^T; C is not read when beta is 0.  With lower != 0 only tiles
// with tile row >= tile column are computed (the lower triangle of a symmetric update; the
// diagonal tiles whole).
__global__ __launch_bounds__(256) void gemm_kernel(double* __restrict__ C, int64_t ldc, const double* __restrict__ A,
                                                   int64_t lda, int transA, const double* __restrict__ B, int64_t ldb,
                                                   int transB, int64_t M, int64_t N, int64_t K, double alpha, double beta,
                                                   int lower) {
  const int64_t ti = blockIdx.y, tj = blockIdx.x;
  if (lower && tj > ti) return;
  __shared__ double As[16][64 + 1];
  __shared__ double Bs[16][64 + 1];
  const int tid = threadIdx.x;
  const int tr = tid / 16, tc = tid % 16;  // 16 x 16 threads, 4 x 4 outputs each
  double acc[4][4] = {};
  const int64_t i0 = ti * 64, j0 = tj * 64;
  for (int64_t k0 = 0; k0 < K; k0 += 16) {
    for (int t = tid; t < 16 * 64; t += 256) {
      const int kk = t / 64, ii = t % 64;
      const int64_t gi = i0 + ii, gk = k0 + kk;
      double av = 0.0;
      if (gi < M && gk < K) av = transA ? A[gk * lda + gi] : A[gi * lda + gk];
      As[kk][ii] = av;
      const int64_t gj = j0 + ii;
      double bv = 0.0;
      if (gj < N && gk < K) bv = transB ? B[gj * ldb + gk] : B[gk * ldb + gj];
      Bs[kk][ii] = bv;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = As[kk][tr + 16 * u];
        b[u] = Bs[kk][tc + 16 * u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t gi = i0 + tr + 16 * u, gj = j0 + tc + 16 * v;
      if (gi < M && gj < N) {
        double* c = &C[gi * ldc + gj];
        *c = beta == 0.0 ? alpha * acc[u][v] : beta * *c + alpha * acc[u][v];
      }
    }
}

int gemm(double* C, int64_t ldc, const double* A, int64_t lda, int transA, const double* B, int64_t ldb, int transB,
         int64_t M, int64_t N, int64_t K, double alpha, double beta, int lower, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return kOk;
  HCIPY_REQUIRE(ceil_div(M, 64) <= 65535, "matrix of %lld rows", (long long)M);
  const dim3 g((unsigned)ceil_div(N, 64), (unsigned)ceil_div(M, 64));
  hipLaunchKernelGGL(gemm_kernel, g, dim3(256), 0, s, C, ldc, A, lda, transA, B, ldb, transB, M, N, K, alpha, beta,
                     lower);
  return check_hip(hipGetLastError(), "gemm_kernel launch");
}

unsigned grid_for(int64_t total) { return (unsigned)std::min<int64_t>(ceil_div(total, 256), 65536); }

struct TikWs {
  size_t a = 0, b = 0, c = 0, s = 0, x = 0, y = 0, hist = 0, info = 0, linv = 0, tmp = 0, total = 0;
  int64_t ld = 0, rows_a = 0;
};
constexpr int kPowerMax = 20000;
constexpr int kPowerChunk = 50;

TikWs tik_layout(int64_t m, int64_t n) {
  TikWs w;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (size_t)round_up((int64_t)bytes, 256);
    return at;
  };
  w.ld = round_up(2 * ceil_div(m, 2), 128);  // reals per operand row, whole 128-element k-tiles
  w.rows_a = round_up(n, 64);
  w.a = take((size_t)w.rows_a * w.ld * 8);
  w.b = take((size_t)n * w.ld * 8);
  w.c = take((size_t)n * n * 16);
  w.s = take((size_t)n * n * 8);
  w.x = take((size_t)n * 8);
  w.y = take((size_t)n * 8);
  w.hist = take((size_t)kPowerMax * 8);
  w.info = take(sizeof(int));
  w.linv = take((size_t)kNb * kNb * 8);
  w.tmp = take((size_t)kNb * std::max(n, m) * 8);  // a panel (rows x 64) or a block row (64 x m)
  w.total = off;
  return w;
}

}  // namespace
}  // namespace hcipy

using namespace hcipy;

extern "C" {

int hcipy_inverse_tikhonov_workspace(int64_t m, int64_t n, size_t* bytes) {
  HCIPY_REQUIRE(bytes != nullptr && m > 0 && n > 0, "bad argument");
  *bytes = tik_layout(m, n).total;
  return kOk;
}

int hcipy_inverse_tikhonov(const double* M, int64_t m, int64_t n, double rcond, double* out, double* s_max,
                           void* workspace, size_t workspace_bytes, void* stream) {
  HCIPY_REQUIRE(M != nullptr && out != nullptr && m > 0 && n > 0, "bad argument");
  HCIPY_REQUIRE(rcond >= 0.0 && std::isfinite(rcond), "rcond must be finite and non-negative");
  HCIPY_REQUIRE(n <= 16384 && m <= ((int64_t)1 << 28), "matrix of %lld x %lld is too large", (long long)m, (long long)n);
  const TikWs w = tik_layout(m, n);
  HCIPY_REQUIRE(workspace != nullptr && workspace_bytes >= w.total, "workspace too small: %zu < %zu bytes",
                workspace_bytes, w.total);
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* ws = static_cast<char*>(workspace);
  double* A = reinterpret_cast<double*>(ws + w.a);
  double* B = reinterpret_cast<double*>(ws + w.b);
  double2* C = reinterpret_cast<double2*>(ws + w.c);
  double* S = reinterpret_cast<double*>(ws + w.s);
  double* x = reinterpret_cast<double*>(ws + w.x);
  double* y = reinterpret_cast<double*>(ws + w.y);
  double* hist = reinterpret_cast<double*>(ws + w.hist);
  int* info = reinterpret_cast<int*>(ws + w.info);
  double* Linv = reinterpret_cast<double*>(ws + w.linv);
  double* T = reinterpret_cast<double*>(ws + w.tmp);

  // 1. Gram matrix M^T M: complex Gram of the paired columns on the FP64 stage kernel.
  hipLaunchKernelGGL(gram_operands_kernel, dim3(grid_for(w.rows_a * w.ld)), dim3(256), 0, s, M, m, n, A, B, w.ld,
                     w.rows_a);
  HCIPY_HIP(hipGetLastError());
  StageArgs g{};
  g.a = A;
  g.lda = w.ld;
  g.b = B;
  g.ldb = w.ld;
  g.M = (int)(2 * n);
  g.N = (int)n;
  g.K = (int)w.ld;
  g.out = C;
  g.ldo = n;
  g.out_mode = kOutC128;
  g.alpha_re = 1.0;
  g.tile_cfg = kTileAuto;
  if (int st = launch_stage(kF64, g, 1, 0 /* rows padded */, s)) return st;
  hipLaunchKernelGGL(real_part_kernel, dim3(grid_for(n * n)), dim3(256), 0, s, C, S, n, 0.0);
  HCIPY_HIP(hipGetLastError());

  // 2. max(s)^2 = largest eigenvalue of M^T M by power iteration (Rayleigh quotients, converged to
  //    1e-15 relative between chunks of iterations).
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, s, x, n);
  HCIPY_HIP(hipGetLastError());
  double lam_max = 0.0;
  std::vector<double> h(kPowerChunk);
  bool converged = false;
  for (int it0 = 0; it0 < kPowerMax && !converged; it0 += kPowerChunk) {
    for (int it = it0; it < it0 + kPowerChunk; ++it) {
      hipLaunchKernelGGL(symv_kernel, dim3((unsigned)ceil_div(n, 4)), dim3(256), 0, s, S, x, y, n);
      hipLaunchKernelGGL(power_step_kernel, dim3(1), dim3(1024), 0, s, x, y, n, hist, it);
    }
    HCIPY_HIP(hipGetLastError());
    HCIPY_HIP(hipMemcpyAsync(h.data(), hist + it0, kPowerChunk * sizeof(double), hipMemcpyDeviceToHost, s));
    HCIPY_HIP(hipStreamSynchronize(s));
    const double q1 = h[kPowerChunk - 1], q0 = h[kPowerChunk - 2];
    lam_max = q1;
    converged = !(q1 > 0.0) || std::fabs(q1 - q0) <= 1e-15 * q1;
  }
  HCIPY_REQUIRE(converged, "power iteration for the largest singular value did not converge in %d steps", kPowerMax);
  const double smax = lam_max > 0.0 ? std::sqrt(lam_max) : 0.0;
  if (s_max != nullptr) *s_max = smax;
  HCIPY_REQUIRE(smax > 0.0, "the matrix is zero");
  const double lam = rcond * smax;

  // 3. Cholesky of M^T M + lam^2 I (lower, in place).
  hipLaunchKernelGGL(real_part_kernel, dim3(grid_for(n * n)), dim3(256), 0, s, C, S, n, lam * lam);
  HCIPY_HIP(hipMemsetAsync(info, 0, sizeof(int), s));
  for (int64_t k0 = 0; k0 < n; k0 += kNb) {
    const int kb = (int)std::min<int64_t>(kNb, n - k0);
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(256), 0, s, S, n, k0, kb, info);
    const int64_t rest = n - k0 - kb;
    if (rest > 0) {
      // L21 = A21 L11^-T through the explicit inverse of the diagonal block
      hipLaunchKernelGGL(trtri_kernel, dim3(1), dim3(kNb), 0, s, S, n, k0, kb, Linv);
      double* L21 = S + (k0 + kb) * n + k0;
      HCIPY_HIP(hipMemcpy2DAsync(T, kNb * 8, L21, (size_t)n * 8, (size_t)kb * 8, (size_t)rest, hipMemcpyDeviceToDevice,
                                 s));
      if (int st = gemm(L21, n, T, kNb, 0, Linv, kNb, 1, rest, kb, kb, 1.0, 0.0, 0, s)) return st;
      if (int st = gemm(S + (k0 + kb) * n + k0 + kb, n, L21, n, 0, L21, n, 1, rest, rest, kb, -1.0, 1.0, 1, s))
        return st;
    }
    HCIPY_HIP(hipGetLastError());
  }
  int hinfo = 0;
  HCIPY_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
  HCIPY_HIP(hipStreamSynchronize(s));
  HCIPY_REQUIRE(hinfo == 0, "M^T M + lam^2 I is not positive definite at row %d (rcond too small)", hinfo - 1);

  // 4. out = L^-T L^-1 M^T: M^T as right-hand sides (n x m), forward then backward block solves.
  {
    const int64_t total = n * m;
    // transpose of M through the B operand (row a of B holds column a of M)
    auto copy_rhs = [&]() -> int {
      for (int64_t a0 = 0; a0 < n; a0 += 65535) {
        const int64_t rows = std::min<int64_t>(65535, n - a0);
        HCIPY_HIP(hipMemcpy2DAsync(out + a0 * m, (size_t)m * 8, B + a0 * w.ld, (size_t)w.ld * 8, (size_t)m * 8,
                                   (size_t)rows, hipMemcpyDeviceToDevice, s));
      }
      return kOk;
    };
    if (int st = copy_rhs()) return st;
    (void)total;
  }
  // block row k of X: X_k <- L_kk^-1 X_k (forward) or L_kk^-T X_k (backward), through a copy
  auto diag_solve = [&](int64_t k0, int kb, int transpose) -> int {
    hipLaunchKernelGGL(trtri_kernel, dim3(1), dim3(kNb), 0, s, S, n, k0, kb, Linv);
    HCIPY_HIP(hipGetLastError());
    HCIPY_HIP(hipMemcpyAsync(T, out + k0 * m, (size_t)kb * m * 8, hipMemcpyDeviceToDevice, s));
    return gemm(out + k0 * m, m, Linv, kNb, transpose, T, m, 0, kb, m, kb, 1.0, 0.0, 0, s);
  };
  for (int64_t k0 = 0; k0 < n; k0 += kNb) {
    const int kb = (int)std::min<int64_t>(kNb, n - k0);
    if (int st = diag_solve(k0, kb, 0)) return st;
    const int64_t rest = n - k0 - kb;
    if (rest > 0)  // rows below: X[k0+kb:] -= L[k0+kb:, k0:k0+kb] X_k
      if (int st = gemm(out + (k0 + kb) * m, m, S + (k0 + kb) * n + k0, n, 0, out + k0 * m, m, 0, rest, m, kb, -1.0,
                        1.0, 0, s))
        return st;
  }
  for (int64_t k0 = (n - 1) / kNb * kNb; k0 >= 0; k0 -= kNb) {
    const int kb = (int)std::min<int64_t>(kNb, n - k0);
    if (int st = diag_solve(k0, kb, 1)) return st;
    if (k0 > 0)  // rows above: X[0:k0] -= L[k0:k0+kb, 0:k0]^T X_k
      if (int st = gemm(out, m, S + k0 * n, n, 1, out + k0 * m, m, 0, k0, m, kb, -1.0, 1.0, 0, s)) return st;
  }
  return kOk;
}

}  // extern "C"
