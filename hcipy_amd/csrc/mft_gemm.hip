// MFMA stage kernel, data-split kernel and Fourier-matrix generator of the MFT.
// The algebra and operand conventions are described in mft_gemm.hpp.
#include "mft_gemm.hpp"

#include <algorithm>
#include <cfloat>
#include <type_traits>

namespace hcipy {

namespace {

union Vec16 {
  uint4 u;
  double d[2];
  float f[4];
  _Float16 h[8];
};

template <typename T>
__device__ inline T zero_of() { return T(0); }

// Row index (inside one MFMA tile) of accumulator register r held by `lane`.
template <int MT>
__device__ inline int acc_row(int r, int lane) {
  if constexpr (MT == 32) {
    return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
  } else {
    return (lane >> 4) * 4 + r;
  }
}

}  // namespace

// One CTA (4 waves) computes a BM x BN tile of C = A . B^T.  Tiles of A and B are staged through
// LDS; the global loads of the next k-tile are issued before the MFMAs of the current one.
template <int P, int BM, int BN, int BK, int WM, int WN, bool B_ALIGNED>
__global__ __launch_bounds__(256) void mft_stage_kernel(StageArgs p) {
  using TR = PrecTraits<P>;
  using T = typename TR::T;
  constexpr int MT = TR::MT;
  constexpr int KVEC = TR::KVEC;
  constexpr int KPL = TR::KPL;
  constexpr int NTHR = 256;
  constexpr int VEC = 16 / (int)sizeof(T);
  constexpr int LDK = BK + VEC;  // 16-byte row padding against LDS bank conflicts
  constexpr int NARR = TR::SPLIT ? 2 : 1;
  constexpr int WTM = BM / WM;
  constexpr int WTN = BN / WN;
  constexpr int TM = WTM / MT;
  constexpr int TN = WTN / MT;
  constexpr int CPR = BK / VEC;  // 16-byte chunks per tile row
  constexpr int ACH = BM * CPR;
  constexpr int BCH = BN * CPR;
  constexpr int AL = (ACH + NTHR - 1) / NTHR;
  constexpr int BL = (BCH + NTHR - 1) / NTHR;
  static_assert(WM * WN == 4, "4 waves per CTA");
  static_assert(TM >= 1 && TN >= 1, "wave tile smaller than MFMA tile");
  static_assert(BK % KVEC == 0, "BK must hold whole k-vectors");

  __shared__ __attribute__((aligned(16))) T sA[NARR][BM * LDK];
  __shared__ __attribute__((aligned(16))) T sB[NARR][BN * LDK];
  __shared__ float s_t[BN];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / WN;
  const int wn = wave % WN;
  const int m0 = blockIdx.x * BM;
  const int n0 = blockIdx.y * BN;
  const int z = blockIdx.z;
  const int M = p.M, N = p.N, K = p.K;

  const T* Ag[2] = {static_cast<const T*>(p.a), static_cast<const T*>(p.a_lo)};
  const T* Bg[2] = {static_cast<const T*>(p.b), static_cast<const T*>(p.b_lo)};
  if constexpr (NARR == 2) Bg[1] += (int64_t)z * p.b_zstride;
  Bg[0] += (int64_t)z * p.b_zstride;

  // Split output: per folded item scale t = min(row_scale over the item) * 2^-ceil(log2 K).
  const bool split_out = (p.out_mode == kOutSplit);
  int item0 = 0;
  if (split_out) {
    const int nend = min(n0 + BN, N);
    item0 = n0 / p.fold;
    const int nitems = (nend - 1) / p.fold - item0 + 1;
    for (int it = wave; it < nitems; it += 4) {
      const int r0 = (item0 + it) * p.fold;
      float mn = FLT_MAX;
      for (int r = lane; r < p.fold; r += 64) mn = fminf(mn, p.row_scale[r0 + r]);
      mn = wave_min(mn);
      if (lane == 0) {
        const float t = ldexpf(mn, -p.t_log2k);
        s_t[it] = t;
        if (blockIdx.x == 0 && p.t_out != nullptr) p.t_out[item0 + it] = t;
      }
    }
  }

  Vec16 ra[NARR][AL];
  Vec16 rb[NARR][BL];

  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int arr = 0; arr < NARR; ++arr) {
#pragma unroll
      for (int i = 0; i < AL; ++i) {
        const int ch = tid + i * NTHR;
        if (ACH % NTHR == 0 || ch < ACH) {
          const int row = ch / CPR;
          const int kc = ch % CPR;
          ra[arr][i].u = *reinterpret_cast<const uint4*>(Ag[arr] + (int64_t)(m0 + row) * p.lda + k0 + kc * VEC);
        }
      }
#pragma unroll
      for (int i = 0; i < BL; ++i) {
        const int ch = tid + i * NTHR;
        if (BCH % NTHR == 0 || ch < BCH) {
          const int row = n0 + ch / CPR;
          const int kk = k0 + (ch % CPR) * VEC;
          Vec16 v;
          v.u = make_uint4(0, 0, 0, 0);
          if (row < N) {
            const T* src = Bg[arr] + (int64_t)row * p.ldb + kk;
            if (B_ALIGNED && kk + VEC <= K) {
              v.u = *reinterpret_cast<const uint4*>(src);
            } else {
              T* e = reinterpret_cast<T*>(&v);
#pragma unroll
              for (int j = 0; j < VEC; ++j) e[j] = (kk + j < K) ? src[j] : zero_of<T>();
            }
          }
          rb[arr][i] = v;
        }
      }
    }
  };

  auto store_tiles = [&]() {
#pragma unroll
    for (int arr = 0; arr < NARR; ++arr) {
#pragma unroll
      for (int i = 0; i < AL; ++i) {
        const int ch = tid + i * NTHR;
        if (ACH % NTHR == 0 || ch < ACH) {
          *reinterpret_cast<uint4*>(&sA[arr][(ch / CPR) * LDK + (ch % CPR) * VEC]) = ra[arr][i].u;
        }
      }
#pragma unroll
      for (int i = 0; i < BL; ++i) {
        const int ch = tid + i * NTHR;
        if (BCH % NTHR == 0 || ch < BCH) {
          *reinterpret_cast<uint4*>(&sB[arr][(ch / CPR) * LDK + (ch % CPR) * VEC]) = rb[arr][i].u;
        }
      }
    }
  };

  // Accumulators.
  using AccT = typename std::conditional<P == kF64, doublex4_t, floatx16_t>::type;
  constexpr int NACC = TR::NACC;
  AccT acc[TM][TN];
  AccT accx[TR::SPLIT ? TM : 1][TR::SPLIT ? TN : 1];
  double acc64[TR::SPLIT ? TM : 1][TR::SPLIT ? TN : 1][TR::SPLIT ? NACC : 1];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
#pragma unroll
      for (int r = 0; r < NACC; ++r) acc[i][j][r] = 0;
      if constexpr (TR::SPLIT) {
#pragma unroll
        for (int r = 0; r < NACC; ++r) {
          accx[i][j][r] = 0;
          acc64[i][j][r] = 0.0;
        }
      }
    }

  const int nk = (K + BK - 1) / BK;
  load_tiles(0);
  for (int kt = 0; kt < nk; ++kt) {
    __syncthreads();  // previous compute finished reading LDS
    store_tiles();
    __syncthreads();
    if (kt + 1 < nk) load_tiles((kt + 1) * BK);

#pragma unroll
    for (int kv = 0; kv < BK / KVEC; ++kv) {
      const int kofs = kv * KVEC + (lane / MT) * KPL;
      if constexpr (P == kF16x2) {
        half8_t fa[2][TM], fb[2][TN];
#pragma unroll
        for (int arr = 0; arr < 2; ++arr) {
#pragma unroll
          for (int i = 0; i < TM; ++i)
            fa[arr][i] = *reinterpret_cast<const half8_t*>(&sA[arr][(wm * WTM + i * MT + lane % MT) * LDK + kofs]);
#pragma unroll
          for (int j = 0; j < TN; ++j)
            fb[arr][j] = *reinterpret_cast<const half8_t*>(&sB[arr][(wn * WTN + j * MT + lane % MT) * LDK + kofs]);
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[0][i], fb[0][j], acc[i][j], 0, 0, 0);
            accx[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[0][i], fb[1][j], accx[i][j], 0, 0, 0);
            accx[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[1][i], fb[0][j], accx[i][j], 0, 0, 0);
          }
      } else if constexpr (P == kF32) {
        floatx4_t fa[TM], fb[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i)
          fa[i] = *reinterpret_cast<const floatx4_t*>(&sA[0][(wm * WTM + i * MT + lane % MT) * LDK + kofs]);
#pragma unroll
        for (int j = 0; j < TN; ++j)
          fb[j] = *reinterpret_cast<const floatx4_t*>(&sB[0][(wn * WTN + j * MT + lane % MT) * LDK + kofs]);
#pragma unroll
        for (int s = 0; s < KPL; ++s)
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[i][s], fb[j][s], acc[i][j], 0, 0, 0);
      } else {
        double fa[TM][KPL], fb[TN][KPL];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          const double* src = &sA[0][(wm * WTM + i * MT + lane % MT) * LDK + kofs];
#pragma unroll
          for (int s = 0; s < KPL; ++s) fa[i][s] = src[s];
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const double* src = &sB[0][(wn * WTN + j * MT + lane % MT) * LDK + kofs];
#pragma unroll
          for (int s = 0; s < KPL; ++s) fb[j][s] = src[s];
        }
#pragma unroll
        for (int s = 0; s < KPL; ++s)
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i][s], fb[j][s], acc[i][j], 0, 0, 0);
      }
    }

    if constexpr (TR::SPLIT) {
      // Flush the FP32 partial sums of this k-tile into FP64.
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int r = 0; r < NACC; ++r) {
            acc64[i][j][r] += (double)acc[i][j][r] + (double)accx[i][j][r] * kLoScale;
            acc[i][j][r] = 0.0f;
            accx[i][j][r] = 0.0f;
          }
    }
  }

  // Epilogue: complex values (rows 2c, 2c+1) of column n.
  const double inv_z = (p.z_scale != nullptr) ? 1.0 / (double)p.z_scale[z] : 1.0;
#pragma unroll
  for (int i = 0; i < TM; ++i) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wn * WTN + j * MT + lane % MT;
      if (n >= N) continue;
      double sc = inv_z;
      if (p.row_scale != nullptr) sc /= (double)p.row_scale[n];
      int64_t base;
      int64_t cstride;
      float t = 1.0f;
      if (p.fold > 0) {
        const int item = n / p.fold;
        const int y = n - item * p.fold;
        base = (int64_t)item * p.fold_stride + y;
        cstride = p.ldo;
        if (split_out) t = s_t[item - item0];
      } else if (p.nc_layout) {
        base = (int64_t)z * p.out_zstride + (int64_t)n * p.ldo;
        cstride = 1;
      } else {
        base = (int64_t)z * p.out_zstride + n;
        cstride = p.ldo;
      }
#pragma unroll
      for (int r = 0; r < NACC; r += 2) {
        const int m = m0 + wm * WTM + i * MT + acc_row<MT>(r, lane);
        if (m >= M) continue;
        const int c = m >> 1;
        double zr, zi;
        if constexpr (TR::SPLIT) {
          zr = acc64[i][j][r];
          zi = acc64[i][j][r + 1];
        } else {
          zr = (double)acc[i][j][r];
          zi = (double)acc[i][j][r + 1];
        }
        zr *= sc;
        zi *= sc;
        const double orr = zr * p.alpha_re - zi * p.alpha_im;
        const double oii = zr * p.alpha_im + zi * p.alpha_re;
        const int64_t idx = base + (int64_t)c * cstride;
        if (p.out_mode == kOutC128) {
          reinterpret_cast<double2*>(p.out)[idx] = make_double2(orr, oii);
        } else if (p.out_mode == kOutC64) {
          reinterpret_cast<float2*>(p.out)[idx] = make_float2((float)orr, (float)oii);
        } else {
          const double xr = orr * (double)t;
          const double xi = oii * (double)t;
          const _Float16 hr = (_Float16)xr;
          const _Float16 hi = (_Float16)xi;
          half2_t h = {hr, hi};
          half2_t l = {(_Float16)((xr - (double)hr) * (double)kLoInv), (_Float16)((xi - (double)hi) * (double)kLoInv)};
          reinterpret_cast<half2_t*>(p.out)[idx] = h;
          reinterpret_cast<half2_t*>(p.out_lo)[idx] = l;
        }
      }
    }
  }
}

// Splits rows of interleaved complex64 data into FP16 hi/lo arrays, with a power-of-two row scale
// s = 2^-e chosen so that max|x| * s < 1.  One CTA per row; row padding is zero-filled.
__global__ __launch_bounds__(256) void split_rows_kernel(const float* __restrict__ in, int64_t ld_in, int K,
                                                         _Float16* __restrict__ hi, _Float16* __restrict__ lo,
                                                         int64_t ld_out, float* __restrict__ row_scale) {
  __shared__ float red[4];
  const int64_t row = blockIdx.x;
  const float* x = in + row * ld_in;
  const int tid = threadIdx.x;
  float mx = 0.0f;
  for (int k = tid; k < K; k += 256) mx = fmaxf(mx, fabsf(x[k]));
  mx = wave_max(mx);
  if ((tid & 63) == 0) red[tid >> 6] = mx;
  __syncthreads();
  mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  float s = 1.0f;
  if (mx > 0.0f && mx <= FLT_MAX) {
    int e;
    frexpf(mx, &e);  // mx = f * 2^e, f in [0.5, 1)  =>  mx * 2^-e < 1
    s = ldexpf(1.0f, -e);
  }
  _Float16* h = hi + row * ld_out;
  _Float16* l = lo + row * ld_out;
  for (int k = tid; k < ld_out; k += 256) {
    const float v = (k < K) ? x[k] * s : 0.0f;
    const _Float16 vh = (_Float16)v;
    h[k] = vh;
    l[k] = (_Float16)((v - (float)vh) * kLoInv);
  }
  if (tid == 0) row_scale[row] = s;
}

// Expanded Fourier matrix: m(c, k) = exp(i * sign * a[c] * b[k]), phase product and sincos in FP64.
// One thread per (c, k) complex entry writes the 2x2 real block.  Entries outside (nc, nk) are 0.
__global__ void gen_matrix_kernel(const double* __restrict__ a, int nc, const double* __restrict__ b, int nk,
                                  double sign, int prec, void* out, void* out_lo, int64_t lda, int rows_pad) {
  const int64_t kc = lda / 2;
  const int64_t total = (int64_t)(rows_pad / 2) * kc;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(idx / kc);
    const int k = (int)(idx - (int64_t)c * kc);
    double mr = 0.0, mi = 0.0;
    if (c < nc && k < nk) {
      const double ph = a[c] * b[k];
      double sn, cs;
      sincos(ph, &sn, &cs);
      mr = cs;
      mi = sign * sn;
    }
    const double v[4] = {mr, -mi, mi, mr};  // (2c,2k) (2c,2k+1) (2c+1,2k) (2c+1,2k+1)
    const int64_t o0 = (int64_t)(2 * c) * lda + 2 * k;
    const int64_t off[4] = {o0, o0 + 1, o0 + lda, o0 + lda + 1};
    if (prec == kF64) {
      double* o = static_cast<double*>(out);
      for (int q = 0; q < 4; ++q) o[off[q]] = v[q];
    } else if (prec == kF32) {
      float* o = static_cast<float*>(out);
      for (int q = 0; q < 4; ++q) o[off[q]] = (float)v[q];
    } else {
      _Float16* oh = static_cast<_Float16*>(out);
      _Float16* ol = static_cast<_Float16*>(out_lo);
      for (int q = 0; q < 4; ++q) {
        const _Float16 hv = (_Float16)v[q];
        oh[off[q]] = hv;
        ol[off[q]] = (_Float16)((v[q] - (double)hv) * (double)kLoInv);
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Host-side launchers.

template <int P, int BM, int BN, int BK, bool AL>
static void launch_stage_t(const StageArgs& a, int nz, hipStream_t s) {
  dim3 grid((unsigned)ceil_div(a.M, BM), (unsigned)ceil_div(a.N, BN), (unsigned)nz);
  hipLaunchKernelGGL((mft_stage_kernel<P, BM, BN, BK, 2, 2, AL>), grid, dim3(256), 0, s, a);
}

int launch_stage(int prec, const StageArgs& a, int nz, bool b_aligned, hipStream_t s) {
  if (a.M <= 0 || a.N <= 0 || a.K <= 0 || nz <= 0) return kOk;
  if (prec == kF64) {
    if (b_aligned) launch_stage_t<kF64, 64, 64, 16, true>(a, nz, s);
    else launch_stage_t<kF64, 64, 64, 16, false>(a, nz, s);
  } else if (prec == kF32) {
    if (b_aligned) launch_stage_t<kF32, 64, 64, 32, true>(a, nz, s);
    else launch_stage_t<kF32, 64, 64, 32, false>(a, nz, s);
  } else if (prec == kF16x2) {
    if (!b_aligned) {
      set_last_error("split-FP16 stage requires an aligned, padded data operand");
      return kInvalidValue;
    }
    launch_stage_t<kF16x2, 64, 64, 64, true>(a, nz, s);
  } else {
    set_last_error("unknown precision %d", prec);
    return kInvalidValue;
  }
  return check_hip(hipGetLastError(), "mft_stage_kernel launch");
}

int launch_split_rows(const float* in, int64_t ld_in, int K, int64_t rows, _Float16* hi, _Float16* lo,
                      int64_t ld_out, float* row_scale, hipStream_t s) {
  if (rows <= 0) return kOk;
  hipLaunchKernelGGL(split_rows_kernel, dim3((unsigned)rows), dim3(256), 0, s, in, ld_in, K, hi, lo, ld_out,
                     row_scale);
  return check_hip(hipGetLastError(), "split_rows_kernel launch");
}

int launch_gen_matrix(const double* a, int nc, const double* b, int nk, double sign, int prec, void* out,
                      void* out_lo, int64_t lda, int rows_pad, hipStream_t s) {
  const int64_t total = (int64_t)(rows_pad / 2) * (lda / 2);
  const int64_t blocks = std::min<int64_t>(ceil_div(total, 256), 65536);
  hipLaunchKernelGGL(gen_matrix_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, nc, b, nk, sign, prec, out,
                     out_lo, lda, rows_pad);
  return check_hip(hipGetLastError(), "gen_matrix_kernel launch");
}

}  // namespace hcipy
