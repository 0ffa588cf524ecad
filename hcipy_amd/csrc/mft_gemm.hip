// MFMA stage kernel, data-split kernel and Fourier-matrix generator of the MFT.
// The algebra and operand conventions are described in mft_gemm.hpp.
#include "mft_gemm.hpp"

#include <algorithm>
#include <cfloat>
#include <type_traits>

namespace hcipy {

namespace {

template <typename T>
__device__ inline T zero_of() { return T(0); }

// Row index (inside one MFMA tile, in expanded-matrix row order) of accumulator register r held
// by `lane`.
//  32x32 (f16, f32): register r of lane l holds MFMA row (r/4)*8 + (l/32)*4 + r%4, so registers
//                    (r, r+1), r even, are rows (2c, 2c+1) of one complex value.
//  16x16x4 f64     : register r of lane l holds MFMA row 4r + l/16.  The A fragment is loaded
//                    with rows permuted by a_row_perm so that the lane again holds expanded rows
//                    4 (l/16) + r, i.e. complex pairs in registers (0,1) and (2,3).
template <int MT>
__device__ inline int acc_row(int r, int lane) {
  if constexpr (MT == 32) {
    return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
  } else {
    return (lane >> 4) * 4 + r;
  }
}

// Expanded-matrix row that MFMA row i of a tile reads (identity for 32x32 tiles).
template <int MT>
__device__ inline int a_row_perm(int i) {
  if constexpr (MT == 32) {
    return i;
  } else {
    return 4 * (i & 3) + (i >> 2);
  }
}

}  // namespace

// Zeroes the elements of a 16-byte vector of T at positions >= valid (valid may exceed the count).
template <typename T>
__device__ inline u32x4_t mask_tail(u32x4_t v, int valid) {
  constexpr int VEC = 16 / (int)sizeof(T);
  if (valid >= VEC) return v;
  unsigned w[4] = {v[0], v[1], v[2], v[3]};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if constexpr (sizeof(T) == 2) {
      const int e0 = 2 * c, e1 = 2 * c + 1;
      unsigned m = 0;
      if (e0 < valid) m |= 0x0000ffffu;
      if (e1 < valid) m |= 0xffff0000u;
      w[c] &= m;
    } else if constexpr (sizeof(T) == 4) {
      if (c >= valid) w[c] = 0;
    } else {
      if (c / 2 >= valid) w[c] = 0;
    }
  }
  return u32x4_t{w[0], w[1], w[2], w[3]};
}

// Element-wise masked load of a 16-byte vector of T (no alignment needed, no read past valid).
template <typename T>
__device__ inline u32x4_t load_masked(const T* src, int valid) {
  unsigned w[4] = {0, 0, 0, 0};
  if constexpr (sizeof(T) == 2) {
    const unsigned short* p = reinterpret_cast<const unsigned short*>(src);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j < valid) w[j / 2] |= (unsigned)p[j] << (16 * (j % 2));
  } else if constexpr (sizeof(T) == 4) {
    const unsigned* p = reinterpret_cast<const unsigned*>(src);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (j < valid) w[j] = p[j];
  } else {
    const unsigned long long* p = reinterpret_cast<const unsigned long long*>(src);
#pragma unroll
    for (int j = 0; j < 2; ++j)
      if (j < valid) {
        const unsigned long long q = p[j];
        w[2 * j] = (unsigned)q;
        w[2 * j + 1] = (unsigned)(q >> 32);
      }
  }
  return u32x4_t{w[0], w[1], w[2], w[3]};
}

// Data-operand row modes of the stage kernel.
enum BMode : int {
  kRowsPadded = 0,   // rows padded to a multiple of 64 elements: 16-byte loads never leave the row
  kRowsAligned = 1,  // 16-byte aligned rows of exactly K elements: vector loads inside K, scalar tail
  kRowsUnaligned = 2 // element loads only
};

// XCD-aware tile order.  Consecutive workgroups are dispatched round-robin over the 8 XCDs; the
// hardware id h is remapped so that XCD x owns a contiguous chunk of the logical tile order, and
// that order walks GM tile rows at a time, so each XCD's L2 sees a compact block of A and B rows.
__device__ inline void tile_of(int h, int gx, int gy, int& bx, int& by) {
  const int T = gx * gy;
  const int L = (T % 8 == 0) ? (h % 8) * (T / 8) + h / 8 : h;
  constexpr int GM = 4;
  const int per_group = GM * gy;
  const int group = L / per_group;
  const int first = group * GM;
  const int gm = min(gx - first, GM);
  const int r = L - group * per_group;
  bx = first + r % gm;
  by = r / gm;
}

// Expansion of one 16-byte chunk of a compact constant row (interleaved re, im of m) into the
// chunk of the expanded real row 2c (re, -im, ...) or 2c+1 (im, re, ...).
template <typename T>
__device__ inline u32x4_t expand_even(u32x4_t v) {
  if constexpr (sizeof(T) == 2) {
    return v ^ u32x4_t{0x80000000u, 0x80000000u, 0x80000000u, 0x80000000u};
  } else if constexpr (sizeof(T) == 4) {
    return v ^ u32x4_t{0u, 0x80000000u, 0u, 0x80000000u};
  } else {
    return v ^ u32x4_t{0u, 0u, 0u, 0x80000000u};
  }
}
template <typename T>
__device__ inline u32x4_t expand_odd(u32x4_t v) {
  if constexpr (sizeof(T) == 2) {
    return (v >> 16) | (v << 16);
  } else if constexpr (sizeof(T) == 4) {
    return u32x4_t{v[1], v[0], v[3], v[2]};
  } else {
    return u32x4_t{v[2], v[3], v[0], v[1]};
  }
}

// One workgroup (4 waves) computes a BM x BN tile of C = A . B^T.  The waves form a
// WM x WN x WK grid: WM x WN split the tile, WK split the k-vectors of every k-tile (their
// partial sums are added through LDS at the end).
//
// Pipeline: global loads of the constant operand (compact complex rows) and of the data operand
// are issued PF k-tiles ahead into register sets, unconditionally and with clamped addresses, so
// that the compiler's wait counters stay exact; masking of rows >= N / k >= K and the expansion
// of the constant into real 2x2 blocks happen when a set is written to the LDS double buffer.
// The k loop is unrolled by PF so register sets and accumulator sets are compile-time indices.
// Split path: the cross accumulator runs over the whole k range; the main (hi.hi) accumulator
// alternates between two sets per k-tile, and the set of tile t-1 is flushed into FP64 while
// the MFMAs of tile t run.
template <int P, int BM, int BN, int BK, int WM, int WN, int WK, int BMODE, int PF>
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
  constexpr int BMC = BM / 2;    // compact constant rows per tile
  constexpr int WTM = BM / WM;
  constexpr int WTN = BN / WN;
  constexpr int TM = WTM / MT;
  constexpr int TN = WTN / MT;
  constexpr int CPR = BK / VEC;  // 16-byte chunks per tile row
  constexpr int ACH = BMC * CPR;
  constexpr int BCH = BN * CPR;
  constexpr int AL = (ACH + NTHR - 1) / NTHR;
  constexpr int BL = (BCH + NTHR - 1) / NTHR;
  constexpr int KV_PER_TILE = BK / KVEC;
  constexpr int KV_PER_WAVE = KV_PER_TILE / WK;
  static_assert(PF == 2 || PF == 4, "k-tiles in flight: 2 or 4 (even, so accumulator sets alternate with t)");
  constexpr int ABUF = NARR * BM * LDK;
  constexpr int BUF = ABUF + NARR * BN * LDK;  // elements per LDS buffer
  static_assert(WM * WN * WK == 4, "4 waves per workgroup");
  static_assert(TM >= 1 && TN >= 1, "wave tile smaller than MFMA tile");
  static_assert(KV_PER_TILE % WK == 0, "k-vectors must split evenly over WK");
  static_assert(ACH % NTHR == 0 && BCH % NTHR == 0, "tiles must split evenly over the threads");

  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);
  __shared__ float s_t[BN];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wk = wave / (WM * WN);
  const int wm = (wave % (WM * WN)) / WN;
  const int wn = wave % WN;
  const int gx = (p.M + BM - 1) / BM;
  const int gy = (p.N + BN - 1) / BN;
  int bx, by;
  tile_of(blockIdx.x, gx, gy, bx, by);
  const int m0 = bx * BM;
  const int mc0 = m0 / 2;
  const int n0 = by * BN;
  const int z = blockIdx.z;
  const int M = p.M, N = p.N, K = p.K;
  const int nk = (K + BK - 1) / BK;

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
        if (bx == 0 && p.t_out != nullptr) p.t_out[item0 + it] = t;
      }
    }
  }

  // Per-thread chunk geometry (fixed over the k loop).
  int a_off[AL];
  int b_row[BL];
  int64_t b_off[BL];
  int b_kc[BL];
#pragma unroll
  for (int i = 0; i < AL; ++i) {
    const int ch = tid + i * NTHR;
    a_off[i] = (ch / CPR) * LDK + (ch % CPR) * VEC;  // LDS offset of the compact chunk's row (x2 below)
  }
#pragma unroll
  for (int i = 0; i < BL; ++i) {
    const int ch = tid + i * NTHR;
    b_row[i] = n0 + ch / CPR;
    b_kc[i] = (ch % CPR) * VEC;
    b_off[i] = (int64_t)min(b_row[i], N - 1) * p.ldb;
  }
  const int nk_full = K / BK;  // k-tiles with no element beyond K
  // Workgroups start their k loop at different k-tiles (when there is no partial k-tile) so that
  // concurrent workgroups read different row offsets and spread over the L2 channels.
  const int rot = (K % BK == 0) ? (int)((bx * 7 + by * 3) % nk) : 0;
  // Largest k offset a full-vector load may start at: padded rows allow every chunk of the last
  // k-tile; exact rows (K a multiple of VEC) are clamped to the last chunk.
  const int kmax_vec = (BMODE == kRowsPadded) ? (nk * BK - VEC) : (K - VEC);

  u32x4_t ra[PF][NARR][AL];
  u32x4_t rb[PF][NARR][BL];

  auto load_set = [&](auto sc, int tile) __attribute__((always_inline)) {
    constexpr int S = decltype(sc)::value;
    int pt = min(tile, nk - 1) + rot;
    pt -= (pt >= nk) ? nk : 0;
    const int k0 = pt * BK;
#pragma unroll
    for (int arr = 0; arr < NARR; ++arr) {
#pragma unroll
      for (int i = 0; i < AL; ++i) {
        const int ch = tid + i * NTHR;
        ra[S][arr][i] = *reinterpret_cast<const u32x4_t*>(Ag[arr] + (int64_t)(mc0 + ch / CPR) * p.lda + k0 +
                                                          (ch % CPR) * VEC);
      }
#pragma unroll
      for (int i = 0; i < BL; ++i) {
        if constexpr (BMODE == kRowsUnaligned) {
          const int kk = k0 + b_kc[i];
          rb[S][arr][i] = (b_row[i] < N && kk < K) ? load_masked<T>(Bg[arr] + b_off[i] + kk, K - kk)
                                                   : u32x4_t{0, 0, 0, 0};
        } else {
          const int kk = min(k0 + b_kc[i], kmax_vec);
          rb[S][arr][i] = *reinterpret_cast<const u32x4_t*>(Bg[arr] + b_off[i] + kk);
        }
      }
    }
  };

  // Writes register set S (k-tile `tile`) to LDS buffer `buf`.  MASK zeroes data elements with
  // k >= K (only the last, partial k-tile needs it).  Data rows >= N are not masked: they only
  // feed output columns that are never written.
  auto store_set = [&](auto sc, auto mc, int tile, int buf) __attribute__((always_inline)) {
    constexpr int S = decltype(sc)::value;
    constexpr bool MASK = decltype(mc)::value;
    T* base = smem + buf * BUF;
#pragma unroll
    for (int arr = 0; arr < NARR; ++arr) {
      T* sA = base + arr * BM * LDK;
      T* sB = base + ABUF + arr * BN * LDK;
#pragma unroll
      for (int i = 0; i < AL; ++i) {
        // compact row r -> expanded rows 2r (re, -im) and 2r+1 (im, re)
        const int o = a_off[i] + (a_off[i] / LDK) * LDK;
        *reinterpret_cast<u32x4_t*>(&sA[o]) = expand_even<T>(ra[S][arr][i]);
        *reinterpret_cast<u32x4_t*>(&sA[o + LDK]) = expand_odd<T>(ra[S][arr][i]);
      }
#pragma unroll
      for (int i = 0; i < BL; ++i) {
        const int ch = tid + i * NTHR;
        u32x4_t v = rb[S][arr][i];
        if constexpr (MASK && BMODE != kRowsUnaligned) v = mask_tail<T>(v, K - (tile * BK + b_kc[i]));
        *reinterpret_cast<u32x4_t*>(&sB[(ch / CPR) * LDK + (ch % CPR) * VEC]) = v;
      }
    }
  };

  using AccT = typename std::conditional<P == kF64, doublex4_t, floatx16_t>::type;
  constexpr int NACC = TR::NACC;
  constexpr int NSET = TR::SPLIT ? 2 : 1;
  AccT acc[NSET][TM][TN];
  // cross products Ah.Bl and Al.Bh in separate accumulators: the three MFMAs of a k-vector
  // are independent and issue back to back
  AccT accx[TR::SPLIT ? 2 : 1][TR::SPLIT ? TM : 1][TR::SPLIT ? TN : 1];
  double acc64[TR::SPLIT ? TM : 1][TR::SPLIT ? TN : 1][TR::SPLIT ? NACC : 1];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
#pragma unroll
      for (int r = 0; r < NACC; ++r) {
#pragma unroll
        for (int q = 0; q < NSET; ++q) acc[q][i][j][r] = 0;
        if constexpr (TR::SPLIT) {
          accx[0][i][j][r] = 0;
          accx[1][i][j][r] = 0;
          acc64[i][j][r] = 0.0;
        }
      }
    }

  int arow[TM];
#pragma unroll
  for (int i = 0; i < TM; ++i) arow[i] = wm * WTM + i * MT + a_row_perm<MT>(lane % MT);

  // MFMAs of one k-tile.  Fragments of k-vector q+1 are read from LDS into the second fragment
  // buffer while the MFMAs of k-vector q run.
  auto compute_tile = [&](auto qc, const T* base) __attribute__((always_inline)) {
    constexpr int Q = decltype(qc)::value;
    if constexpr (P == kF16x2) {
      half8_t fa[2][2][TM], fb[2][2][TN];  // [buffer][hi/lo][tile]
      auto read_frags = [&](int buf, int q) __attribute__((always_inline)) {
        const int kofs = (q * WK + wk) * KVEC + (lane / MT) * KPL;
#pragma unroll
        for (int arr = 0; arr < 2; ++arr) {
          const T* sA = base + arr * BM * LDK;
          const T* sB = base + ABUF + arr * BN * LDK;
#pragma unroll
          for (int i = 0; i < TM; ++i) fa[buf][arr][i] = *reinterpret_cast<const half8_t*>(&sA[arow[i] * LDK + kofs]);
#pragma unroll
          for (int j = 0; j < TN; ++j)
            fb[buf][arr][j] = *reinterpret_cast<const half8_t*>(&sB[(wn * WTN + j * MT + lane % MT) * LDK + kofs]);
        }
      };
      read_frags(0, 0);
#pragma unroll
      for (int q = 0; q < KV_PER_WAVE; ++q) {
        const int cb = q & 1;
        if (q + 1 < KV_PER_WAVE) read_frags(cb ^ 1, q + 1);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            acc[Q][i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[cb][0][i], fb[cb][0][j], acc[Q][i][j], 0, 0, 0);
            accx[0][i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[cb][0][i], fb[cb][1][j], accx[0][i][j], 0, 0, 0);
            accx[1][i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[cb][1][i], fb[cb][0][j], accx[1][i][j], 0, 0, 0);
          }
      }
    } else if constexpr (P == kF32) {
      floatx4_t fa[2][TM], fb[2][TN];
      auto read_frags = [&](int buf, int q) __attribute__((always_inline)) {
        const int kofs = (q * WK + wk) * KVEC + (lane / MT) * KPL;
#pragma unroll
        for (int i = 0; i < TM; ++i) fa[buf][i] = *reinterpret_cast<const floatx4_t*>(&base[arow[i] * LDK + kofs]);
#pragma unroll
        for (int j = 0; j < TN; ++j)
          fb[buf][j] = *reinterpret_cast<const floatx4_t*>(&base[ABUF + (wn * WTN + j * MT + lane % MT) * LDK + kofs]);
      };
      read_frags(0, 0);
#pragma unroll
      for (int q = 0; q < KV_PER_WAVE; ++q) {
        const int cb = q & 1;
        if (q + 1 < KV_PER_WAVE) read_frags(cb ^ 1, q + 1);
#pragma unroll
        for (int s = 0; s < KPL; ++s)
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              acc[0][i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[cb][i][s], fb[cb][j][s], acc[0][i][j], 0, 0, 0);
      }
    } else {
      doublex4_t fa[2][TM], fb[2][TN];
      auto read_frags = [&](int buf, int q) __attribute__((always_inline)) {
        const int kofs = (q * WK + wk) * KVEC + (lane / MT) * KPL;
#pragma unroll
        for (int i = 0; i < TM; ++i) fa[buf][i] = *reinterpret_cast<const doublex4_t*>(&base[arow[i] * LDK + kofs]);
#pragma unroll
        for (int j = 0; j < TN; ++j)
          fb[buf][j] = *reinterpret_cast<const doublex4_t*>(&base[ABUF + (wn * WTN + j * MT + lane % MT) * LDK + kofs]);
      };
      read_frags(0, 0);
#pragma unroll
      for (int q = 0; q < KV_PER_WAVE; ++q) {
        const int cb = q & 1;
        if (q + 1 < KV_PER_WAVE) read_frags(cb ^ 1, q + 1);
#pragma unroll
        for (int s = 0; s < KPL; ++s)
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[cb][i][s], fb[cb][j][s], acc[0][i][j], 0, 0, 0);
      }
    }
  };

  // Flush main accumulator set Q into FP64 and clear it.
  auto flush = [&](auto qc) __attribute__((always_inline)) {
    if constexpr (TR::SPLIT) {
      constexpr int Q = decltype(qc)::value;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int r = 0; r < NACC; ++r) {
            acc64[i][j][r] += (double)acc[Q][i][j][r];
            acc[Q][i][j][r] = 0.0f;
          }
    }
  };

  // One k-tile t = kb + D; D is a compile-time position inside the PF-unrolled loop.  The main
  // loop (guarded = false) only stores full k-tiles; the tail checks bounds and masks.
  auto step = [&](int kb, auto dc, auto gc) __attribute__((always_inline)) {
    constexpr int D = decltype(dc)::value;
    constexpr bool GUARDED = decltype(gc)::value;
    constexpr int NXT = (D + 1) % PF;          // register set of tile t + 1 and of t + 1 + PF
    constexpr int Q = D % NSET;                // accumulator set of tile t
    constexpr int QP = (D + NSET - 1) % NSET;  // set of tile t - 1
    const int t = kb + D;
    if (GUARDED && t >= nk) return;
    compute_tile(std::integral_constant<int, Q>{}, smem + (t & 1) * BUF);
    if constexpr (NSET == 2) flush(std::integral_constant<int, QP>{});
    if (!GUARDED || t + 1 < nk)
      store_set(std::integral_constant<int, NXT>{}, std::integral_constant<bool, GUARDED>{}, t + 1, (t + 1) & 1);
    load_set(std::integral_constant<int, NXT>{}, t + 1 + PF);
    __syncthreads();
  };
  using kFalse = std::integral_constant<bool, false>;
  using kTrue = std::integral_constant<bool, true>;

  // Prologue: tile 0 into LDS buffer 0, tiles 1..PF in flight (tile t lives in register set t % PF).
  load_set(std::integral_constant<int, 0>{}, 0);
  store_set(std::integral_constant<int, 0>{}, kTrue{}, 0, 0);
  load_set(std::integral_constant<int, 1 % PF>{}, 1);
  if constexpr (PF == 4) {
    load_set(std::integral_constant<int, 2>{}, 2);
    load_set(std::integral_constant<int, 3>{}, 3);
  }
  load_set(std::integral_constant<int, 0>{}, PF);
  __syncthreads();

  // Main loop: every tile it stores (up to kb + PF) is a full k-tile.
  int kb = 0;
  for (; kb + PF < nk_full; kb += PF) {
    step(kb, std::integral_constant<int, 0>{}, kFalse{});
    step(kb, std::integral_constant<int, 1>{}, kFalse{});
    if constexpr (PF == 4) {
      step(kb, std::integral_constant<int, 2>{}, kFalse{});
      step(kb, std::integral_constant<int, 3>{}, kFalse{});
    }
  }
  for (; kb < nk; kb += PF) {
    step(kb, std::integral_constant<int, 0>{}, kTrue{});
    step(kb, std::integral_constant<int, 1>{}, kTrue{});
    if constexpr (PF == 4) {
      step(kb, std::integral_constant<int, 2>{}, kTrue{});
      step(kb, std::integral_constant<int, 3>{}, kTrue{});
    }
  }
  if constexpr (TR::SPLIT) {
    flush(std::integral_constant<int, 0>{});
    flush(std::integral_constant<int, 1>{});
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < NACC; ++r)
          acc64[i][j][r] += ((double)accx[0][i][j][r] + (double)accx[1][i][j][r]) * kLoScale;
  }

  // Reduce the WK partial sums through LDS (the operand buffers are free now).
  if constexpr (WK > 1) {
    using RT = typename std::conditional<TR::SPLIT || P == kF64, double, float>::type;
    constexpr int NR = TM * TN * NACC;
    RT* red = reinterpret_cast<RT*>(smem_raw);
    const int slot = (wave % (WM * WN)) * 64 + lane;
    if (wk > 0) {
      RT* dst = red + ((wk - 1) * (WM * WN) * 64 + slot) * NR;
      int q = 0;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int r = 0; r < NACC; ++r, ++q) {
            if constexpr (TR::SPLIT) dst[q] = acc64[i][j][r];
            else dst[q] = (RT)acc[0][i][j][r];
          }
    }
    __syncthreads();
    if (wk > 0) return;
#pragma unroll
    for (int g = 1; g < WK; ++g) {
      const RT* srcp = red + ((g - 1) * (WM * WN) * 64 + slot) * NR;
      int q = 0;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int r = 0; r < NACC; ++r, ++q) {
            if constexpr (TR::SPLIT) acc64[i][j][r] += srcp[q];
            else acc[0][i][j][r] += srcp[q];
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
          zr = (double)acc[0][i][j][r];
          zi = (double)acc[0][i][j][r + 1];
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
        } else if (p.out_mode == kOutI32) {
          reinterpret_cast<float*>(p.out)[idx] = (float)(orr * orr + oii * oii);
        } else if (p.out_mode == kOutI64) {
          reinterpret_cast<double*>(p.out)[idx] = orr * orr + oii * oii;
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
// s = 2^-e chosen so that max|x| * s < 1.  One workgroup per row, 8 consecutive floats per thread
// (two 16-byte loads, two 16-byte stores); row padding up to ld_out is zero-filled.
__global__ __launch_bounds__(256) void split_rows_kernel(const float* __restrict__ in, int64_t ld_in, int K,
                                                         _Float16* __restrict__ hi, _Float16* __restrict__ lo,
                                                         int64_t ld_out, float* __restrict__ row_scale) {
  __shared__ float red[4];
  const int64_t row = blockIdx.x;
  const float* x = in + row * ld_in;
  const int tid = threadIdx.x;
  const bool vec_ok = ((ld_in & 3) == 0) && ((reinterpret_cast<uintptr_t>(in) & 15) == 0);
  float mx = 0.0f;
  for (int k0 = tid * 8; k0 < K; k0 += 256 * 8) {
    if (vec_ok && k0 + 8 <= K) {
      const float4 a = *reinterpret_cast<const float4*>(x + k0);
      const float4 b = *reinterpret_cast<const float4*>(x + k0 + 4);
      mx = fmaxf(mx, fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w))));
      mx = fmaxf(mx, fmaxf(fmaxf(fabsf(b.x), fabsf(b.y)), fmaxf(fabsf(b.z), fabsf(b.w))));
    } else {
      for (int j = 0; j < 8; ++j)
        if (k0 + j < K) mx = fmaxf(mx, fabsf(x[k0 + j]));
    }
  }
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
  for (int k0 = tid * 8; k0 < ld_out; k0 += 256 * 8) {
    float v[8];
    if (vec_ok && k0 + 8 <= K) {
      const float4 a = *reinterpret_cast<const float4*>(x + k0);
      const float4 b = *reinterpret_cast<const float4*>(x + k0 + 4);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = (k0 + j < K) ? x[k0 + j] : 0.0f;
    }
    half8_t vh, vl;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float sv = v[j] * s;
      const _Float16 q = (_Float16)sv;
      vh[j] = q;
      vl[j] = (_Float16)((sv - (float)q) * kLoInv);
    }
    // ld_out is a multiple of 64 halves, so k0 + 8 <= ld_out
    *reinterpret_cast<half8_t*>(h + k0) = vh;
    *reinterpret_cast<half8_t*>(l + k0) = vl;
  }
  if (tid == 0) row_scale[row] = s;
}

// Fast path of the split: one wave per row, the whole row (K <= 64 * PL floats) in registers,
// 16-byte loads, 8-byte stores, a wave-only max reduction.  Needs 16-byte aligned rows with
// K and ld_in multiples of 4.
template <int PL>
__global__ __launch_bounds__(256) void split_rows_wave_kernel(const float* __restrict__ in, int64_t ld_in, int K,
                                                              _Float16* __restrict__ hi, _Float16* __restrict__ lo,
                                                              int64_t ld_out, float* __restrict__ row_scale,
                                                              int64_t rows) {
  constexpr int NV = PL / 4;  // float4 chunks per lane
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  const float* x = in + row * ld_in;
  floatx4_t v[NV];
  float mx = 0.0f;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const int k = (j * 64 + lane) * 4;
    v[j] = (k < K) ? *reinterpret_cast<const floatx4_t*>(x + k) : floatx4_t{0.f, 0.f, 0.f, 0.f};
    mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v[j][0]), fabsf(v[j][1])), fmaxf(fabsf(v[j][2]), fabsf(v[j][3]))));
  }
  mx = wave_max(mx);
  float s = 1.0f;
  if (mx > 0.0f && mx <= FLT_MAX) {
    int e;
    frexpf(mx, &e);
    s = ldexpf(1.0f, -e);
  }
  typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
  _Float16* h = hi + row * ld_out;
  _Float16* l = lo + row * ld_out;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const int k = (j * 64 + lane) * 4;
    if (k < ld_out) {
      half4_t vh, vl;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float sv = v[j][e] * s;
        const _Float16 q = (_Float16)sv;
        vh[e] = q;
        vl[e] = (_Float16)((sv - (float)q) * kLoInv);
      }
      *reinterpret_cast<half4_t*>(h + k) = vh;
      *reinterpret_cast<half4_t*>(l + k) = vl;
    }
  }
  // zero padding beyond the register-held chunks
  for (int k = NV * 256 + lane * 4; k < ld_out; k += 256) {
    typedef _Float16 half4z_t __attribute__((ext_vector_type(4)));
    const half4z_t zq = {0, 0, 0, 0};
    *reinterpret_cast<half4z_t*>(h + k) = zq;
    *reinterpret_cast<half4z_t*>(l + k) = zq;
  }
  if (lane == 0) row_scale[row] = s;
}

// Compact Fourier matrix: m(c, k) = exp(i * sign * a[c] * b[k]), phase product and sincos in FP64,
// stored as interleaved (re, im) rows of lda reals (the stage kernel expands it per fragment).
// One thread per complex entry; entries outside (nc, nk) are 0.
__global__ void gen_matrix_kernel(const double* __restrict__ a, int nc, const double* __restrict__ b, int nk,
                                  double sign, int prec, void* out, void* out_lo, int64_t lda, int rows_pad) {
  const int64_t kc = lda / 2;
  const int64_t total = (int64_t)rows_pad * kc;
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
    const int64_t o = (int64_t)c * lda + 2 * k;
    if (prec == kF64) {
      double* q = static_cast<double*>(out);
      q[o] = mr;
      q[o + 1] = mi;
    } else if (prec == kF32) {
      float* q = static_cast<float*>(out);
      q[o] = (float)mr;
      q[o + 1] = (float)mi;
    } else {
      _Float16* oh = static_cast<_Float16*>(out);
      _Float16* ol = static_cast<_Float16*>(out_lo);
      const _Float16 hr = (_Float16)mr, hi = (_Float16)mi;
      oh[o] = hr;
      oh[o + 1] = hi;
      ol[o] = (_Float16)((mr - (double)hr) * (double)kLoInv);
      ol[o + 1] = (_Float16)((mi - (double)hi) * (double)kLoInv);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Host-side launchers.

template <int P, int BM, int BN, int BK, int WM, int WN, int WK, int BMODE, int PF>
static int launch_stage_t(const StageArgs& a, int nz, hipStream_t s) {
  using T = typename PrecTraits<P>::T;
  constexpr int VEC = 16 / (int)sizeof(T);
  constexpr int NARR = PrecTraits<P>::SPLIT ? 2 : 1;
  constexpr size_t lds = 2 * (size_t)NARR * (BM + BN) * (BK + VEC) * sizeof(T);
  static_assert(lds <= 160 * 1024, "LDS budget of one CU");
  auto kern = mft_stage_kernel<P, BM, BN, BK, WM, WN, WK, BMODE, PF>;
  static bool configured = false;  // per instantiation
  if (!configured) {
    HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
    configured = true;
  }
  const int gx = (int)ceil_div(a.M, BM);
  const int gy = (int)ceil_div(a.N, BN);
  hipLaunchKernelGGL(kern, dim3((unsigned)(gx * gy), 1, (unsigned)nz), dim3(256), lds, s, a);
  return kOk;
}

// Tile shape: 64 x 64 tiles (2 x 2 waves) when they give at least 256 workgroups, else 64 x 32
// tiles with the k-vectors split over 2 waves.  BK / PF per precision keep 32-64 KiB of operands
// in flight per workgroup.
template <int P, int BK, int BMODE, int PF>
static int launch_prec(const StageArgs& a, int nz, hipStream_t s) {
  const int64_t tiles64 = ceil_div(a.M, 64) * ceil_div(a.N, 64) * nz;
  if (tiles64 >= 256) return launch_stage_t<P, 64, 64, BK, 2, 2, 1, BMODE, PF>(a, nz, s);
  return launch_stage_t<P, 64, 32, BK, 2, 1, 2, BMODE, PF>(a, nz, s);
}

int launch_stage(int prec, const StageArgs& a, int nz, int b_mode, hipStream_t s) {
  if (a.M <= 0 || a.N <= 0 || a.K <= 0 || nz <= 0) return kOk;
  int st = kOk;
  if (prec == kF64) {
    if (b_mode == kRowsPadded) st = launch_prec<kF64, 32, kRowsPadded, 4>(a, nz, s);
    else if (b_mode == kRowsAligned) st = launch_prec<kF64, 32, kRowsAligned, 4>(a, nz, s);
    else st = launch_prec<kF64, 32, kRowsUnaligned, 4>(a, nz, s);
  } else if (prec == kF32) {
    if (b_mode == kRowsPadded) st = launch_prec<kF32, 64, kRowsPadded, 2>(a, nz, s);
    else if (b_mode == kRowsAligned) st = launch_prec<kF32, 64, kRowsAligned, 2>(a, nz, s);
    else st = launch_prec<kF32, 64, kRowsUnaligned, 2>(a, nz, s);
  } else if (prec == kF16x2) {
    if (b_mode != kRowsPadded) {
      set_last_error("split-FP16 stage requires a padded data operand");
      return kInvalidValue;
    }
    st = launch_prec<kF16x2, 128, kRowsPadded, 2>(a, nz, s);
  } else {
    set_last_error("unknown precision %d", prec);
    return kInvalidValue;
  }
  if (st) return st;
  return check_hip(hipGetLastError(), "mft_stage_kernel launch");
}

int launch_split_rows(const float* in, int64_t ld_in, int K, int64_t rows, _Float16* hi, _Float16* lo,
                      int64_t ld_out, float* row_scale, hipStream_t s) {
  if (rows <= 0) return kOk;
  const bool vec = ((reinterpret_cast<uintptr_t>(in) & 15) == 0) && (ld_in % 4 == 0) && (K % 4 == 0);
  const dim3 g((unsigned)ceil_div(rows, 4));
  if (vec && K <= 64 * 16 && ld_out <= 64 * 16 + 256)
    hipLaunchKernelGGL(split_rows_wave_kernel<16>, g, dim3(256), 0, s, in, ld_in, K, hi, lo, ld_out, row_scale, rows);
  else if (vec && K <= 64 * 32)
    hipLaunchKernelGGL(split_rows_wave_kernel<32>, g, dim3(256), 0, s, in, ld_in, K, hi, lo, ld_out, row_scale, rows);
  else if (vec && K <= 64 * 64)
    hipLaunchKernelGGL(split_rows_wave_kernel<64>, g, dim3(256), 0, s, in, ld_in, K, hi, lo, ld_out, row_scale, rows);
  else
    hipLaunchKernelGGL(split_rows_kernel, dim3((unsigned)rows), dim3(256), 0, s, in, ld_in, K, hi, lo, ld_out,
                       row_scale);
  return check_hip(hipGetLastError(), "split_rows_kernel launch");
}

int launch_gen_matrix(const double* a, int nc, const double* b, int nk, double sign, int prec, void* out,
                      void* out_lo, int64_t lda, int rows_pad, hipStream_t s) {
  const int64_t total = (int64_t)rows_pad * (lda / 2);
  const int64_t blocks = std::min<int64_t>(ceil_div(total, 256), 65536);
  hipLaunchKernelGGL(gen_matrix_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, nc, b, nk, sign, prec, out,
                     out_lo, lda, rows_pad);
  return check_hip(hipGetLastError(), "gen_matrix_kernel launch");
}

}  // namespace hcipy
