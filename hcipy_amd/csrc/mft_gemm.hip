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
  kRowsUnaligned = 2,// element loads only
  kRealPadded = 3,   // padded data rows, and a real constant of BM plain rows per tile (no 2x2
                     // expansion): C[m][n] is real and the epilogue writes single values
  kRealComb = 4      // parity stage 2 with recombination: BM = 64 constant rows (32 of each plane),
                     // BN = 128 data rows (32 of each parity item), wave w computes item w
};

// Parity recombination of one entry (v', u'): the four products of the parity items (ee, eo, oe,
// oo; x = real part, y = imaginary part) give the outputs at the up to four mirror positions
//   O(sv v', su u') = ee + i s su eo + i s sv oe - su sv oo     (s = -1 forward, +1 backward)
// times alpha, stored in out_mode at out[ob + v * nu + u].
__device__ inline void parity_store(double2 ee, double2 eo, double2 oe, double2 oo, int vp, int up, const ParAxis& au,
                                    const ParAxis& av, int nu, double sign, double ar, double ai, int out_mode,
                                    void* __restrict__ out, int64_t ob) {
  const int vis[2] = {par_plus(av, vp), par_minus(av, vp)};
  const int uis[2] = {par_plus(au, up), par_minus(au, up)};
#pragma unroll
  for (int qv = 0; qv < 2; ++qv) {
    const double sv = qv ? -1.0 : 1.0;
    const int vi = vis[qv];
    if (vi < 0) continue;
#pragma unroll
    for (int qu = 0; qu < 2; ++qu) {
      const double su = qu ? -1.0 : 1.0;
      const int ui = uis[qu];
      if (ui < 0) continue;
      const double a = sign * su, c = sign * sv, d = su * sv;
      const double orr = ee.x - a * eo.y - c * oe.y - d * oo.x;
      const double oii = ee.y + a * eo.x + c * oe.x - d * oo.y;
      const double zr = orr * ar - oii * ai, zi = orr * ai + oii * ar;
      const int64_t o = ob + (int64_t)vi * nu + ui;
      if (out_mode == kOutC64) static_cast<float2*>(out)[o] = make_float2((float)zr, (float)zi);
      else if (out_mode == kOutC128) static_cast<double2*>(out)[o] = make_double2(zr, zi);
      else if (out_mode == kOutI32) static_cast<float*>(out)[o] = (float)(zr * zr + zi * zi);
      else static_cast<double*>(out)[o] = zr * zr + zi * zi;
    }
  }
}

// XCD-aware tile order.  Consecutive workgroups are dispatched round-robin over the 8 XCDs; the
// hardware id h is remapped so that XCD x owns a contiguous chunk of the logical tile order, and
// that order walks GM tile rows at a time, so each XCD's L2 sees a compact block of A and B rows.
// slot_off is the XCD slot of the workgroup with h = 0 (nonzero inside a grouped launch, where an
// item's first workgroup need not sit on a multiple of 8); within each run of 8 consecutive h the
// slots (h + slot_off) mod 8 are all distinct, so the map stays a bijection on [0, T).
__device__ inline void tile_of(int h, int gx, int gy, int& bx, int& by, int slot_off = 0) {
  const int T = gx * gy;
  int L = h;
  if (T % 8 == 0) {
    L = ((h + slot_off) & 7) * (T / 8) + (h >> 3);
  } else if (slot_off == 0) {
    // chunks of q + 1 (the first r XCDs) or q logical tiles: still a bijection on [0, T)
    const int q = T >> 3, r = T & 7, x = h & 7;
    L = x * q + min(x, r) + (h >> 3);
  }
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
// Body of the stage kernel for workgroup tile h (in the XCD-aware order) of batch slice z.
// NYQK: the kernel may form an extra stage-1 row itself (StageArgs::nyq_coef; grouped launches).
template <int P, int BM, int BN, int BK, int WM, int WN, int WK, int BMODE, int PF, bool NYQK = false>
__device__ __forceinline__ void stage_body(const StageArgs& p, int h, int z, int slot_off = 0) {
  using TR = PrecTraits<P>;
  using T = typename TR::T;
  constexpr int MT = TR::MT;
  constexpr int KVEC = TR::KVEC;
  constexpr int KPL = TR::KPL;
  constexpr int NTHR = 64 * WM * WN * WK;
  constexpr int VEC = 16 / (int)sizeof(T);
  constexpr int LDK = BK + VEC;  // 16-byte row padding against LDS bank conflicts
  constexpr int NARR = TR::NARR;
  constexpr bool COMB = (BMODE == kRealComb);
  constexpr bool REAL = (BMODE == kRealPadded || COMB);
  static_assert(!COMB || (BM == 64 && BN == 128 && WM == 2 && WN == 2 && (WK == 1 || WK == 2) && TR::SPLIT),
                "recombining stage: 2 planes x 32 rows, 4 items x 32 rows, one wave (or two, splitting the "
                "k-vectors) per item");
  constexpr int BMC = REAL ? BM : BM / 2;  // constant rows loaded per tile (compact complex, or real)
  constexpr int WTM = BM / WM;
  constexpr int WTN = COMB ? 32 : BN / WN;
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
  static_assert(NTHR == 256 || NTHR == 512, "4 or 8 waves per workgroup");
  static_assert(TM >= 1 && TN >= 1, "wave tile smaller than MFMA tile");
  static_assert(KV_PER_TILE % WK == 0, "k-vectors must split evenly over WK");
  // Tiles with fewer 16-byte chunks than threads x AL (BL) leave the last threads without a chunk:
  // they neither load nor store it (A_ALL / B_ALL: every thread has every chunk).
  constexpr bool A_ALL = (ACH % NTHR == 0), B_ALL = (BCH % NTHR == 0);

  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);
  __shared__ float s_t[BN];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wk = wave / (WM * WN);
  // recombining stage: wave w takes constant plane w >> 1 (cosines for ee / eo, sines for oe / oo)
  // and parity item w
  const int wm = COMB ? ((wave & 3) >> 1) : (wave % (WM * WN)) / WN;
  const int wn = COMB ? (wave & 3) : wave % WN;
  constexpr int TBM = COMB ? 32 : BM;  // output rows and columns per workgroup
  constexpr int TBN = COMB ? 32 : BN;
  const int gx = (p.M + TBM - 1) / TBM;
  const int gy = (p.N + TBN - 1) / TBN;
  int bx, by;
  tile_of(h, gx, gy, bx, by, slot_off);
  const int m0 = bx * TBM;
  const int mc0 = m0 / 2;
  const int n0 = by * TBN;
  const int M = p.M, N = p.N, K = p.K;
  const int nk = (K + BK - 1) / BK;

  const T* Ag[3] = {static_cast<const T*>(p.a), static_cast<const T*>(p.a_lo), static_cast<const T*>(p.a_lo2)};
  const T* Bg[3] = {static_cast<const T*>(p.b), static_cast<const T*>(p.b_lo), static_cast<const T*>(p.b_lo2)};
  const int64_t a_plane = COMB ? 0 : (int64_t)((z >> p.a_zshift) & 1) * p.a_zstride;
#pragma unroll
  for (int arr = 0; arr < NARR; ++arr) {
    Ag[arr] += a_plane;
    Bg[arr] += (int64_t)z * (COMB ? 4 : 1) * p.b_zstride;
  }

  const bool split_out = (p.out_mode == kOutSplit || p.out_mode == kOutSplitT);
  int item0 = 0;

  // Per-thread chunk geometry (fixed over the k loop).
  int a_off[AL];
  int64_t a_goff[AL];
  int b_row[BL];
  int64_t b_off[BL];
  int b_kc[BL];
#pragma unroll
  for (int i = 0; i < AL; ++i) {
    const int ch = A_ALL ? tid + i * NTHR : min(tid + i * NTHR, ACH - 1);  // clamped: in bounds, unused
    const int r = ch / CPR;
    a_off[i] = r * LDK + (ch % CPR) * VEC;  // LDS offset of the compact chunk's row (x2 below)
    if constexpr (COMB) a_goff[i] = (int64_t)(r >> 5) * p.a_zstride + (int64_t)(m0 + (r & 31)) * p.lda;
    else a_goff[i] = (int64_t)((REAL ? m0 : mc0) + r) * p.lda;
    a_goff[i] += (ch % CPR) * VEC;
  }
#pragma unroll
  for (int i = 0; i < BL; ++i) {
    const int ch = B_ALL ? tid + i * NTHR : min(tid + i * NTHR, BCH - 1);
    const int r = ch / CPR;
    b_row[i] = n0 + (COMB ? (r & 31) : r);
    b_kc[i] = (ch % CPR) * VEC;
    b_off[i] = (int64_t)min(b_row[i], N - 1) * p.ldb;
    if constexpr (COMB) b_off[i] += (int64_t)(r >> 5) * p.b_zstride;
  }
  const int nk_full = K / BK;  // k-tiles with no element beyond K
  // Largest k offset a full-vector load may start at: padded rows allow every chunk of the last
  // k-tile; exact rows (K a multiple of VEC) are clamped to the last chunk.
  const int kmax_vec = (BMODE == kRowsPadded || REAL) ? (nk * BK - VEC) : (K - VEC);

  u32x4_t ra[PF][NARR][AL];
  u32x4_t rb[PF][NARR][BL];

  auto load_set = [&](auto sc, int tile) __attribute__((always_inline)) {
    constexpr int S = decltype(sc)::value;
    const int k0 = min(tile, nk - 1) * BK;
#pragma unroll
    for (int arr = 0; arr < NARR; ++arr) {
#pragma unroll
      for (int i = 0; i < AL; ++i)
        if (A_ALL || tid + i * NTHR < ACH) ra[S][arr][i] = *reinterpret_cast<const u32x4_t*>(Ag[arr] + a_goff[i] + k0);
#pragma unroll
      for (int i = 0; i < BL; ++i) {
        if (!B_ALL && tid + i * NTHR >= BCH) continue;
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
        if (!A_ALL && tid + i * NTHR >= ACH) continue;
        if constexpr (REAL) {
          *reinterpret_cast<u32x4_t*>(&sA[a_off[i]]) = ra[S][arr][i];
        } else {
          // compact row r -> expanded rows 2r (re, -im) and 2r+1 (im, re)
          const int o = a_off[i] + (a_off[i] / LDK) * LDK;
          *reinterpret_cast<u32x4_t*>(&sA[o]) = expand_even<T>(ra[S][arr][i]);
          *reinterpret_cast<u32x4_t*>(&sA[o + LDK]) = expand_odd<T>(ra[S][arr][i]);
        }
      }
#pragma unroll
      for (int i = 0; i < BL; ++i) {
        const int ch = tid + i * NTHR;
        if (!B_ALL && ch >= BCH) continue;
        u32x4_t v = rb[S][arr][i];
        if constexpr (MASK && BMODE != kRowsUnaligned) v = mask_tail<T>(v, K - (tile * BK + b_kc[i]));
        *reinterpret_cast<u32x4_t*>(&sB[(ch / CPR) * LDK + (ch % CPR) * VEC]) = v;
      }
    }
  };

  using AccT = typename std::conditional<P == kF64, doublex4_t, floatx16_t>::type;
  constexpr int NACC = TR::NACC;
  // The hi.hi FP32 accumulators are flushed into FP64 after every kFlushK real k values a wave
  // has accumulated (FE k-tiles).  With FE = 1 and a single MFMA tile per wave, two sets
  // alternate over k-tiles so that the flush of one overlaps the MFMAs of the other.
  constexpr int KPW = KV_PER_WAVE * KVEC;  // real k per wave per k-tile
  // The parity form flushes twice as often: its contiguous k ranges are coherent (the
  // products near the peak add up with one sign), so FP32 partial sums grow faster.
  constexpr int FLUSHK = REAL ? kFlushK / 2 : kFlushK;
  constexpr int FE = (KPW >= FLUSHK) ? 1 : FLUSHK / KPW;
  static_assert(FE * KPW == FLUSHK || FE == 1, "flush interval must be a whole number of k-tiles");
  constexpr int NSET = (TR::SPLIT && FE == 1 && TM * TN < 4 && !REAL) ? 2 : 1;
  // Cross products go to two accumulators for a single tile (independent MFMAs back to back),
  // to one when the wave has several tiles.
  constexpr int NXA = (TR::SPLIT && TM * TN < 2) ? 2 : 1;
  AccT acc[NSET][TM][TN];
  // cross products Ah.Bl and Al.Bh in separate accumulators: the three MFMAs of a k-vector
  // are independent and issue back to back
  AccT accx[NXA][TR::SPLIT ? TM : 1][TR::SPLIT ? TN : 1];
  // second-order cross products A0.B2 + A1.B1 + A2.B0 of the three-way BF16 split
  constexpr bool THREE = (NARR == 3);
  AccT accy[THREE ? TM : 1][THREE ? TN : 1];
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
#pragma unroll
          for (int q = 0; q < NXA; ++q) accx[q][i][j][r] = 0;
          acc64[i][j][r] = 0.0;
        }
        if constexpr (THREE) accy[i][j][r] = 0;
      }
    }

  int arow[TM];
#pragma unroll
  for (int i = 0; i < TM; ++i) arow[i] = wm * WTM + i * MT + a_row_perm<MT>(lane % MT);

  // MFMAs of one k-tile.  Fragments of k-vector q+1 are read from LDS into the second fragment
  // buffer while the MFMAs of k-vector q run.
  auto compute_tile = [&](auto qc, const T* base) __attribute__((always_inline)) {
    constexpr int Q = decltype(qc)::value;
    if constexpr (P == kF16x2 || P == kBF16x3) {
      using FragT = typename std::conditional<P == kF16x2, half8_t, bf16x8_t>::type;
      FragT fa[2][NARR][TM], fb[2][NARR][TN];  // [buffer][split array][tile]
      auto mfma = [](FragT a, FragT b, AccT c) __attribute__((always_inline)) {
        if constexpr (P == kF16x2) return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
        else return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
      };
      auto read_frags = [&](int buf, int q) __attribute__((always_inline)) {
        const int kofs = (q * WK + wk) * KVEC + (lane / MT) * KPL;
#pragma unroll
        for (int arr = 0; arr < NARR; ++arr) {
          const T* sA = base + arr * BM * LDK;
          const T* sB = base + ABUF + arr * BN * LDK;
#pragma unroll
          for (int i = 0; i < TM; ++i) fa[buf][arr][i] = *reinterpret_cast<const FragT*>(&sA[arow[i] * LDK + kofs]);
#pragma unroll
          for (int j = 0; j < TN; ++j)
            fb[buf][arr][j] = *reinterpret_cast<const FragT*>(&sB[(wn * WTN + j * MT + lane % MT) * LDK + kofs]);
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
            acc[Q][i][j] = mfma(fa[cb][0][i], fb[cb][0][j], acc[Q][i][j]);
            accx[0][i][j] = mfma(fa[cb][0][i], fb[cb][1][j], accx[0][i][j]);
            accx[NXA - 1][i][j] = mfma(fa[cb][1][i], fb[cb][0][j], accx[NXA - 1][i][j]);
            if constexpr (THREE) {
              accy[i][j] = mfma(fa[cb][0][i], fb[cb][NARR - 1][j], accy[i][j]);
              accy[i][j] = mfma(fa[cb][1][i], fb[cb][1][j], accy[i][j]);
              accy[i][j] = mfma(fa[cb][NARR - 1][i], fb[cb][0][j], accy[i][j]);
            }
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

  // Extra output row (first row tile of a real split-output stage): data row r of the tile is
  // shared by NQT threads, each summing NQK consecutive k of every k-tile in FP64.
  constexpr bool NYQ_STORE = REAL && TR::SPLIT && !COMB;  // extra row from split-formed values
  constexpr bool NYQ_OK = NYQ_STORE && NYQK;               // extra row formed here
  constexpr int NQT = NTHR / BN;
  constexpr int NQK = BK / NQT;
  static_assert(!NYQ_OK || (NQT * NQK == BK && NQK % 8 == 0 && NQT <= 64), "extra-row thread layout");
  const bool nyq_k = NYQ_OK && p.nyq_coef != nullptr && m0 == 0 && p.out_mode == kOutSplitT;
  float* coef_s = reinterpret_cast<float*>(smem + 2 * BUF);
  double nacc = 0.0;
  auto nyq_tile = [&](int t, const T* base) __attribute__((always_inline)) {
    if constexpr (NYQ_OK) {
      const int r = tid / NQT;
      const int kq = (tid % NQT) * NQK;
      const T* row = base + ABUF + r * LDK + kq;
      const float* cf = coef_s + t * BK + kq;
#pragma unroll
      for (int c = 0; c < NQK; c += 8) {
        T v[NARR][8];
#pragma unroll
        for (int arr = 0; arr < NARR; ++arr)
          *reinterpret_cast<u32x4_t*>(v[arr]) = *reinterpret_cast<const u32x4_t*>(row + arr * BN * LDK + c);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          double x = (double)(float)v[0][e];
          if constexpr (NARR == 2) x += (double)(float)v[1][e] * kLoScale;
          if constexpr (NARR == 3) x += (double)(float)v[1][e] + (double)(float)v[2][e];
          nacc += (double)cf[c + e] * x;
        }
      }
    }
  };

  // One k-tile t = kb + D; D is a compile-time position inside the PF-unrolled loop.  The main
  // loop (guarded = false) only stores full k-tiles; the tail checks bounds and masks.
  // Order: tile t + 1 goes from registers to the other LDS buffer (no wave reads it during this
  // step), the freed registers take the loads of tile t + 1 + PF, and the MFMAs of tile t then
  // run while those loads pass through the address units.  One barrier ends the step: it
  // publishes tile t + 1 and frees buffer t for the store of the next step.
  auto step = [&](int kb, auto dc, auto gc) __attribute__((always_inline)) {
    constexpr int D = decltype(dc)::value;
    constexpr bool GUARDED = decltype(gc)::value;
    constexpr int NXT = (D + 1) % PF;          // register set of tile t + 1 and of t + 1 + PF
    constexpr int Q = D % NSET;                // accumulator set of tile t
    constexpr int QP = (D + NSET - 1) % NSET;  // set of tile t - 1
    const int t = kb + D;
    if (GUARDED && t >= nk) return;
    if (!GUARDED || t + 1 < nk)
      store_set(std::integral_constant<int, NXT>{}, std::integral_constant<bool, GUARDED>{}, t + 1, (t + 1) & 1);
    load_set(std::integral_constant<int, NXT>{}, t + 1 + PF);
    compute_tile(std::integral_constant<int, Q>{}, smem + (t & 1) * BUF);
    if constexpr (NYQ_OK) {
      if (nyq_k) nyq_tile(t, smem + (t & 1) * BUF);
    }
    if constexpr (NSET == 2) flush(std::integral_constant<int, QP>{});
    else if constexpr (FE == 1) flush(std::integral_constant<int, 0>{});
    else if ((t + 1) % FE == 0) flush(std::integral_constant<int, 0>{});
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

  // Split output: per folded item scale t = min(row_scale over the item) * 2^-ceil(log2 K).
  // All threads load the item's row scales (independent loads, issued after the operand
  // prologue so that their latency overlaps it).
  if (split_out) {
    __shared__ float s_red[NTHR / 64];
    const int nend = min(n0 + BN, N);
    item0 = n0 / p.fold;
    const int nitems = (nend - 1) / p.fold - item0 + 1;
    for (int it = 0; it < nitems; ++it) {
      const float* rs = p.row_scale + (int64_t)z * p.rs_zstride + (int64_t)(item0 + it) * p.fold;
      float mn = FLT_MAX;
      int r = tid;
      for (; r + 3 * NTHR < p.fold; r += 4 * NTHR)
        mn = fminf(mn, fminf(fminf(rs[r], rs[r + NTHR]), fminf(rs[r + 2 * NTHR], rs[r + 3 * NTHR])));
      for (; r < p.fold; r += NTHR) mn = fminf(mn, rs[r]);
      mn = wave_min(mn);
      if (lane == 0) s_red[wave] = mn;
      __syncthreads();
      if (tid == 0) {
        float m = s_red[0];
#pragma unroll
        for (int w = 1; w < NTHR / 64; ++w) m = fminf(m, s_red[w]);
        const float t = ldexpf(m, -p.t_log2k);
        s_t[it] = t;
        if (bx == 0 && p.t_out != nullptr) p.t_out[(int64_t)z * p.t_zstride + item0 + it] = t;
      }
      __syncthreads();
    }
  }
  if constexpr (NYQ_OK) {
    if (nyq_k) {
      // the extra constant row, zero-padded to whole k-tiles, behind the operand buffers
      const float* cf = p.nyq_coef + (int64_t)((z >> p.a_zshift) & 1) * K;
      for (int k = tid; k < nk * BK; k += NTHR) coef_s[k] = k < K ? cf[k] : 0.0f;
    }
  }
  __syncthreads();

  // The extra output row formed before this kernel (FP64 values in p.nyq): stored by the first
  // row tile with the same scales and split as the product rows.
  if constexpr (NYQ_STORE) {
    if (p.nyq != nullptr && m0 == 0 && p.out_mode == kOutSplitT) {
      const double zs = (p.z_scale != nullptr) ? 1.0 / (double)p.z_scale[z] : 1.0;
      const int nend = min(n0 + BN, N);
      for (int n = n0 + tid; n < nend; n += NTHR) {
        const double v = p.nyq[(int64_t)z * p.nyq_zstride + n] * zs * (double)s_t[n / p.fold - item0];
        const int64_t idx = (int64_t)z * p.out_zstride + (int64_t)(2 * p.nyq_row + (n & 1)) * p.ldo + (n >> 1);
        if constexpr (P == kBF16x3) {
          const __bf16 q0 = f32_to_bf16((float)v);
          const double r1 = v - (double)(float)q0;
          const __bf16 q1 = f32_to_bf16((float)r1);
          static_cast<__bf16*>(p.out)[idx] = q0;
          static_cast<__bf16*>(p.out_lo)[idx] = q1;
          static_cast<__bf16*>(p.out_lo2)[idx] = f32_to_bf16((float)(r1 - (double)(float)q1));
        } else {
          const _Float16 hq = f32_to_f16((float)v);
          static_cast<_Float16*>(p.out)[idx] = hq;
          static_cast<_Float16*>(p.out_lo)[idx] = f32_to_f16((float)((v - (double)hq) * (double)kLoInv));
        }
      }
    }
  }

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
    if constexpr (NSET == 2) {
      flush(std::integral_constant<int, 0>{});
      flush(std::integral_constant<int, 1>{});
    } else if constexpr (FE > 1) {
      flush(std::integral_constant<int, 0>{});  // k-tiles after the last full flush interval
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < NACC; ++r) {
          double x = (double)accx[0][i][j][r];
          if constexpr (NXA == 2) x += (double)accx[1][i][j][r];
          if constexpr (THREE) acc64[i][j][r] += x + (double)accy[i][j][r];
          else acc64[i][j][r] += x * kLoScale;
        }
  }
  if constexpr (NYQ_OK) {
    if (nyq_k) {
#pragma unroll
      for (int off = NQT / 2; off >= 1; off >>= 1) nacc += __shfl_xor(nacc, off);
      const int n = n0 + tid / NQT;
      if (tid % NQT == 0 && n < N) {
        double v = nacc / (double)p.row_scale[(int64_t)z * p.rs_zstride + n] * (double)s_t[n / p.fold - item0];
        if (p.z_scale != nullptr) v /= (double)p.z_scale[z];
        const int64_t idx = (int64_t)z * p.out_zstride + (int64_t)(2 * p.nyq_row + (n & 1)) * p.ldo + (n >> 1);
        if constexpr (P == kBF16x3) {
          const __bf16 q0 = f32_to_bf16((float)v);
          const double r1 = v - (double)(float)q0;
          const __bf16 q1 = f32_to_bf16((float)r1);
          static_cast<__bf16*>(p.out)[idx] = q0;
          static_cast<__bf16*>(p.out_lo)[idx] = q1;
          static_cast<__bf16*>(p.out_lo2)[idx] = f32_to_bf16((float)(r1 - (double)(float)q1));
        } else {
          const _Float16 hq = f32_to_f16((float)v);
          static_cast<_Float16*>(p.out)[idx] = hq;
          static_cast<_Float16*>(p.out_lo)[idx] = f32_to_f16((float)((v - (double)hq) * (double)kLoInv));
        }
      }
    }
  }

  // Reduce the WK partial sums through LDS (the operand buffers are free now).
  if constexpr (WK > 1) {
    using RT = typename std::conditional<TR::SPLIT || P == kF64, double, float>::type;
    constexpr int NR = TM * TN * NACC;
    static_assert((size_t)(WK - 1) * WM * WN * 64 * NR * sizeof(RT) <= 2 * (size_t)BUF * sizeof(T),
                  "k-split partial sums must fit in the operand buffers");
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
    // the recombining epilogue below has barriers: its k-split waves stay (and add nothing)
    if constexpr (!COMB) {
      if (wk > 0) return;
    }
#pragma unroll
    for (int g = 1; g < WK; ++g) {
      if (COMB && wk > 0) break;
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

  if constexpr (COMB) {
    // Epilogue of the recombining stage: each wave's 32 x 32 products (item w, scaled back) go to
    // LDS, then every thread recombines entries (v', u') of the four items into the output.
    __syncthreads();
    constexpr int HLD = 33;  // padded row of doubles
    static_assert((size_t)4 * 32 * HLD * sizeof(double) <= 2 * (size_t)BUF * sizeof(T), "products fit in LDS");
    double* hs = reinterpret_cast<double*>(smem_raw);
    if (wk == 0) {  // the k-split waves' partial sums are in the first four waves' accumulators
      const double zs = (p.z_scale != nullptr) ? 1.0 / (double)p.z_scale[4 * z + wave] : 1.0;
#pragma unroll
      for (int r = 0; r < NACC; ++r) hs[(wave * 32 + acc_row<MT>(r, lane)) * HLD + lane % MT] = acc64[0][0][r] * zs;
    }
    __syncthreads();
    const int64_t ob = (int64_t)z * p.c_nu * p.c_nv;
    for (int e = tid; e < 32 * 16; e += NTHR) {
      const int mi = e >> 4, ui = e & 15;
      const int vp = m0 + mi, up = (n0 >> 1) + ui;
      if (vp >= M || 2 * up >= N) continue;
      double2 q[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double* row = hs + (c * 32 + mi) * HLD + 2 * ui;
        q[c] = make_double2(row[0], row[1]);
      }
      parity_store(q[0], q[1], q[2], q[3], vp, up, p.cu, p.cv, p.c_nu, p.c_sign, p.alpha_re, p.alpha_im, p.out_mode,
                   p.out, ob);
    }
    return;
  }

  const double inv_z = (p.z_scale != nullptr) ? 1.0 / (double)p.z_scale[z] : 1.0;
  if constexpr (REAL && P == kF16x2 && WK == 1 && TM == 1 && TN == 1) {
    // Split output of a parity stage 1 (kOutSplitT): element (m, n) goes to row 2m + (n & 1),
    // column n >> 1 of the hi and lo arrays.  Written from registers that is one 2-byte store per
    // accumulator and array, four 32-byte pieces of different rows per instruction; staged in LDS
    // (the operand buffers are free), the tile's 2 BM rows of BN / 2 values leave in 16-byte pieces.
    if (p.out_mode == kOutSplitT && BN % 16 == 0) {
      constexpr int RW = BN / 2;     // values per output row of the tile
      constexpr int LDR = RW + 8;    // padded LDS row (halves)
      constexpr int PIECES = RW / 8;  // 16-byte pieces per row
      static_assert((size_t)2 * 2 * BM * LDR * sizeof(_Float16) <= 2 * (size_t)BUF * sizeof(T), "staging fits in LDS");
      __syncthreads();  // every wave is done with the operand buffers
      _Float16* sh = reinterpret_cast<_Float16*>(smem_raw);
      _Float16* sl = sh + 2 * BM * LDR;
      const int nl = wn * WTN + lane % MT;
      const int n = n0 + nl;
      if (n < N) {
        double sc = inv_z;
        if (p.row_scale != nullptr) sc /= (double)p.row_scale[(int64_t)z * p.rs_zstride + n];
        sc *= (double)s_t[n / p.fold - item0];
#pragma unroll
        for (int r = 0; r < NACC; ++r) {
          const int ml = wm * WTM + acc_row<MT>(r, lane);
          const double v = acc64[0][0][r] * sc;
          const _Float16 h = f32_to_f16((float)v);
          const int o = (2 * ml + (nl & 1)) * LDR + (nl >> 1);
          sh[o] = h;
          sl[o] = f32_to_f16((float)((v - (double)h) * (double)kLoInv));
        }
      }
      __syncthreads();
      const int yc0 = n0 >> 1;         // first output column of the tile
      const int ycols = (N + 1) >> 1;  // output columns with data (N is even here)
      for (int t = tid; t < 2 * 2 * BM * PIECES; t += NTHR) {
        const int arr = t / (2 * BM * PIECES);
        const int rem = t - arr * (2 * BM * PIECES);
        const int rl = rem / PIECES, pc = rem - rl * PIECES;
        const int m = m0 + (rl >> 1);
        const int col = yc0 + pc * 8;
        if (m >= M || col >= ycols) continue;
        const _Float16* src = (arr ? sl : sh) + rl * LDR + pc * 8;
        _Float16* dst = static_cast<_Float16*>(arr ? p.out_lo : p.out) + (int64_t)z * p.out_zstride +
                        (int64_t)(2 * m + (rl & 1)) * p.ldo + col;
        // ordinary stores: the 64-byte row pieces of horizontally adjacent tiles meet in L2 and leave
        // as whole lines (G's compulsory 4.21 MB at 1024^2 -> 512^2; nontemporal stores wrote 4.82)
        if (col + 8 <= ycols) {
          *reinterpret_cast<u32x4_t*>(dst) = *reinterpret_cast<const u32x4_t*>(src);
        } else {
          for (int e = 0; e < ycols - col; ++e) dst[e] = src[e];
        }
      }
      return;
    }
  }
  if constexpr (REAL) {
    // Epilogue of a real constant: one real value per accumulator register.
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int n = n0 + wn * WTN + j * MT + lane % MT;
        if (n >= N) continue;
        double sc = inv_z;
        if (p.row_scale != nullptr) sc /= (double)p.row_scale[(int64_t)z * p.rs_zstride + n];
        if (split_out) sc *= (double)s_t[n / p.fold - item0];
#pragma unroll
        for (int r = 0; r < NACC; ++r) {
          const int m = m0 + wm * WTM + i * MT + acc_row<MT>(r, lane);
          if (m >= M) continue;
          double v;
          if constexpr (TR::SPLIT) v = acc64[i][j][r];
          else v = (double)acc[0][i][j][r];
          v *= sc;
          if (p.out_mode == kOutR64) {
            static_cast<double*>(p.out)[(int64_t)z * p.out_zstride + (int64_t)m * p.ldo + n] = v;
          } else if constexpr (TR::SPLIT) {
            const int64_t idx = (int64_t)z * p.out_zstride + (int64_t)(2 * m + (n & 1)) * p.ldo + (n >> 1);
            if constexpr (P == kBF16x3) {
              const __bf16 q0 = f32_to_bf16((float)v);
              const double r1 = v - (double)(float)q0;
              const __bf16 q1 = f32_to_bf16((float)r1);
              static_cast<__bf16*>(p.out)[idx] = q0;
              static_cast<__bf16*>(p.out_lo)[idx] = q1;
              static_cast<__bf16*>(p.out_lo2)[idx] = f32_to_bf16((float)(r1 - (double)(float)q1));
            } else {
              const _Float16 h = f32_to_f16((float)v);
              __builtin_nontemporal_store(h, static_cast<_Float16*>(p.out) + idx);
              __builtin_nontemporal_store(f32_to_f16((float)((v - (double)h) * (double)kLoInv)),
                                          static_cast<_Float16*>(p.out_lo) + idx);
            }
          }
        }
      }
    }
    return;
  }

  // Epilogue: complex values (rows 2c, 2c+1) of column n.
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
        } else if (p.out_mode == kOutI32Base || p.out_mode == kOutI64Base) {
          const double2 bz = static_cast<const double2*>(p.base)[idx];
          const double xr = orr + bz.x, xi = oii + bz.y;
          if (p.out_mode == kOutI32Base) reinterpret_cast<float*>(p.out)[idx] = (float)(xr * xr + xi * xi);
          else reinterpret_cast<double*>(p.out)[idx] = xr * xr + xi * xi;
        } else if constexpr (P == kBF16x3) {
          // b0 through FP32 (a double rounding moves it by one ulp at most); b1 and b2 take the
          // exact FP64 remainders
          const double xr = orr * (double)t;
          const double xi = oii * (double)t;
          const __bf16 r0 = f32_to_bf16((float)xr), i0 = f32_to_bf16((float)xi);
          const double xr1 = xr - (double)(float)r0, xi1 = xi - (double)(float)i0;
          const __bf16 r1 = f32_to_bf16((float)xr1), i1 = f32_to_bf16((float)xi1);
          const __bf16 r2 = f32_to_bf16((float)(xr1 - (double)(float)r1));
          const __bf16 i2 = f32_to_bf16((float)(xi1 - (double)(float)i1));
          reinterpret_cast<bf16x2_t*>(p.out)[idx] = bf16x2_t{r0, i0};
          reinterpret_cast<bf16x2_t*>(p.out_lo)[idx] = bf16x2_t{r1, i1};
          reinterpret_cast<bf16x2_t*>(p.out_lo2)[idx] = bf16x2_t{r2, i2};
        } else {
          // hi through FP32 (native conversions; a double rounding only moves hi by one ulp, and
          // lo below takes the exact remainder either way)
          const double xr = orr * (double)t;
          const double xi = oii * (double)t;
          const _Float16 hr = f32_to_f16((float)xr);
          const _Float16 hi = f32_to_f16((float)xi);
          half2_t h = {hr, hi};
          half2_t l = {f32_to_f16((float)((xr - (double)hr) * (double)kLoInv)),
                       f32_to_f16((float)((xi - (double)hi) * (double)kLoInv))};
          reinterpret_cast<half2_t*>(p.out)[idx] = h;
          reinterpret_cast<half2_t*>(p.out_lo)[idx] = l;
        }
      }
    }
  }
}

template <int P, int BM, int BN, int BK, int WM, int WN, int WK, int BMODE, int PF>
__global__ __launch_bounds__(64 * WM * WN * WK) void mft_stage_kernel(StageArgs p) {
  stage_body<P, BM, BN, BK, WM, WN, WK, BMODE, PF>(p, blockIdx.x, blockIdx.z);
}

// Grouped launch: one workgroup per tile of every item of a table of stage arguments (items may
// differ in matrices, sizes and batch).  Item i owns workgroups tile_start[i] .. tile_start[i+1]-1,
// in (batch slice, tile) order.
template <int P, int BM, int BN, int BK, int WM, int WN, int WK, int BMODE, int PF>
__global__ __launch_bounds__(64 * WM * WN * WK) void mft_stage_grouped_kernel(const StageArgs* __restrict__ table,
                                                                              const int64_t* __restrict__ tile_start,
                                                                              const int* __restrict__ tile_item,
                                                                              int n_items) {
  const int64_t b = blockIdx.x;
  int lo = 0;
  if (tile_item != nullptr) {
    lo = tile_item[b];  // item of every workgroup, tabulated by the caller
  } else {
    int hi = n_items - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tile_start[mid] <= b) lo = mid;
      else hi = mid - 1;
    }
  }
  const StageArgs p = table[lo];
  constexpr int TBM = BMODE == kRealComb ? 32 : BM;  // output tile of a workgroup
  constexpr int TBN = BMODE == kRealComb ? 32 : BN;
  const int64_t tiles_xy = (int64_t)((p.M + TBM - 1) / TBM) * ((p.N + TBN - 1) / TBN);
  const int64_t local = b - tile_start[lo];
  const int h = (int)(local % tiles_xy);
  // the XCD slot of this item's tile h = 0 in this batch slice
  stage_body<P, BM, BN, BK, WM, WN, WK, BMODE, PF, BMODE == kRealPadded>(p, h, (int)(local / tiles_xy),
                                                                        (int)((b - h) & 7));
}

// Split-FP16 stage kernel with the constant operand kept compact.
//
// Same contraction as mft_stage_kernel<kF16x2, ...>, arranged so that neither LDS nor HBM ever
// holds the expanded 2x2 real blocks of the constant: the workgroup tile is BMC compact constant
// rows (complex) x BN data rows, and every wave reads one compact A fragment per k-vector and
// forms in registers the operand of an "even" MFMA tile, rows (re, -im) -> Re of the result, and
// of an "odd" tile, rows (im, re) -> Im of the result (one xor / one rotate per dword).  Both
// tiles share the data fragment.  Per k-vector a wave reads 4 fragments (A hi/lo, B hi/lo) for
// 6 MFMAs per 32 x 32 tile pair, and LDS holds half the constant bytes of the expanded form.
//
// Waves form WM x WN x WK: WM x WN split the tile, WK split the k-vectors of every k-tile.
// Each wave flushes its FP32 hi.hi accumulators into FP64 after kFlushK real k values.
template <int BMC, int BN, int BK, int WM, int WN, int WK, int PF>
__global__ __launch_bounds__(64 * WM * WN * WK) void mft_f16x2_kernel(StageArgs p) {
  using T = _Float16;
  constexpr int NTHR = 64 * WM * WN * WK;
  constexpr int VEC = 8;                  // fp16 elements per 16-byte chunk
  constexpr int LDK = BK + VEC;           // LDS row stride (16-byte pad against bank conflicts)
  constexpr int WTC = BMC / WM;           // compact constant rows per wave
  constexpr int WTN = BN / WN;            // data rows per wave
  constexpr int TMC = WTC / 32;
  constexpr int TN = WTN / 32;
  constexpr int CPR = BK / VEC;           // chunks per tile row
  constexpr int ACH = BMC * CPR;
  constexpr int BCH = BN * CPR;
  constexpr int AL = ACH / NTHR;
  constexpr int BL = BCH / NTHR;
  constexpr int KV_PER_TILE = BK / 16;
  constexpr int KV_PER_WAVE = KV_PER_TILE / WK;
  constexpr int KPW = KV_PER_WAVE * 16;   // real k per wave per k-tile
  constexpr int FE = (KPW >= kFlushK) ? 1 : kFlushK / KPW;
  constexpr int ABUF = 2 * BMC * LDK;     // hi, lo constant rows
  constexpr int BUF = ABUF + 2 * BN * LDK;
  static_assert(NTHR == 256 || NTHR == 512, "4 or 8 waves per workgroup");
  static_assert(TMC >= 1 && TN >= 1 && WTC % 32 == 0 && WTN % 32 == 0, "wave tile of whole 32 x 32 MFMA tiles");
  static_assert(ACH % NTHR == 0 && BCH % NTHR == 0 && AL >= 1 && BL >= 1, "tiles must split evenly over the threads");
  static_assert(KV_PER_TILE % WK == 0 && KV_PER_WAVE >= 1, "k-vectors must split evenly over WK");
  static_assert(FE * KPW == kFlushK || FE == 1, "flush interval must be a whole number of k-tiles");
  static_assert(PF == 2, "two k-tiles in flight");

  // The two operand buffers are separate LDS objects, so the compiler knows that the stores of
  // tile t + 1 and the fragment reads of tile t never alias and may interleave them.
  __shared__ __attribute__((aligned(16))) T lds0[BUF];
  __shared__ __attribute__((aligned(16))) T lds1[BUF];
  __shared__ float s_t[BN];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wk = wave / (WM * WN);
  const int wm = (wave % (WM * WN)) / WN;
  const int wn = wave % WN;
  const int Mc = p.M / 2;  // compact constant rows
  const int gx = (Mc + BMC - 1) / BMC;
  const int gy = (p.N + BN - 1) / BN;
  int bx, by;
  tile_of(blockIdx.x, gx, gy, bx, by);
  const int c0 = bx * BMC;
  const int n0 = by * BN;
  const int z = blockIdx.z;
  const int N = p.N, K = p.K;
  const int nk = (K + BK - 1) / BK;
  const int nk_full = K / BK;

  const T* Ag[2] = {static_cast<const T*>(p.a), static_cast<const T*>(p.a_lo)};
  const T* Bg[2] = {static_cast<const T*>(p.b) + (int64_t)z * p.b_zstride,
                    static_cast<const T*>(p.b_lo) + (int64_t)z * p.b_zstride};
  const bool split_out = (p.out_mode == kOutSplit);
  int item0 = 0;

  // Per-thread chunk geometry.  Constant rows are padded to a multiple of 64 rows and data
  // rows >= N are clamped (they only feed output columns that are never written).
  int a_lds[AL], b_lds[BL];
  int64_t a_off[AL], b_off[BL];
#pragma unroll
  for (int i = 0; i < AL; ++i) {
    const int ch = tid + i * NTHR;
    a_lds[i] = (ch / CPR) * LDK + (ch % CPR) * VEC;
    a_off[i] = (int64_t)(c0 + ch / CPR) * p.lda + (ch % CPR) * VEC;
  }
  int b_kc[BL];
#pragma unroll
  for (int i = 0; i < BL; ++i) {
    const int ch = tid + i * NTHR;
    b_lds[i] = (ch / CPR) * LDK + (ch % CPR) * VEC;
    b_kc[i] = (ch % CPR) * VEC;
    b_off[i] = (int64_t)min(n0 + ch / CPR, N - 1) * p.ldb + b_kc[i];
  }

  u32x4_t ra[PF][2][AL];
  u32x4_t rb[PF][2][BL];
  // Loads of k-tile `tile` (clamped to the last one) into register set S.  Rows are padded to a
  // whole number of k-tiles, so full 16-byte loads of the last k-tile stay inside the row.
  auto load_set = [&](auto sc, int tile) __attribute__((always_inline)) {
    constexpr int S = decltype(sc)::value;
    const int k0 = min(tile, nk - 1) * BK;
#pragma unroll
    for (int arr = 0; arr < 2; ++arr) {
#pragma unroll
      for (int i = 0; i < AL; ++i) ra[S][arr][i] = *reinterpret_cast<const u32x4_t*>(Ag[arr] + a_off[i] + k0);
#pragma unroll
      for (int i = 0; i < BL; ++i) rb[S][arr][i] = *reinterpret_cast<const u32x4_t*>(Bg[arr] + b_off[i] + k0);
    }
  };
  // Register set S (k-tile `tile`) to LDS buffer `buf`; MASK zeroes data elements with k >= K.
  auto store_set = [&](auto sc, auto mc, int tile, T* base) __attribute__((always_inline)) {
    constexpr int S = decltype(sc)::value;
    constexpr bool MASK = decltype(mc)::value;
#pragma unroll
    for (int arr = 0; arr < 2; ++arr) {
      T* sA = base + arr * BMC * LDK;
      T* sB = base + ABUF + arr * BN * LDK;
#pragma unroll
      for (int i = 0; i < AL; ++i) *reinterpret_cast<u32x4_t*>(&sA[a_lds[i]]) = ra[S][arr][i];
#pragma unroll
      for (int i = 0; i < BL; ++i) {
        u32x4_t v = rb[S][arr][i];
        if constexpr (MASK) v = mask_tail<T>(v, K - (tile * BK + b_kc[i]));
        *reinterpret_cast<u32x4_t*>(&sB[b_lds[i]]) = v;
      }
    }
  };

  // Accumulators: [constant tile][0: Re (even rows), 1: Im (odd rows)][data tile].
  floatx16_t acc[TMC][2][TN];   // hi.hi
  floatx16_t accx[TMC][2][TN];  // hi.lo + lo.hi
  double acc64[TMC][2][TN][16];
#pragma unroll
  for (int i = 0; i < TMC; ++i)
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          acc[i][e][j][r] = 0.0f;
          accx[i][e][j][r] = 0.0f;
          acc64[i][e][j][r] = 0.0;
        }

  // Fragments of k-vector q + 1 are read while the MFMAs of q run (two fragment buffers) with 4
  // waves; with 8 waves the other wave of the SIMD covers the LDS latency and one buffer saves
  // the registers.
  constexpr int FB = (NTHR == 256) ? 2 : 1;
  auto compute_tile = [&](const T* base) __attribute__((always_inline)) {
    half8_t fa[FB][2][TMC], fb[FB][2][TN];  // [buffer][hi/lo][tile]
    auto read_frags = [&](int buf, int q) __attribute__((always_inline)) {
      const int kofs = (q * WK + wk) * 16 + (lane >> 5) * 8;
#pragma unroll
      for (int arr = 0; arr < 2; ++arr) {
        const T* sA = base + arr * BMC * LDK;
        const T* sB = base + ABUF + arr * BN * LDK;
#pragma unroll
        for (int i = 0; i < TMC; ++i)
          fa[buf][arr][i] = *reinterpret_cast<const half8_t*>(&sA[(wm * WTC + i * 32 + (lane & 31)) * LDK + kofs]);
#pragma unroll
        for (int j = 0; j < TN; ++j)
          fb[buf][arr][j] = *reinterpret_cast<const half8_t*>(&sB[(wn * WTN + j * 32 + (lane & 31)) * LDK + kofs]);
      }
    };
    if constexpr (FB == 2) read_frags(0, 0);
#pragma unroll
    for (int q = 0; q < KV_PER_WAVE; ++q) {
      const int cb = q % FB;
      if constexpr (FB == 2) {
        if (q + 1 < KV_PER_WAVE) read_frags(cb ^ 1, q + 1);
      } else {
        read_frags(0, q);
      }
      half8_t ae[2][TMC], ao[2][TMC];
#pragma unroll
      for (int arr = 0; arr < 2; ++arr)
#pragma unroll
        for (int i = 0; i < TMC; ++i) {
          const u32x4_t v = __builtin_bit_cast(u32x4_t, fa[cb][arr][i]);
          ae[arr][i] = __builtin_bit_cast(half8_t, v ^ u32x4_t{0x80000000u, 0x80000000u, 0x80000000u, 0x80000000u});
          ao[arr][i] = __builtin_bit_cast(half8_t, (v >> 16) | (v << 16));
        }
#pragma unroll
      for (int i = 0; i < TMC; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc[i][0][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ae[0][i], fb[cb][0][j], acc[i][0][j], 0, 0, 0);
          acc[i][1][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ao[0][i], fb[cb][0][j], acc[i][1][j], 0, 0, 0);
        }
#pragma unroll
      for (int i = 0; i < TMC; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          accx[i][0][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ae[0][i], fb[cb][1][j], accx[i][0][j], 0, 0, 0);
          accx[i][1][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ao[0][i], fb[cb][1][j], accx[i][1][j], 0, 0, 0);
        }
#pragma unroll
      for (int i = 0; i < TMC; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          accx[i][0][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ae[1][i], fb[cb][0][j], accx[i][0][j], 0, 0, 0);
          accx[i][1][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ao[1][i], fb[cb][0][j], accx[i][1][j], 0, 0, 0);
        }
    }
  };

  auto flush = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < TMC; ++i)
#pragma unroll
      for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            acc64[i][e][j][r] += (double)acc[i][e][j][r];
            acc[i][e][j][r] = 0.0f;
          }
  };

  // Instruction interleaving of one main-loop step: the LDS stores of tile t + 1 and the global
  // loads of tile t + 1 + PF are spread between the MFMAs of tile t, so that the LDS, the address
  // units and the matrix cores work at the same time instead of one after the other (all waves
  // of a workgroup run the same phase between two barriers).
  constexpr int NMF = 6 * TMC * TN;       // MFMAs per k-vector
  constexpr int NDR = 2 * (TMC + TN);     // fragment reads per k-vector
  constexpr int NST = 2 * (AL + BL);      // LDS stores (= global loads) per step
  auto interleave = [&]() __attribute__((always_inline)) {
    // one LDS store and one global load after each of the first NST MFMAs, fragment reads of
    // the next k-vector in between
#pragma unroll
    for (int q = 0; q < KV_PER_WAVE; ++q) {
#pragma unroll
      for (int m = 0; m < NMF; ++m) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
        if (q * NMF + m < NST) {
          __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // LDS store
          __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // global load
        }
        if (m < NDR && q + 1 < KV_PER_WAVE) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // fragment read
      }
    }
  };

  // Step t: tile t + 1 to the other LDS buffer, loads of tile t + 1 + PF into the freed
  // registers, MFMAs of tile t, flush at the end of a flush interval, one barrier.
  auto step = [&](int kb, auto dc, auto gc) __attribute__((always_inline)) {
    constexpr int D = decltype(dc)::value;
    constexpr bool GUARDED = decltype(gc)::value;
    constexpr int NXT = (D + 1) % PF;
    const int t = kb + D;  // kb is a multiple of PF = 2, so tile t lives in buffer D & 1
    T* cur = (D & 1) ? lds1 : lds0;
    T* nxt = (D & 1) ? lds0 : lds1;
    if (GUARDED && t >= nk) return;
    if (!GUARDED || t + 1 < nk)
      store_set(std::integral_constant<int, NXT>{}, std::integral_constant<bool, GUARDED>{}, t + 1, nxt);
    load_set(std::integral_constant<int, NXT>{}, t + 1 + PF);
    compute_tile(cur);
    // kb is a multiple of PF: for FE dividing PF the flush test is a compile-time one
    if constexpr (FE == 1) flush();
    else if constexpr (PF % FE == 0) { if constexpr ((D + 1) % FE == 0) flush(); }
    else if ((t + 1) % FE == 0) flush();
    if constexpr (!GUARDED) interleave();
    __syncthreads();
  };
  using kFalse = std::integral_constant<bool, false>;
  using kTrue = std::integral_constant<bool, true>;

  // Prologue: tile 0 into LDS buffer 0, tiles 1 and 2 in flight.
  load_set(std::integral_constant<int, 0>{}, 0);
  store_set(std::integral_constant<int, 0>{}, kTrue{}, 0, lds0);
  load_set(std::integral_constant<int, 1>{}, 1);
  load_set(std::integral_constant<int, 0>{}, 2);

  // Split output: per folded item scale t = min(row_scale over the item) * 2^-ceil(log2 K).
  if (split_out) {
    __shared__ float s_red[NTHR / 64];
    const int nend = min(n0 + BN, N);
    item0 = n0 / p.fold;
    const int nitems = (nend - 1) / p.fold - item0 + 1;
    for (int it = 0; it < nitems; ++it) {
      const float* rs = p.row_scale + (int64_t)(item0 + it) * p.fold;
      float mn = FLT_MAX;
      int r = tid;
      for (; r + 3 * NTHR < p.fold; r += 4 * NTHR)
        mn = fminf(mn, fminf(fminf(rs[r], rs[r + NTHR]), fminf(rs[r + 2 * NTHR], rs[r + 3 * NTHR])));
      for (; r < p.fold; r += NTHR) mn = fminf(mn, rs[r]);
      mn = wave_min(mn);
      if (lane == 0) s_red[wave] = mn;
      __syncthreads();
      if (tid == 0) {
        float m = s_red[0];
#pragma unroll
        for (int w = 1; w < NTHR / 64; ++w) m = fminf(m, s_red[w]);
        const float tt = ldexpf(m, -p.t_log2k);
        s_t[it] = tt;
        if (bx == 0 && p.t_out != nullptr) p.t_out[item0 + it] = tt;
      }
      __syncthreads();
    }
  }
  __syncthreads();

  int kb = 0;
  for (; kb + PF < nk_full; kb += PF) {
    step(kb, std::integral_constant<int, 0>{}, kFalse{});
    step(kb, std::integral_constant<int, 1>{}, kFalse{});
  }
  for (; kb < nk; kb += PF) {
    step(kb, std::integral_constant<int, 0>{}, kTrue{});
    step(kb, std::integral_constant<int, 1>{}, kTrue{});
  }
  if constexpr (FE > 1) flush();  // k-tiles after the last full flush interval
#pragma unroll
  for (int i = 0; i < TMC; ++i)
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc64[i][e][j][r] += (double)accx[i][e][j][r] * kLoScale;

  // Sum the WK partial results through LDS, one k-group per round (the operand buffers are free
  // after the last barrier of the k loop).
  if constexpr (WK > 1) {
    constexpr int NR = TMC * 2 * TN * 16;
    static_assert((size_t)WM * WN * 64 * NR * sizeof(double) <= (size_t)BUF * sizeof(T),
                  "one k-group of partial sums must fit in an operand buffer");
    double* red = reinterpret_cast<double*>(lds0) + ((wave % (WM * WN)) * 64 + lane) * NR;
#pragma unroll 1
    for (int g = 1; g < WK; ++g) {
      if (wk == g) {
        int q = 0;
#pragma unroll
        for (int i = 0; i < TMC; ++i)
#pragma unroll
          for (int e = 0; e < 2; ++e)
#pragma unroll
            for (int j = 0; j < TN; ++j)
#pragma unroll
              for (int r = 0; r < 16; ++r, ++q) red[q] = acc64[i][e][j][r];
      }
      __syncthreads();
      if (wk == 0) {
        int q = 0;
#pragma unroll
        for (int i = 0; i < TMC; ++i)
#pragma unroll
          for (int e = 0; e < 2; ++e)
#pragma unroll
            for (int j = 0; j < TN; ++j)
#pragma unroll
              for (int r = 0; r < 16; ++r, ++q) acc64[i][e][j][r] += red[q];
      }
      __syncthreads();
    }
    if (wk > 0) return;
  }

  // Epilogue: register r of the Re / Im tiles holds compact row (r/4)*8 + (lane/32)*4 + r%4 of
  // data column lane%32.
  const double inv_z = (p.z_scale != nullptr) ? 1.0 / (double)p.z_scale[z] : 1.0;
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int n = n0 + wn * WTN + j * 32 + (lane & 31);
    if (n >= N) continue;
    double sc = inv_z;
    if (p.row_scale != nullptr) sc /= (double)p.row_scale[n];
    int64_t base, cstride;
    float tt = 1.0f;
    if (p.fold > 0) {
      const int item = n / p.fold;
      base = (int64_t)item * p.fold_stride + (n - item * p.fold);
      cstride = p.ldo;
      if (split_out) tt = s_t[item - item0];
    } else if (p.nc_layout) {
      base = (int64_t)z * p.out_zstride + (int64_t)n * p.ldo;
      cstride = 1;
    } else {
      base = (int64_t)z * p.out_zstride + n;
      cstride = p.ldo;
    }
    const double ar = p.alpha_re * sc, ai = p.alpha_im * sc;
#pragma unroll
    for (int i = 0; i < TMC; ++i) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int c = c0 + wm * WTC + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (c >= Mc) continue;
        const double zr = acc64[i][0][j][r], zi = acc64[i][1][j][r];
        const double orr = zr * ar - zi * ai;
        const double oii = zr * ai + zi * ar;
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
          const double xr = orr * (double)tt;
          const double xi = oii * (double)tt;
          const _Float16 hr = f32_to_f16((float)xr);
          const _Float16 hi = f32_to_f16((float)xi);
          half2_t h = {hr, hi};
          half2_t l = {f32_to_f16((float)((xr - (double)hr) * (double)kLoInv)),
                       f32_to_f16((float)((xi - (double)hi) * (double)kLoInv))};
          reinterpret_cast<half2_t*>(p.out)[idx] = h;
          reinterpret_cast<half2_t*>(p.out_lo)[idx] = l;
        }
      }
    }
  }
}

// LDS DMA issue and wait, written as asm so that the compiler's wait-count pass, which cannot
// tell the LDS stages apart and would wait for every outstanding DMA before any LDS read, leaves
// the pipeline alone.  The DMA writes no VGPR (its data goes to LDS at M0 + 16 * lane); its
// address VGPRs are read at issue.  Kernels using these issue no other vector-memory loads while
// a DMA is outstanding and drain with dma_wait<0>() before any.
// `lds_addr` must be uniform across the wave (it goes to M0).
__device__ __forceinline__ void dma_load16(const void* gaddr, unsigned lds_addr) {
  const unsigned m = __builtin_amdgcn_readfirstlane(lds_addr);
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" : : "v"(gaddr), "s"(m) : "memory", "m0");
}
template <int N>
__device__ __forceinline__ void dma_wait() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  asm volatile("s_waitcnt vmcnt(%0)" : : "n"(N) : "memory");
}

// Split-FP16 stage kernel fed by LDS DMA.
//
// Same arithmetic and output as mft_f16x2_kernel (compact constant, Re / Im MFMA tiles formed in
// registers), but the operand tiles go from HBM/L2 straight into LDS with global_load_lds_dwordx4:
// no VGPR staging and no LDS store instructions.  Three LDS stages (tile t computed while tiles
// t+1 and t+2 are in flight); each row of a stage holds the 16 16-byte chunks of one k-tile with
// the chunk index XOR-ed by (row mod 16), so that the 32 rows a fragment read touches fall in
// distinct banks.  A DMA wave instruction fills one contiguous KiB of LDS (lane i -> bytes 16 i),
// so the swizzle is applied by choosing each lane's source chunk.  The compiler does not track
// LDS DMA, so the waits are explicit: before computing tile t every wave waits until only its
// loads of tile t+1 are outstanding, then one barrier makes tile t visible and frees the stage
// of tile t-1 for the loads of tile t+2.  Requires K to be a multiple of the 128-element k-tile.
template <int WN, int WK>
__global__ __launch_bounds__(64 * WN * WK) void mft_f16x2_dma_kernel(StageArgs p) {
  using T = _Float16;
  constexpr int NTHR = 64 * WN * WK;
  constexpr int NW = NTHR / 64;
  constexpr int BMC = 32;                    // compact constant rows per tile
  constexpr int BN = 32 * WN;                // data rows per tile
  constexpr int BK = 128;
  constexpr int CPR = 16;                    // 16-byte chunks per row per k-tile
  constexpr int A_CH = BMC * CPR, B_CH = BN * CPR;
  constexpr int INSTR = 2 * (A_CH + B_CH) / 64;  // DMA wave instructions per k-tile
  static_assert(INSTR % NW == 0, "DMA instructions must split evenly over the waves");
  constexpr int IPW = INSTR / NW;
  constexpr int OFF_ALO = A_CH * 16, OFF_BHI = 2 * A_CH * 16, OFF_BLO = OFF_BHI + B_CH * 16;
  constexpr int STAGE = 2 * (A_CH + B_CH) * 16;
  constexpr int NS = 3;
  constexpr int KV_PER_WAVE = (BK / 16) / WK;
  constexpr int KPW = KV_PER_WAVE * 16;
  constexpr int FE = (KPW >= kFlushK) ? 1 : kFlushK / KPW;
  static_assert(KV_PER_WAVE >= 1 && (BK / 16) % WK == 0, "k-vectors must split evenly over WK");

  __shared__ __attribute__((aligned(1024))) unsigned char lds[NS * STAGE];
  __shared__ float s_t[BN];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wk = wave / WN;
  const int wn = wave % WN;
  const int Mc = p.M / 2;
  const int gx = (Mc + BMC - 1) / BMC;
  const int gy = (p.N + BN - 1) / BN;
  int bx, by;
  tile_of(blockIdx.x, gx, gy, bx, by);
  const int c0 = bx * BMC;
  const int n0 = by * BN;
  const int z = blockIdx.z;
  const int N = p.N;
  const int nk = p.K / BK;
  const T* Ag[2] = {static_cast<const T*>(p.a), static_cast<const T*>(p.a_lo)};
  const T* Bg[2] = {static_cast<const T*>(p.b) + (int64_t)z * p.b_zstride,
                    static_cast<const T*>(p.b_lo) + (int64_t)z * p.b_zstride};
  const bool split_out = (p.out_mode == kOutSplit);
  int item0 = 0;

  // Split output: per folded item scale t = min(row_scale over the item) * 2^-ceil(log2 K).
  if (split_out) {
    __shared__ float s_red[NTHR / 64];
    const int nend = min(n0 + BN, N);
    item0 = n0 / p.fold;
    const int nitems = (nend - 1) / p.fold - item0 + 1;
    for (int it = 0; it < nitems; ++it) {
      const float* rs = p.row_scale + (int64_t)(item0 + it) * p.fold;
      float mn = FLT_MAX;
      int r = tid;
      for (; r + 3 * NTHR < p.fold; r += 4 * NTHR)
        mn = fminf(mn, fminf(fminf(rs[r], rs[r + NTHR]), fminf(rs[r + 2 * NTHR], rs[r + 3 * NTHR])));
      for (; r < p.fold; r += NTHR) mn = fminf(mn, rs[r]);
      mn = wave_min(mn);
      if (lane == 0) s_red[wave] = mn;
      __syncthreads();
      if (tid == 0) {
        float m = s_red[0];
#pragma unroll
        for (int w = 1; w < NTHR / 64; ++w) m = fminf(m, s_red[w]);
        const float tt = ldexpf(m, -p.t_log2k);
        s_t[it] = tt;
        if (bx == 0 && p.t_out != nullptr) p.t_out[item0 + it] = tt;
      }
      __syncthreads();
    }
  }

  // DMA sources of this wave's instructions: instruction j = wave + u * NW covers LDS chunks
  // 64 j .. 64 j + 63 of a stage (one of the four arrays), lane i the chunk 64 j + i.
  const T* src[IPW];
  int dst[IPW];
#pragma unroll
  for (int u = 0; u < IPW; ++u) {
    const int j = wave + u * NW;
    int q = j * 64 + lane;
    int arr;
    if (q < A_CH) {
      arr = 0;
    } else if (q < 2 * A_CH) {
      arr = 1;
      q -= A_CH;
    } else if (q < 2 * A_CH + B_CH) {
      arr = 2;
      q -= 2 * A_CH;
    } else {
      arr = 3;
      q -= 2 * A_CH + B_CH;
    }
    const int row = q / CPR;
    const int c = (q % CPR) ^ (row & 15);
    if (arr < 2) src[u] = Ag[arr] + (int64_t)(c0 + row) * p.lda + c * 8;
    else src[u] = Bg[arr - 2] + (int64_t)min(n0 + row, N - 1) * p.ldb + c * 8;
    dst[u] = j * 1024;
  }
  const unsigned lds_base = (unsigned)reinterpret_cast<uintptr_t>((__attribute__((address_space(3))) unsigned char*)lds);
  auto issue = [&](int tile, int s) __attribute__((always_inline)) {
    const int k0 = min(tile, nk - 1) * BK;
#pragma unroll
    for (int u = 0; u < IPW; ++u) dma_load16(src[u] + k0, lds_base + (unsigned)(s * STAGE + dst[u]));
  };

  floatx16_t acc[2], accx[2];  // [Re, Im]
  double acc64[2][16];
#pragma unroll
  for (int e = 0; e < 2; ++e)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      acc[e][r] = 0.0f;
      accx[e][r] = 0.0f;
      acc64[e][r] = 0.0;
    }

  const int arow = lane & 31;
  const int brow = wn * 32 + (lane & 31);
  auto frag = [&](const unsigned char* st, int off, int row, int c) __attribute__((always_inline)) {
    return *reinterpret_cast<const half8_t*>(st + off + (row * CPR + (c ^ (row & 15))) * 16);
  };
  auto compute = [&](const unsigned char* st) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < KV_PER_WAVE; ++q) {
      const int c = 2 * (q * WK + wk) + (lane >> 5);
      const half8_t ah = frag(st, 0, arow, c), al = frag(st, OFF_ALO, arow, c);
      const half8_t bh = frag(st, OFF_BHI, brow, c), bl = frag(st, OFF_BLO, brow, c);
      const u32x4_t vh = __builtin_bit_cast(u32x4_t, ah), vl = __builtin_bit_cast(u32x4_t, al);
      const u32x4_t sgn = {0x80000000u, 0x80000000u, 0x80000000u, 0x80000000u};
      const half8_t aeh = __builtin_bit_cast(half8_t, vh ^ sgn), ael = __builtin_bit_cast(half8_t, vl ^ sgn);
      const half8_t aoh = __builtin_bit_cast(half8_t, (vh >> 16) | (vh << 16));
      const half8_t aol = __builtin_bit_cast(half8_t, (vl >> 16) | (vl << 16));
      acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(aeh, bh, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(aoh, bh, acc[1], 0, 0, 0);
      accx[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(aeh, bl, accx[0], 0, 0, 0);
      accx[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(aoh, bl, accx[1], 0, 0, 0);
      accx[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ael, bh, accx[0], 0, 0, 0);
      accx[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(aol, bh, accx[1], 0, 0, 0);
    }
  };
  auto flush = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        acc64[e][r] += (double)acc[e][r];
        acc[e][r] = 0.0f;
      }
  };
  auto step = [&](int kb, auto dc) __attribute__((always_inline)) {
    constexpr int D = decltype(dc)::value;  // kb is a multiple of NS: tile t is in stage D
    const int t = kb + D;
    if (t >= nk) return;
    dma_wait<IPW>();  // own loads of tile t done (those of tile t + 1 may be in flight)
    __syncthreads();    // everyone's loads of tile t done; stage of tile t - 1 no longer read
    issue(t + 2, (D + 2) % NS);
    compute(lds + D * STAGE);
    if (FE == 1 || (t + 1) % FE == 0) flush();
  };

  issue(0, 0);
  issue(1, 1);
  for (int kb = 0; kb < nk; kb += NS) {
    step(kb, std::integral_constant<int, 0>{});
    step(kb, std::integral_constant<int, 1>{});
    step(kb, std::integral_constant<int, 2>{});
  }
  dma_wait<0>();
  __syncthreads();
  if constexpr (FE > 1) flush();
#pragma unroll
  for (int e = 0; e < 2; ++e)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc64[e][r] += (double)accx[e][r] * kLoScale;

  // Sum the WK partial results through LDS, one k-group per round.
  if constexpr (WK > 1) {
    static_assert((size_t)WN * 64 * 32 * sizeof(double) <= (size_t)NS * STAGE, "partial sums fit in LDS");
    double* red = reinterpret_cast<double*>(lds) + (wn * 64 + lane) * 32;
#pragma unroll 1
    for (int g = 1; g < WK; ++g) {
      if (wk == g) {
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
          for (int r = 0; r < 16; ++r) red[e * 16 + r] = acc64[e][r];
      }
      __syncthreads();
      if (wk == 0) {
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
          for (int r = 0; r < 16; ++r) acc64[e][r] += red[e * 16 + r];
      }
      __syncthreads();
    }
    if (wk > 0) return;
  }

  // Epilogue (same output formats as the other stage kernels).
  const int n = n0 + wn * 32 + (lane & 31);
  if (n >= N) return;
  const double inv_z = (p.z_scale != nullptr) ? 1.0 / (double)p.z_scale[z] : 1.0;
  double sc = inv_z;
  if (p.row_scale != nullptr) sc /= (double)p.row_scale[n];
  int64_t base, cstride;
  float tt = 1.0f;
  if (p.fold > 0) {
    const int item = n / p.fold;
    base = (int64_t)item * p.fold_stride + (n - item * p.fold);
    cstride = p.ldo;
    if (split_out) tt = s_t[item - item0];
  } else if (p.nc_layout) {
    base = (int64_t)z * p.out_zstride + (int64_t)n * p.ldo;
    cstride = 1;
  } else {
    base = (int64_t)z * p.out_zstride + n;
    cstride = p.ldo;
  }
  const double ar = p.alpha_re * sc, ai = p.alpha_im * sc;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int c = c0 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
    if (c >= Mc) continue;
    const double zr = acc64[0][r], zi = acc64[1][r];
    const double orr = zr * ar - zi * ai;
    const double oii = zr * ai + zi * ar;
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
      const double xr = orr * (double)tt;
      const double xi = oii * (double)tt;
      const _Float16 hr = f32_to_f16((float)xr);
      const _Float16 hi = f32_to_f16((float)xi);
      half2_t h = {hr, hi};
      half2_t l = {f32_to_f16((float)((xr - (double)hr) * (double)kLoInv)),
                   f32_to_f16((float)((xi - (double)hi) * (double)kLoInv))};
      reinterpret_cast<half2_t*>(p.out)[idx] = h;
      reinterpret_cast<half2_t*>(p.out_lo)[idx] = l;
    }
  }
}

// Split formats of the data operand: FP16 hi/lo (x = hi + lo * 2^-11) or BF16 b0/b1/b2
// (x = b0 + b1 + b2, each the bfloat16 rounding of the remaining remainder).
enum SplitFmt : int { kSplitF16x2 = 0, kSplitBF16x3 = 1 };

// Splits 4 consecutive floats (already multiplied by the row scale) and stores them at element k
// of each split row (8-byte stores).
template <int FMT>
__device__ inline void split_store4(floatx4_t v, void* const* outs, int64_t row_off, int k) {
  if constexpr (FMT == kSplitF16x2) {
    typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
    half4_t vh, vl;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const _Float16 q = (_Float16)v[e];
      vh[e] = q;
      vl[e] = (_Float16)((v[e] - (float)q) * kLoInv);
    }
    *reinterpret_cast<half4_t*>(static_cast<_Float16*>(outs[0]) + row_off + k) = vh;
    *reinterpret_cast<half4_t*>(static_cast<_Float16*>(outs[1]) + row_off + k) = vl;
  } else {
    typedef __bf16 bf4_t __attribute__((ext_vector_type(4)));
    bf4_t b0, b1, b2;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      // v - b0 and r - b1 are exact in FP32 (b0, b1 keep the leading 8 bits of their operands)
      const __bf16 q0 = (__bf16)v[e];
      const float r = v[e] - (float)q0;
      const __bf16 q1 = (__bf16)r;
      b0[e] = q0;
      b1[e] = q1;
      b2[e] = (__bf16)(r - (float)q1);
    }
    *reinterpret_cast<bf4_t*>(static_cast<__bf16*>(outs[0]) + row_off + k) = b0;
    *reinterpret_cast<bf4_t*>(static_cast<__bf16*>(outs[1]) + row_off + k) = b1;
    *reinterpret_cast<bf4_t*>(static_cast<__bf16*>(outs[2]) + row_off + k) = b2;
  }
}

// Splits rows of interleaved complex64 data into the split arrays of format FMT, with a
// power-of-two row scale s = 2^-e chosen so that max|x| * s < 1.  One workgroup per row, any
// length and alignment; row padding up to ld_out is zero-filled.
template <int FMT>
__global__ __launch_bounds__(256) void split_rows_kernel(const float* __restrict__ in, int64_t ld_in, int K,
                                                         SplitOut o, int64_t ld_out, float* __restrict__ row_scale) {
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
  // ld_out is a multiple of 64 elements, so k0 + 4 <= ld_out
  for (int k0 = tid * 4; k0 < ld_out; k0 += 256 * 4) {
    floatx4_t v;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = (k0 + j < K) ? x[k0 + j] * s : 0.0f;
    split_store4<FMT>(v, o.p, row * ld_out, k0);
  }
  if (tid == 0) row_scale[row] = s;
}

// Fast path of the split: one wave per row, the whole row (K <= 64 * PL floats) in registers,
// 16-byte loads, 8-byte stores, a wave-only max reduction.  Needs 16-byte aligned rows with
// K and ld_in multiples of 4.
template <int PL, int FMT>
__global__ __launch_bounds__(256) void split_rows_wave_kernel(const float* __restrict__ in, int64_t ld_in, int K,
                                                              SplitOut o, int64_t ld_out,
                                                              float* __restrict__ row_scale, int64_t rows) {
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
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const int k = (j * 64 + lane) * 4;
    if (k < ld_out) split_store4<FMT>(v[j] * s, o.p, row * ld_out, k);
  }
  // zero padding beyond the register-held chunks
  for (int k = NV * 256 + lane * 4; k < ld_out; k += 256) split_store4<FMT>(floatx4_t{0.f, 0.f, 0.f, 0.f}, o.p, row * ld_out, k);
  if (lane == 0) row_scale[row] = s;
}

// Compact Fourier matrix: m(c, k) = exp(i * sign * a[c] * b[k]), phase product and sincos in FP64,
// stored as interleaved (re, im) rows of lda reals (the stage kernel expands it per fragment), in
// the operand format of `prec` (split precisions write their 2 or 3 arrays).
// One thread per complex entry; entries outside (nc, nk) are 0.
__global__ void gen_matrix_kernel(const double* __restrict__ a, int nc, const double* __restrict__ b, int nk,
                                  double sign, int prec, SplitOut out, int64_t lda, int rows_pad) {
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
      double* q = static_cast<double*>(out.p[0]);
      q[o] = mr;
      q[o + 1] = mi;
    } else if (prec == kF32) {
      float* q = static_cast<float*>(out.p[0]);
      q[o] = (float)mr;
      q[o + 1] = (float)mi;
    } else if (prec == kF16x2) {
      _Float16* oh = static_cast<_Float16*>(out.p[0]);
      _Float16* ol = static_cast<_Float16*>(out.p[1]);
      const _Float16 hr = (_Float16)mr, hi = (_Float16)mi;
      oh[o] = hr;
      oh[o + 1] = hi;
      ol[o] = (_Float16)((mr - (double)hr) * (double)kLoInv);
      ol[o + 1] = (_Float16)((mi - (double)hi) * (double)kLoInv);
    } else {
      const double v[2] = {mr, mi};
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const __bf16 q0 = f32_to_bf16((float)v[e]);
        const double r = v[e] - (double)(float)q0;
        const __bf16 q1 = f32_to_bf16((float)r);
        static_cast<__bf16*>(out.p[0])[o + e] = q0;
        static_cast<__bf16*>(out.p[1])[o + e] = q1;
        static_cast<__bf16*>(out.p[2])[o + e] = f32_to_bf16((float)(r - (double)(float)q1));
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Parity form of the 2-D transform between grids that are symmetric about 0 on every axis.
//
// Each axis is described by magnitudes P[j] >= 0 and the grid indices of the samples at +P[j]
// and -P[j] (either may be absent: the origin has no separate mirror, and an FFT-style grid's
// first sample -n/2 d has none).  With y', x' indexing the magnitudes of the input axes, the
// four mirror samples p = E(+y', +x'), q = E(+y', -x'), r = E(-y', +x'), s = E(-y', -x') (absent
// samples count as 0) give
//   F_ee = p + q + r + s,  F_eo = p - q + r - s,  F_oe = p + q - r - s,  F_oo = p - q - r + s
// (first letter the y parity, second the x parity), and with C = cos(u x), S = sin(u x) per axis
//   H_ab[v'][u'] = sum_y' Y_a[v'][y'] sum_x' X_b[u'][x'] F_ab[y'][x']      (Y_e/X_e = C, Y_o/X_o = S)
//   O(sv v', su u') = H_ee + i s su H_eo + i s sv H_oe - su sv H_oo      (s = -1 forward, +1 backward)
// The eight real half-size products do about a quarter of the multiply-adds of the complex form.

// Sum (MAX false) or maximum (MAX true) of eight values per lane over the 64 lanes of a wave, by
// halving exchanges: lanes keep half of their values at each of the offsets 32, 16, 8 and send the
// other half, then the one value left is reduced over offsets 4, 2, 1 (ten exchanges instead of
// 48).  Afterwards every lane l holds the result for value index l >> 3.
template <bool MAX>
__device__ inline double wave_reduce8(const double (&v)[8], int lane) {
  auto op = [](double a, double b) __attribute__((always_inline)) { return MAX ? fmax(a, b) : a + b; };
  const bool b5 = (lane & 32) != 0, b4 = (lane & 16) != 0, b3 = (lane & 8) != 0;
  double k4[4], k2[2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double keep = b5 ? v[4 + i] : v[i];
    const double send = b5 ? v[i] : v[4 + i];
    k4[i] = op(keep, __shfl_xor(send, 32));
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double keep = b4 ? k4[2 + i] : k4[i];
    const double send = b4 ? k4[i] : k4[2 + i];
    k2[i] = op(keep, __shfl_xor(send, 16));
  }
  double k1 = op(b3 ? k2[1] : k2[0], __shfl_xor(b3 ? k2[0] : k2[1], 8));
  k1 = op(k1, __shfl_xor(k1, 4));
  k1 = op(k1, __shfl_xor(k1, 2));
  k1 = op(k1, __shfl_xor(k1, 1));
  return k1;
}

// Splits NV (2 or 4) FP64 values (already scaled into (-1, 1)) into the split arrays of format FMT
// at element k of a row (one vector store per array): hi = fp16(v), lo = fp16((v - hi) * 2^11), or
// the three bfloat16 parts of the remainders, all taken from the FP64 value.
template <int FMT, int NV>
__device__ inline void split_store_d(const double* v, void* const* outs, int64_t row_off, int k) {
  if constexpr (FMT == kSplitF16x2) {
    typedef _Float16 hv_t __attribute__((ext_vector_type(NV)));
    hv_t vh, vl;
#pragma unroll
    for (int e = 0; e < NV; ++e) {
      const _Float16 q = f32_to_f16((float)v[e]);
      vh[e] = q;
      vl[e] = f32_to_f16((float)((v[e] - (double)q) * (double)kLoInv));
    }
    __builtin_nontemporal_store(vh, reinterpret_cast<hv_t*>(static_cast<_Float16*>(outs[0]) + row_off + k));
    __builtin_nontemporal_store(vl, reinterpret_cast<hv_t*>(static_cast<_Float16*>(outs[1]) + row_off + k));
  } else {
    typedef __bf16 bv_t __attribute__((ext_vector_type(NV)));
    bv_t b0, b1, b2;
#pragma unroll
    for (int e = 0; e < NV; ++e) {
      const __bf16 q0 = f32_to_bf16((float)v[e]);
      const double r = v[e] - (double)(float)q0;
      const __bf16 q1 = f32_to_bf16((float)r);
      b0[e] = q0;
      b1[e] = q1;
      b2[e] = f32_to_bf16((float)(r - (double)(float)q1));
    }
    *reinterpret_cast<bv_t*>(static_cast<__bf16*>(outs[0]) + row_off + k) = b0;
    *reinterpret_cast<bv_t*>(static_cast<__bf16*>(outs[1]) + row_off + k) = b1;
    *reinterpret_cast<bv_t*>(static_cast<__bf16*>(outs[2]) + row_off + k) = b2;
  }
}

// Parity split of `batch` c64 fields [ny][nx]: item z = 4 b + parity (ee, eo, oe, oo) holds rows
// 2 y' (real parts) and 2 y' + 1 (imaginary parts) over x' < ax.h, each multiplied by a
// power-of-two scale (max |value| * scale < 1) and split; row padding up to ld_out is zero.  The
// parity sums are formed in FP64 (exact for FP32 inputs) and split from FP64.
// With nyq_out set, the split also forms one row of the stage-1 product in FP64: for every row
// (z, 2 y' + c), sum_x' X_b[x'] F_ab[y'][x'] with X_e = nyq_coef[x'] (cosines) and X_o =
// nyq_coef[ax.h + x'] (sines), written unscaled to nyq_out[z * 2 ay.h + 2 y' + c].
// A workgroup of 256 threads takes 256 / TPR values of y' (TPR threads each); a thread holds the
// four mirror samples of XPT consecutive x' in registers (XPT * TPR >= ld_out), so the field is
// read once.
// Body of the split for row block `blk` of `nblk` (one workgroup of 256 threads) of field b.
template <int FMT, int XPT, int TPR>
__device__ __forceinline__ void parity_split_body(const float2* __restrict__ in, int nx, int ny, const ParAxis& ax,
                                                  const ParAxis& ay, const SplitOut& o, int64_t ld_out,
                                                  float* __restrict__ row_scale, const double* __restrict__ nyq_coef,
                                                  double* __restrict__ nyq_out, int blk, int nblk, int64_t b) {
  constexpr int SV = XPT % 4 == 0 ? 4 : 2;  // elements per store
  static_assert(XPT % SV == 0 && (TPR == 64 || TPR == 128 || TPR == 256), "thread layout");
  constexpr int TEAMS = 256 / TPR, WPT = TPR / 64;
  __shared__ double red[TEAMS][WPT][8];
  __shared__ double red_n[TEAMS][WPT][8];
  const int team = threadIdx.x / TPR;
  const int tt = threadIdx.x % TPR;
  // XCD-aware row order: consecutive workgroups go round-robin over the 8 XCDs, and XCD x takes
  // the x-th eighth of the y' rows, the rows the stage-1 workgroups on that XCD read (their tile
  // order gives each XCD a contiguous range of data rows)
  const int rb = (nblk % 8 == 0) ? (blk & 7) * (nblk >> 3) + (blk >> 3) : blk;
  const int yp = rb * TEAMS + team;
  const int hx = ax.h, hy = ay.h;
  const bool live = yp < hy;
  const int jp = live ? par_plus(ay, yp) : -1, jm = live ? par_minus(ay, yp) : -1;
  const float2* rp = in + (b * ny + (jp >= 0 ? jp : 0)) * (int64_t)nx;
  const float2* rm = in + (b * ny + (jm >= 0 ? jm : 0)) * (int64_t)nx;
  const int x0 = tt * XPT;
  float2 P[XPT], Q[XPT], R[XPT], W[XPT];
  const float2 zero = make_float2(0.f, 0.f);
  // indices first, then every load unconditionally at a clamped index (so that they issue back to
  // back), then the absent samples are zeroed by selects
  int cps[XPT], cms[XPT];
#pragma unroll
  for (int j = 0; j < XPT; ++j) {
    const int xp = min(x0 + j, hx - 1);
    cps[j] = x0 + j < hx ? par_plus(ax, xp) : -1;
    cms[j] = x0 + j < hx ? par_minus(ax, xp) : -1;
  }
#pragma unroll
  for (int j = 0; j < XPT; ++j) {
    const int cp = max(cps[j], 0), cm = max(cms[j], 0);
    P[j] = rp[cp];
    Q[j] = rp[cm];
    R[j] = rm[cp];
    W[j] = rm[cm];
  }
  // coefficients of the extra stage-1 row (issued with the data loads)
  const bool nyq = nyq_out != nullptr;
  double ncs[XPT], nsn[XPT];
#pragma unroll
  for (int j = 0; j < XPT; ++j) {
    const int xj = min(x0 + j, hx - 1);
    ncs[j] = nyq ? nyq_coef[xj] : 0.0;
    nsn[j] = nyq ? nyq_coef[hx + xj] : 0.0;
  }
#pragma unroll
  for (int j = 0; j < XPT; ++j) {
    const bool vp = cps[j] >= 0, vm = cms[j] >= 0;
    P[j] = (jp >= 0 && vp) ? P[j] : zero;
    Q[j] = (jp >= 0 && vm) ? Q[j] : zero;
    R[j] = (jm >= 0 && vp) ? R[j] : zero;
    W[j] = (jm >= 0 && vm) ? W[j] : zero;
  }
  // the eight values (parity ab, part c) at index 2 ab + c, in FP64
  auto combos = [&](int j, double* v) __attribute__((always_inline)) {
    const double pqx = (double)P[j].x + Q[j].x, pqy = (double)P[j].y + Q[j].y;
    const double pmx = (double)P[j].x - Q[j].x, pmy = (double)P[j].y - Q[j].y;
    const double rwx = (double)R[j].x + W[j].x, rwy = (double)R[j].y + W[j].y;
    const double rmx = (double)R[j].x - W[j].x, rmy = (double)R[j].y - W[j].y;
    v[0] = pqx + rwx; v[1] = pqy + rwy;  // ee
    v[2] = pmx + rmx; v[3] = pmy + rmy;  // eo
    v[4] = pqx - rwx; v[5] = pqy - rwy;  // oe
    v[6] = pmx - rmx; v[7] = pmy - rmy;  // oo
  };
  double mx[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < XPT; ++j) {
    double v[8];
    combos(j, v);
#pragma unroll
    for (int e = 0; e < 8; ++e) mx[e] = fmax(mx[e], fabs(v[e]));
  }
  const int wv = tt >> 6;
  const int lane = tt & 63;
  {
    const double m = wave_reduce8<true>(mx, lane);
    if ((lane & 7) == 0) red[team][wv][lane >> 3] = m;
  }
  if (nyq) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < XPT; ++j) {
      double v[8];
      combos(j, v);
#pragma unroll
      for (int e = 0; e < 8; ++e) acc[e] += v[e] * (((e >> 1) & 1) ? nsn[j] : ncs[j]);  // x parity: cos / sin
    }
    const double a = wave_reduce8<false>(acc, lane);
    if ((lane & 7) == 0) red_n[team][wv][lane >> 3] = a;
  }
  __syncthreads();
  if (!live) return;
  if (nyq && tt < 8) {
    double a = red_n[team][0][tt];
    for (int w = 1; w < WPT; ++w) a += red_n[team][w][tt];
    nyq_out[(b * 4 + (tt >> 1)) * 2 * (int64_t)hy + 2 * yp + (tt & 1)] = a;
  }
  double sc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    double m = red[team][0][e];
#pragma unroll
    for (int w = 1; w < WPT; ++w) m = fmax(m, red[team][w][e]);
    double s = 1.0;
    if (m > 0.0 && m <= DBL_MAX) s = ldexp(1.0, -(ilogb(m) + 1));  // m * s in [0.5, 1)
    sc[e] = s;
  }
  if (tt < 8) {
    double v = 1.0;
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (e == tt) v = sc[e];  // a static register index, no private-memory array
    row_scale[((b * 4 + (tt >> 1)) * hy + yp) * 2 + (tt & 1)] = (float)v;
  }
  if (x0 >= ld_out) return;
#pragma unroll
  for (int j0 = 0; j0 < XPT; j0 += SV) {
    if (x0 + j0 >= ld_out) break;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      double w[SV];
#pragma unroll
      for (int j = 0; j < SV; ++j) {
        double v[8];
        combos(j0 + j, v);  // recomputed per row: a few adds, fewer live registers
        w[j] = v[e] * sc[e];
      }
      const int64_t row = ((b * 4 + (e >> 1)) * hy + yp) * 2 + (e & 1);
      split_store_d<FMT, SV>(w, o.p, row * ld_out, x0 + j0);
    }
  }
}

// One workgroup per row block (x) and field (y).
template <int FMT, int XPT, int TPR>
__global__ __launch_bounds__(256) void parity_split_kernel(const float2* __restrict__ in, int nx, int ny, ParAxis ax,
                                                           ParAxis ay, SplitOut o, int64_t ld_out,
                                                           float* __restrict__ row_scale,
                                                           const double* __restrict__ nyq_coef,
                                                           double* __restrict__ nyq_out) {
  parity_split_body<FMT, XPT, TPR>(in, nx, ny, ax, ay, o, ld_out, row_scale, nyq_coef, nyq_out, blockIdx.x, gridDim.x,
                                   blockIdx.y);
}

// The same split for up to kSplitItemsMax single fields of one geometry: workgroup row blockIdx.y
// takes item blockIdx.y.
template <int FMT, int XPT, int TPR>
__global__ __launch_bounds__(256) void parity_split_items_kernel(SplitItems it, int nx, int ny, ParAxis ax, ParAxis ay,
                                                                 int64_t ld_out) {
  const int i = blockIdx.y;
  parity_split_body<FMT, XPT, TPR>(static_cast<const float2*>(it.in[i]), nx, ny, ax, ay, it.o[i], ld_out, it.rs[i],
                                   nullptr, nullptr, blockIdx.x, gridDim.x, 0);
}

// Real parity matrices of one stage, two planes of rows_pad x lda: plane 0 cos(c[i] k[j]),
// plane 1 sin(c[i] k[j]) for i < nc, j < nk (c, k: the positive halves of the row and
// contraction coordinates), zero elsewhere; phase and sincos in FP64, stored in the operand
// format of `prec`.
__global__ void gen_parity_kernel(const double* __restrict__ c, int nc, const double* __restrict__ k, int nk, int prec,
                                  SplitOut out, int64_t lda, int rows_pad) {
  const int64_t plane = (int64_t)rows_pad * lda;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < 2 * plane;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int pl = (int)(idx / plane);
    const int64_t rem = idx - pl * plane;
    const int i = (int)(rem / lda);
    const int j = (int)(rem - (int64_t)i * lda);
    double v = 0.0;
    if (i < nc && j < nk) {
      double sn, cs;
      sincos(c[i] * k[j], &sn, &cs);
      v = pl ? sn : cs;
    }
    if (prec == kF64) {
      static_cast<double*>(out.p[0])[idx] = v;
    } else if (prec == kF32) {
      static_cast<float*>(out.p[0])[idx] = (float)v;
    } else if (prec == kF16x2) {
      const _Float16 h = (_Float16)v;
      static_cast<_Float16*>(out.p[0])[idx] = h;
      static_cast<_Float16*>(out.p[1])[idx] = (_Float16)((v - (double)h) * (double)kLoInv);
    } else {
      const __bf16 q0 = f32_to_bf16((float)v);
      const double r = v - (double)(float)q0;
      const __bf16 q1 = f32_to_bf16((float)r);
      static_cast<__bf16*>(out.p[0])[idx] = q0;
      static_cast<__bf16*>(out.p[1])[idx] = q1;
      static_cast<__bf16*>(out.p[2])[idx] = f32_to_bf16((float)(r - (double)(float)q1));
    }
  }
}

// Recombination of the four parity products H (item 4 b + ab: au.h ... rows av.h of au.h complex
// FP64 values) into output field b ([nv][nu]) at the grid indices of +-v', +-u' that exist, times
// alpha, in the output format.
__global__ __launch_bounds__(256) void parity_combine_kernel(const double2* __restrict__ H, ParAxis au, ParAxis av,
                                                             int nu, int nv, double sign, double ar, double ai,
                                                             int out_mode, void* __restrict__ out, int64_t batch) {
  const int hu = au.h, hv = av.h;
  const int64_t per = (int64_t)hu * hv;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= batch * per) return;
  const int64_t b = idx / per;
  const int64_t rem = idx - b * per;
  const int vp = (int)(rem / hu);
  const int up = (int)(rem - (int64_t)vp * hu);
  const double2* Hb = H + b * 4 * per + rem;
  parity_store(Hb[0], Hb[per], Hb[2 * per], Hb[3 * per], vp, up, au, av, nu, sign, ar, ai, out_mode, out,
               b * (int64_t)nv * nu);
}

// ---------------------------------------------------------------------------------------------
// The parity form as one persistent launch (FusedArgs in mft_gemm.hpp describes the task order
// and the sync block).

// Longest wait for a dependency, in ticks of the 100 MHz real-time counter (s_memrealtime): 10 ms,
// some 300 times a whole 1024^2 -> 512^2 execution, so it only ends a wait that can never be
// satisfied.  The bound is in time, not in polls: a poll's latency grows with the number of
// workgroups polling the same line.
constexpr uint64_t kWaitTicks = 1000000;

// Tasks with linear index = q mod 8 among the first n of a level.
__device__ inline int64_t queue_share(int64_t n, int q) { return n > q ? (n - q + 7) / 8 : 0; }

// Control flow of the persistent loop is kept uniform: a branch on threadIdx.x == 0 around the
// queue grab, the counter increments or the polls lets the compiler route the other lanes past it,
// and with the loop's barriers that made lanes 1..63 of wave 0 and waves 1..3 return to the
// barrier and reread the task index before thread 0 took the next task (the previous task repeated
// forever).  So every lane of wave 0 runs these steps (wave 0 is selected by a scalar), each atomic
// is issued by all 64 lanes with lane 0 alone adding 1, and loop conditions are readfirstlane'd.
__device__ inline bool wave0() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0; }

// Adds 1 to *c (all lanes of wave 0 call it); returns the value before the add.
__device__ inline unsigned wave_add1(unsigned* c, int memorder) {
  const unsigned one = (threadIdx.x & 63) == 0 ? 1u : 0u;
  unsigned old;
  if (memorder == __ATOMIC_RELEASE) old = __hip_atomic_fetch_add(c, one, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  else old = __hip_atomic_fetch_add(c, one, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __builtin_amdgcn_readfirstlane(old);  // lane 0's: the value before its own add
}

// The workgroup's stores become visible at agent scope (every XCD), then the counter is raised.
__device__ inline void fused_signal(unsigned* c) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (wave0()) (void)wave_add1(c, __ATOMIC_RELEASE);
}

// Wave 0 polls until every counter c[i * kSyncLine] (i < nc) reaches target(i), then acquires at
// agent scope; the barrier orders the workgroup's later loads after it.  A wait longer than
// kWaitTicks counts a timeout in `err` and returns.  Polls back off (about 0.25 us apart).
template <typename TargetFn>
__device__ inline void fused_wait(const unsigned* c, int nc, TargetFn target, unsigned* err) {
  if (wave0()) {
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    bool late = false;
    for (int i = 0; i < nc && !late; ++i) {
      const unsigned want = target(i);
      for (;;) {
        const unsigned seen = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(c + i * kSyncLine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (seen >= want) break;
        if (__builtin_amdgcn_s_memrealtime() - t0 > kWaitTicks) {
          late = true;
          break;
        }
        __builtin_amdgcn_s_sleep(8);
      }
    }
    if (late) (void)wave_add1(err, __ATOMIC_RELAXED);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Split-FP16 parity execution: split (FMT, XPT, TPR as the split kernel), stage 1 on 64 x 64 tiles
// with 64-deep k-tiles, stage 2 recombining 32 x 32 tiles of the four items.
template <int XPT, int TPR>
__global__ __launch_bounds__(256) void parity_fused_kernel(FusedArgs f) {
  constexpr int TEAMS = 256 / TPR;
  (void)TEAMS;
  const int q = blockIdx.x & 7;
  unsigned* const sync = f.sync;
  unsigned* const err = sync + 17 * kSyncLine;
  const int gx1 = (f.s1.M + 63) / 64, gy1 = (f.s1.N + 63) / 64, t1 = gx1 * gy1;
  const int gx2 = (f.s2.M + 31) / 32, gy2 = (f.s2.N + 31) / 32, t2 = gx2 * gy2;
  const int64_t c0 = queue_share(f.n0, q), c1 = queue_share(f.n1, q), c2 = queue_share(f.n2, q);
  const int64_t nfield = f.n2 / t2;
  __shared__ unsigned s_task;
  for (;;) {
    if (wave0()) s_task = wave_add1(sync + q * kSyncLine, __ATOMIC_RELAXED);  // every lane stores the same
    __syncthreads();
    int64_t j = (int64_t)__builtin_amdgcn_readfirstlane(s_task);  // a scalar: the branches below are uniform
    __syncthreads();
    if (j < c0) {
      const int64_t L = q + 8 * j;
      const int64_t b = L / f.nblk;
      parity_split_body<kSplitF16x2, XPT, TPR>(static_cast<const float2*>(f.in), f.nx, f.ny, f.ax, f.ay, f.fo,
                                               f.ld_out, f.row_scale, f.nyq_coef, f.nyq_out, (int)(L - b * f.nblk),
                                               f.nblk, b);
      fused_signal(sync + (8 + q) * kSyncLine);
    } else if ((j -= c0) < c1) {
      // every split: the per-item scale of the intermediate is the minimum over all row scales
      fused_wait(sync + 8 * kSyncLine, 8, [&](int i) { return (unsigned)queue_share(f.n0, i); }, err);
      const int64_t L = q + 8 * j;
      const int z = (int)(L / t1);
      const int h = (int)(L - (int64_t)z * t1);
      const int so = (int)((L - h) & 7);
      stage_body<kF16x2, 64, 64, 64, 2, 2, 1, kRealPadded, 2>(f.s1, h, z, so);
      int bx, by;
      tile_of(h, gx1, gy1, bx, by, so);
      fused_signal(sync + (kSyncFixedLines + bx) * kSyncLine);
    } else if ((j -= c1) < c2) {
      const int64_t L = q + 8 * j;
      const int z = (int)(L / t2);
      const int h = (int)(L - (int64_t)z * t2);
      const int so = (int)((L - h) & 7);
      int bx, by;
      tile_of(h, gx2, gy2, bx, by, so);
      // the stage-1 row tile holding this tile's u' (16 of them, inside one 64-row tile), and row
      // tile 0, which writes the item scales and the extra row of an FFT-style axis
      const int u0 = 16 * by;
      const int rt = u0 < f.s1.M ? u0 / 64 : 0;
      const unsigned want = (unsigned)(4 * nfield * gy1);
      unsigned* const c = sync + kSyncFixedLines * kSyncLine;
      fused_wait(c, 1, [&](int) { return want; }, err);
      if (rt != 0) fused_wait(c + rt * kSyncLine, 1, [&](int) { return want; }, err);
      stage_body<kF16x2, 64, 128, 64, 2, 2, 1, kRealComb, 2>(f.s2, h, z, so);
      __syncthreads();
    } else {
      break;
    }
  }
  // the last workgroup to finish clears the counters for the next execution
  if (wave0()) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned done = wave_add1(sync + 16 * kSyncLine, __ATOMIC_RELAXED);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (done == gridDim.x - 1) {
      for (int i = 0; i < 16; ++i) __hip_atomic_store(sync + i * kSyncLine, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int i = 0; i < gx1; ++i)
        __hip_atomic_store(sync + (kSyncFixedLines + i) * kSyncLine, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(sync + 16 * kSyncLine, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Host-side launchers.

// LDS behind the operand buffers for the extra constant row of a real split-output stage (one
// float per k, whole k-tiles): at most kNyqLdsMax, i.e. contractions of up to 4096 values.
constexpr size_t kNyqLdsMax = 16 * 1024;
inline size_t nyq_lds_bytes(int K, int BK) { return K > 0 ? (size_t)ceil_div(K, BK) * BK * sizeof(float) : 0; }

template <int P, int BM, int BN, int BK, int BMODE>
constexpr size_t stage_lds_max() {
  using T = typename PrecTraits<P>::T;
  constexpr int VEC = 16 / (int)sizeof(T);
  constexpr size_t lds = 2 * (size_t)PrecTraits<P>::NARR * (BM + BN) * (BK + VEC) * sizeof(T);
  constexpr size_t budget = 156 * 1024;  // one CU's LDS, less the kernel's static arrays
  constexpr size_t want = lds + ((BMODE == kRealPadded && PrecTraits<P>::SPLIT) ? kNyqLdsMax : 0);
  return want < budget ? want : (lds < budget ? budget : lds);
}

template <int P, int BM, int BN, int BK, int WM, int WN, int WK, int BMODE, int PF>
static int launch_stage_t(const StageArgs& a, int nz, hipStream_t s) {
  using T = typename PrecTraits<P>::T;
  constexpr int VEC = 16 / (int)sizeof(T);
  constexpr int NARR = PrecTraits<P>::NARR;
  constexpr size_t lds = 2 * (size_t)NARR * (BM + BN) * (BK + VEC) * sizeof(T);
  static_assert(lds <= 156 * 1024, "LDS budget of one CU (with the kernel's static arrays)");
  auto kern = mft_stage_kernel<P, BM, BN, BK, WM, WN, WK, BMODE, PF>;
  static bool configured = false;  // per instantiation
  if (!configured) {
    HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
    configured = true;
  }
  constexpr int TBM = BMODE == kRealComb ? 32 : BM;  // output tile of a workgroup
  constexpr int TBN = BMODE == kRealComb ? 32 : BN;
  const int gx = (int)ceil_div(a.M, TBM);
  const int gy = (int)ceil_div(a.N, TBN);
  HCIPY_REQUIRE(a.nyq_coef == nullptr, "single stage launches take split-formed extra rows only");
  hipLaunchKernelGGL(kern, dim3((unsigned)(gx * gy), 1, (unsigned)nz), dim3(64 * WM * WN * WK), lds, s, a);
  return kOk;
}

template <int BMC, int BN, int BK, int WM, int WN, int WK>
static int launch_f16x2_t(const StageArgs& a, int nz, hipStream_t s) {
  static_assert(2 * (size_t)(2 * (BMC + BN) * (BK + 8)) * sizeof(_Float16) <= 150 * 1024, "LDS budget of one CU");
  auto kern = mft_f16x2_kernel<BMC, BN, BK, WM, WN, WK, 2>;  // operand buffers are static LDS
  const int gx = (int)ceil_div(a.M / 2, BMC);
  const int gy = (int)ceil_div(a.N, BN);
  hipLaunchKernelGGL(kern, dim3((unsigned)(gx * gy), 1, (unsigned)nz), dim3(64 * WM * WN * WK), 0, s, a);
  return kOk;
}

template <int WN, int WK>
static int launch_f16x2_dma_t(const StageArgs& a, int nz, hipStream_t s) {
  const int gx = (int)ceil_div(a.M / 2, 32);
  const int gy = (int)ceil_div(a.N, 32 * WN);
  hipLaunchKernelGGL((mft_f16x2_dma_kernel<WN, WK>), dim3((unsigned)(gx * gy), 1, (unsigned)nz), dim3(64 * WN * WK), 0,
                     s, a);
  return kOk;
}

bool tile_cfg_supported(int prec, int cfg) {
  if (cfg < 0 || cfg >= kR64x64_w4) return false;  // the complex form's shapes end before the parity ones
  // BF16x3: three operand arrays per matrix; the shapes whose LDS buffers and thread split fit
  if (prec == kBF16x3) return cfg == k64x64_w2x2 || cfg == k64x32_w2x1 || cfg == k128x64 || cfg == k64x32_w8;
  if (cfg >= kC64x64_w4) return prec == kF16x2;  // compact-constant shapes: split-FP16 only
  // FP64: 64 x 64 wave tiles need more LDS for the k-split reduction than the operand buffers
  // hold, and its 32-deep k-tiles have too few k-vectors for a 4-way split
  return prec != kF64 || cfg == k64x64_w2x2 || cfg == k64x32_w2x1 || cfg == k64x64_w8;
}

// Default tile shape: 64 x 64 tiles when they give at least 256 workgroups (one per CU), else
// 64 x 32 tiles.
// FP32 / split-FP16: the 8-wave shapes (two waves per SIMD hide each other's LDS and load
// latencies; measured fastest at 1024^2 -> 512^2).  FP64 keeps the 4-wave shapes.
template <int P>
static int default_tile_cfg(const StageArgs& a, int nz) {
  const int64_t tiles64 = ceil_div(a.M, 64) * ceil_div(a.N, 64) * nz;
  if constexpr (P == kF64) return tiles64 >= 256 ? k64x64_w2x2 : k64x32_w2x1;
  return tiles64 >= 256 ? k64x64_w8 : k64x32_w8;
}

template <int P, int BK, int BMODE, int PF>
static int launch_prec(const StageArgs& a, int nz, hipStream_t s) {
  const int cfg = a.tile_cfg == kTileAuto ? default_tile_cfg<P>(a, nz) : a.tile_cfg;
  switch (cfg) {
    case k64x64_w2x2: return launch_stage_t<P, 64, 64, BK, 2, 2, 1, BMODE, PF>(a, nz, s);
    case k64x32_w2x1: return launch_stage_t<P, 64, 32, BK, 2, 1, 2, BMODE, PF>(a, nz, s);
    case k64x64_w8: return launch_stage_t<P, 64, 64, BK, 2, 2, 2, BMODE, PF>(a, nz, s);
    default: break;
  }
  if constexpr (P == kF16x2) {
    if (cfg == kC64x64_w4) return launch_f16x2_t<32, 64, BK, 1, 2, 2>(a, nz, s);
    if (cfg == kC64x64_w8) return launch_f16x2_t<32, 64, BK, 1, 2, 4>(a, nz, s);
    if (cfg == kC64x32_w4) return launch_f16x2_t<32, 32, BK, 1, 1, 4>(a, nz, s);
    if (cfg == kC128x32_w8) return launch_f16x2_t<64, 32, BK, 2, 1, 4>(a, nz, s);
    if (cfg >= kD64x64_w4 && cfg <= kD64x32_w8) {
      if (a.K % 128 != 0)  // the DMA kernels need whole k-tiles: same arithmetic, compact kernel
        return cfg <= kD64x64_w8 ? launch_f16x2_t<32, 64, BK, 1, 2, 2>(a, nz, s)
                                 : launch_f16x2_t<32, 32, BK, 1, 1, 4>(a, nz, s);
      if (cfg == kD64x64_w4) return launch_f16x2_dma_t<2, 2>(a, nz, s);
      if (cfg == kD64x64_w8) return launch_f16x2_dma_t<2, 4>(a, nz, s);
      if (cfg == kD64x32_w4) return launch_f16x2_dma_t<1, 4>(a, nz, s);
      return launch_f16x2_dma_t<1, 8>(a, nz, s);
    }
  }
  if constexpr (P != kF64) {
    if (cfg == k64x32_w8) return launch_stage_t<P, 64, 32, BK, 2, 1, 4, BMODE, PF>(a, nz, s);
    if (cfg == k64x64_k4) return launch_stage_t<P, 64, 64, BK, 1, 1, 4, BMODE, PF>(a, nz, s);
    if (cfg == k64x32_k4) return launch_stage_t<P, 64, 32, BK, 1, 1, 4, BMODE, PF>(a, nz, s);
    if (cfg == k128x64) return launch_stage_t<P, 128, 64, BK / 2, 2, 2, 1, BMODE, PF>(a, nz, s);
    if (cfg == k64x128) return launch_stage_t<P, 64, 128, BK / 2, 2, 2, 1, BMODE, PF>(a, nz, s);
  }
  set_last_error("tile configuration %d is not available for this precision", cfg);
  return kInvalidValue;
}

// BF16x3 shapes: 64-deep k-tiles for the 4-wave shapes (three operand arrays per matrix), the
// 8-wave 64x32 shape keeps 128-deep k-tiles (its 512 threads need 16 chunks per tile row).
static int launch_bf16x3(const StageArgs& a, int nz, hipStream_t s) {
  const int cfg = a.tile_cfg == kTileAuto ? k64x32_w8 : a.tile_cfg;
  switch (cfg) {
    case k64x64_w2x2: return launch_stage_t<kBF16x3, 64, 64, 64, 2, 2, 1, kRowsPadded, 2>(a, nz, s);
    case k64x32_w2x1: return launch_stage_t<kBF16x3, 64, 32, 64, 2, 1, 2, kRowsPadded, 2>(a, nz, s);
    case k128x64: return launch_stage_t<kBF16x3, 128, 64, 32, 2, 2, 1, kRowsPadded, 2>(a, nz, s);
    case k64x32_w8: return launch_stage_t<kBF16x3, 64, 32, 128, 2, 1, 4, kRowsPadded, 2>(a, nz, s);
    default: break;
  }
  set_last_error("tile configuration %d is not available for BF16x3", cfg);
  return kInvalidValue;
}

template <int P, int BM, int BN, int BK, int WM, int WN, int WK, int BMODE, int PF>
static int launch_grouped_t(const StageArgs* table, const int64_t* tile_start, const int* tile_item, int n, int64_t tiles, hipStream_t s,
                            int nyq_k = 0) {
  using T = typename PrecTraits<P>::T;
  constexpr int VEC = 16 / (int)sizeof(T);
  constexpr int NARR = PrecTraits<P>::NARR;
  constexpr size_t lds = 2 * (size_t)NARR * (BM + BN) * (BK + VEC) * sizeof(T);
  constexpr size_t lds_max = stage_lds_max<P, BM, BN, BK, BMODE>();
  static_assert(lds <= 156 * 1024, "LDS budget of one CU (with the kernel's static arrays)");
  auto kern = mft_stage_grouped_kernel<P, BM, BN, BK, WM, WN, WK, BMODE, PF>;
  static bool configured = false;
  if (!configured) {
    HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_max));
    configured = true;
  }
  HCIPY_REQUIRE(tiles > 0 && tiles <= INT32_MAX, "grouped launch of %lld workgroups", (long long)tiles);
  const size_t extra = nyq_lds_bytes(nyq_k, BK);
  HCIPY_REQUIRE(lds + extra <= lds_max, "extra constant row of %d values exceeds the LDS reserve", nyq_k);
  hipLaunchKernelGGL(kern, dim3((unsigned)tiles), dim3(64 * WM * WN * WK), lds + extra, s, table, tile_start, tile_item, n);
  return check_hip(hipGetLastError(), "mft_stage_grouped_kernel launch");
}

template <int P, int BK, int BMODE, int PF>
static int launch_grouped_prec(int cfg, const StageArgs* table, const int64_t* tile_start, const int* tile_item, int n, int64_t tiles,
                               hipStream_t s) {
  switch (cfg) {
    case k64x64_w2x2: return launch_grouped_t<P, 64, 64, BK, 2, 2, 1, BMODE, PF>(table, tile_start, tile_item, n, tiles, s);
    case k64x32_w2x1: return launch_grouped_t<P, 64, 32, BK, 2, 1, 2, BMODE, PF>(table, tile_start, tile_item, n, tiles, s);
    case k64x64_w8: return launch_grouped_t<P, 64, 64, BK, 2, 2, 2, BMODE, PF>(table, tile_start, tile_item, n, tiles, s);
    default: break;
  }
  if constexpr (P != kF64) {
    if (cfg == k64x32_w8) return launch_grouped_t<P, 64, 32, BK, 2, 1, 4, BMODE, PF>(table, tile_start, tile_item, n, tiles, s);
    if (cfg == k128x64) return launch_grouped_t<P, 128, 64, BK / 2, 2, 2, 1, BMODE, PF>(table, tile_start, tile_item, n, tiles, s);
  }
  if constexpr (P == kF16x2) {
    if (cfg == k128x64_h) return launch_grouped_t<P, 128, 64, 32, 2, 2, 1, BMODE, PF>(table, tile_start, tile_item, n, tiles, s);
  }
  set_last_error("tile configuration %d is not available to grouped launches", cfg);
  return kInvalidValue;
}

// Tile shape of a grouped launch.  Short contractions (every item K <= 512: for example the
// adjoint of a transform onto a small focal grid) are bound by the per-workgroup prologue and
// epilogue, so they take shapes with fewer k-tiles in flight and more output per workgroup
// (64x32 with 4 waves for stage 1, 128x64 half-depth for stage 2); long ones take the 8-wave
// shapes, 64x64 when that still gives two workgroups per CU.
int grouped_tile_cfg(int prec, int stage, int max_k, int64_t tiles64) {
  if (prec == kF64) return stage == 0 ? k64x64_w2x2 : k64x32_w2x1;
  if (prec == kBF16x3) {
    if (max_k <= 512) return stage == 0 ? k64x32_w2x1 : k128x64;
    return k64x32_w8;
  }
  if (max_k <= 512) return stage == 0 ? k64x32_w2x1 : k128x64;
  if (stage == 0 && tiles64 >= 512) return k64x64_w8;
  return k64x32_w8;
}

// Tile shapes launch_stage_grouped instantiates for each precision.
bool grouped_cfg_supported(int prec, int cfg) {
  if (prec == kF64) return cfg == k64x64_w2x2 || cfg == k64x32_w2x1 || cfg == k64x64_w8;
  if (prec == kBF16x3) return cfg == k64x32_w2x1 || cfg == k128x64 || cfg == k64x32_w8;
  if (prec == kF16x2 && cfg == k128x64_h) return true;
  return cfg == k64x64_w2x2 || cfg == k64x32_w2x1 || cfg == k64x64_w8 || cfg == k64x32_w8 || cfg == k128x64;
}

bool is_comb_cfg(int cfg);

// Grouped launches of the parity form: stage 1 (real constant) and the recombining stage 2.
// Stage-1 shapes here leave room in LDS for an extra row of up to 4096 constants.
int grouped_real_cfg(int prec, int stage, int64_t tiles64) {
  // Stage 2: 32-deep k-tiles (61 KB of LDS, two workgroups per CU) over the 64-deep ones (110 KB,
  // one per CU).  Grouped launches run thousands of short-contraction tiles, where the
  // per-workgroup prologue and the recombining epilogue dominate and a second resident workgroup
  // hides them: 20 % more propagations/s on 32 x 32 transforms 256^2 -> 128^2 (contraction 128),
  // 5 % on mixed 256^2 / 512^2 / 1024^2 pupils.  Stage 1 keeps 64-deep k-tiles: its 32-deep
  // shape (kR64x64_h) measured 1 % slower on the former and 3 % on the latter.
  if (stage == 1) return kRC32x32_h;
  if (tiles64 < 128) return prec == kF16x2 ? k64x32_w8 : k64x32_w2x1;
  return prec == kF16x2 ? kR64x64_w4 : k64x64_w2x2;
}

// Output rows x columns per workgroup of the parity shapes grouped launches use.
void real_tile_shape(int cfg, int* bm, int* bn) {
  if (is_comb_cfg(cfg)) {
    *bm = 32;
    *bn = 32;
  } else {
    *bm = 64;
    *bn = (cfg == k64x32_w8 || cfg == k64x32_w2x1) ? 32 : 64;
  }
}

int launch_stage_grouped_real(int prec, int cfg, const StageArgs* table, const int64_t* tile_start,
                              const int* tile_item, int n, int64_t tiles, hipStream_t s, int nyq_k) {
  if (n <= 0 || tiles <= 0) return kOk;
  if (prec == kF16x2) {
    switch (cfg) {
      case kR64x64_w4: return launch_grouped_t<kF16x2, 64, 64, 64, 2, 2, 1, kRealPadded, 2>(table, tile_start, tile_item, n, tiles, s, nyq_k);
      case k64x32_w8: return launch_grouped_t<kF16x2, 64, 32, 128, 2, 1, 4, kRealPadded, 2>(table, tile_start, tile_item, n, tiles, s, nyq_k);
      case kRC32x32_w4: return launch_grouped_t<kF16x2, 64, 128, 64, 2, 2, 1, kRealComb, 2>(table, tile_start, tile_item, n, tiles, s);
      case kRC32x32_h: return launch_grouped_t<kF16x2, 64, 128, 32, 2, 2, 1, kRealComb, 2>(table, tile_start, tile_item, n, tiles, s);
      case kRC32x32_w8: return launch_grouped_t<kF16x2, 64, 128, 64, 2, 2, 2, kRealComb, 2>(table, tile_start, tile_item, n, tiles, s);
      case kR64x64_h: return launch_grouped_t<kF16x2, 64, 64, 32, 2, 2, 1, kRealPadded, 2>(table, tile_start, tile_item, n, tiles, s, nyq_k);
      default: break;
    }
  } else if (prec == kBF16x3) {
    switch (cfg) {
      case k64x64_w2x2: return launch_grouped_t<kBF16x3, 64, 64, 64, 2, 2, 1, kRealPadded, 2>(table, tile_start, tile_item, n, tiles, s, nyq_k);
      case k64x32_w2x1: return launch_grouped_t<kBF16x3, 64, 32, 64, 2, 1, 2, kRealPadded, 2>(table, tile_start, tile_item, n, tiles, s, nyq_k);
      case k64x32_w8: return launch_grouped_t<kBF16x3, 64, 32, 128, 2, 1, 4, kRealPadded, 2>(table, tile_start, tile_item, n, tiles, s, nyq_k);
      case kRC32x32_h: return launch_grouped_t<kBF16x3, 64, 128, 32, 2, 2, 1, kRealComb, 2>(table, tile_start, tile_item, n, tiles, s);
      default: break;
    }
  }
  set_last_error("tile configuration %d is not available to grouped parity launches", cfg);
  return kInvalidValue;
}

void grouped_tile_shape(int cfg, int* bm, int* bn) {
  *bm = (cfg == k128x64 || cfg == k128x64_h) ? 128 : 64;
  *bn = (cfg == k64x32_w2x1 || cfg == k64x32_w8) ? 32 : 64;
}

int launch_stage_grouped(int prec, int cfg, int b_mode, const StageArgs* table, const int64_t* tile_start,
                         const int* tile_item, int n, int64_t tiles, hipStream_t s) {
  if (n <= 0 || tiles <= 0) return kOk;
  if (prec == kF64) {
    if (b_mode == kRowsPadded) return launch_grouped_prec<kF64, 32, kRowsPadded, 4>(cfg, table, tile_start, tile_item, n, tiles, s);
    if (b_mode == kRowsAligned) return launch_grouped_prec<kF64, 32, kRowsAligned, 4>(cfg, table, tile_start, tile_item, n, tiles, s);
    return launch_grouped_prec<kF64, 32, kRowsUnaligned, 4>(cfg, table, tile_start, tile_item, n, tiles, s);
  }
  if (prec == kF32) {
    if (b_mode == kRowsPadded) return launch_grouped_prec<kF32, 64, kRowsPadded, 2>(cfg, table, tile_start, tile_item, n, tiles, s);
    if (b_mode == kRowsAligned) return launch_grouped_prec<kF32, 64, kRowsAligned, 2>(cfg, table, tile_start, tile_item, n, tiles, s);
    return launch_grouped_prec<kF32, 64, kRowsUnaligned, 2>(cfg, table, tile_start, tile_item, n, tiles, s);
  }
  if (prec == kF16x2) {
    HCIPY_REQUIRE(b_mode == kRowsPadded, "split-FP16 stage requires a padded data operand");
    return launch_grouped_prec<kF16x2, 128, kRowsPadded, 2>(cfg, table, tile_start, tile_item, n, tiles, s);
  }
  if (prec == kBF16x3) {
    HCIPY_REQUIRE(b_mode == kRowsPadded, "split-BF16 stage requires a padded data operand");
    if (cfg == k64x32_w2x1) return launch_grouped_t<kBF16x3, 64, 32, 64, 2, 1, 2, kRowsPadded, 2>(table, tile_start, tile_item, n, tiles, s);
    if (cfg == k128x64) return launch_grouped_t<kBF16x3, 128, 64, 32, 2, 2, 1, kRowsPadded, 2>(table, tile_start, tile_item, n, tiles, s);
    if (cfg == k64x32_w8) return launch_grouped_t<kBF16x3, 64, 32, 128, 2, 1, 4, kRowsPadded, 2>(table, tile_start, tile_item, n, tiles, s);
    set_last_error("tile configuration %d is not available to grouped BF16x3 launches", cfg);
    return kInvalidValue;
  }
  set_last_error("unknown precision %d", prec);
  return kInvalidValue;
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
  } else if (prec == kBF16x3) {
    if (b_mode != kRowsPadded) {
      set_last_error("split-BF16 stage requires a padded data operand");
      return kInvalidValue;
    }
    st = launch_bf16x3(a, nz, s);
  } else {
    set_last_error("unknown precision %d", prec);
    return kInvalidValue;
  }
  if (st) return st;
  return check_hip(hipGetLastError(), "mft_stage_kernel launch");
}

int launch_split_rows(int prec, const float* in, int64_t ld_in, int K, int64_t rows, SplitOut o, int64_t ld_out,
                      float* row_scale, hipStream_t s) {
  if (rows <= 0) return kOk;
  const bool vec = ((reinterpret_cast<uintptr_t>(in) & 15) == 0) && (ld_in % 4 == 0) && (K % 4 == 0);
  const dim3 g((unsigned)ceil_div(rows, 4));
#define HCIPY_SPLIT_LAUNCH(FMT)                                                                                     \
  if (vec && K <= 64 * 16 && ld_out <= 64 * 16 + 256)                                                              \
    hipLaunchKernelGGL((split_rows_wave_kernel<16, FMT>), g, dim3(256), 0, s, in, ld_in, K, o, ld_out, row_scale, rows); \
  else if (vec && K <= 64 * 32)                                                                                    \
    hipLaunchKernelGGL((split_rows_wave_kernel<32, FMT>), g, dim3(256), 0, s, in, ld_in, K, o, ld_out, row_scale, rows); \
  else if (vec && K <= 64 * 64)                                                                                    \
    hipLaunchKernelGGL((split_rows_wave_kernel<64, FMT>), g, dim3(256), 0, s, in, ld_in, K, o, ld_out, row_scale, rows); \
  else                                                                                                             \
    hipLaunchKernelGGL((split_rows_kernel<FMT>), dim3((unsigned)rows), dim3(256), 0, s, in, ld_in, K, o, ld_out,     \
                       row_scale);
  if (prec == kF16x2) {
    HCIPY_SPLIT_LAUNCH(kSplitF16x2)
  } else {
    HCIPY_SPLIT_LAUNCH(kSplitBF16x3)
  }
#undef HCIPY_SPLIT_LAUNCH
  return check_hip(hipGetLastError(), "split_rows_kernel launch");
}

int launch_gen_matrix(const double* a, int nc, const double* b, int nk, double sign, int prec, SplitOut out,
                      int64_t lda, int rows_pad, hipStream_t s) {
  const int64_t total = (int64_t)rows_pad * (lda / 2);
  const int64_t blocks = std::min<int64_t>(ceil_div(total, 256), 65536);
  hipLaunchKernelGGL(gen_matrix_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, nc, b, nk, sign, prec, out, lda,
                     rows_pad);
  return check_hip(hipGetLastError(), "gen_matrix_kernel launch");
}

// Parity-form launchers.
bool real_cfg_supported(int prec, int cfg, int stage) {
  if (prec == kF16x2)
    return cfg == k64x64_w2x2 || cfg == k64x32_w2x1 || cfg == k64x64_w8 || cfg == k64x32_w8 || cfg == k128x64 ||
           cfg == k64x128 || cfg == kR64x64_w4 || cfg == kR64x32_w4 || cfg == kR32x64_w4 || cfg == kR64x64_h ||
           (stage == 1 && (cfg == kRC32x32_w4 || cfg == kRC32x32_h || cfg == kRC32x32_w8));
  if (prec == kBF16x3)
    return cfg == k64x64_w2x2 || cfg == k64x32_w2x1 || cfg == k128x64 || cfg == k64x32_w8 || cfg == k64x64_w8 ||
           cfg == kR64x64_h || (stage == 1 && (cfg == kRC32x32_h || cfg == kRC32x32_w8));
  return false;
}

bool is_comb_cfg(int cfg) { return cfg == kRC32x32_w4 || cfg == kRC32x32_h || cfg == kRC32x32_w8; }

int launch_stage_real(int prec, const StageArgs& a, int nz, hipStream_t s) {
  if (a.M <= 0 || a.N <= 0 || a.K <= 0 || nz <= 0) return kOk;
  const int64_t tiles64 = ceil_div(a.M, 64) * ceil_div(a.N, 64) * nz;
  int st = kInvalidValue;
  if (prec == kF16x2) {
    // 64-deep k-tiles, two workgroups per CU: measured fastest on both parity stages at
    // 1024^2 -> 512^2 (K = 512); the 8-wave shapes for small grids
    const int cfg = a.tile_cfg == kTileAuto ? (tiles64 >= 128 ? kR64x64_w4 : k64x32_w8) : a.tile_cfg;
    switch (cfg) {
      case k64x64_w2x2: st = launch_stage_t<kF16x2, 64, 64, 128, 2, 2, 1, kRealPadded, 2>(a, nz, s); break;
      case k64x32_w2x1: st = launch_stage_t<kF16x2, 64, 32, 128, 2, 1, 2, kRealPadded, 2>(a, nz, s); break;
      case k64x64_w8: st = launch_stage_t<kF16x2, 64, 64, 128, 2, 2, 2, kRealPadded, 2>(a, nz, s); break;
      case k64x32_w8: st = launch_stage_t<kF16x2, 64, 32, 128, 2, 1, 4, kRealPadded, 2>(a, nz, s); break;
      case k128x64: st = launch_stage_t<kF16x2, 128, 64, 64, 2, 2, 1, kRealPadded, 2>(a, nz, s); break;
      case k64x128: st = launch_stage_t<kF16x2, 64, 128, 64, 2, 2, 1, kRealPadded, 2>(a, nz, s); break;
      case kR64x64_w4: st = launch_stage_t<kF16x2, 64, 64, 64, 2, 2, 1, kRealPadded, 2>(a, nz, s); break;
      case kR64x32_w4: st = launch_stage_t<kF16x2, 64, 32, 64, 2, 1, 2, kRealPadded, 2>(a, nz, s); break;
      case kR32x64_w4: st = launch_stage_t<kF16x2, 32, 64, 64, 1, 2, 2, kRealPadded, 2>(a, nz, s); break;
      case kRC32x32_w4: st = launch_stage_t<kF16x2, 64, 128, 64, 2, 2, 1, kRealComb, 2>(a, nz, s); break;
      case kRC32x32_h: st = launch_stage_t<kF16x2, 64, 128, 32, 2, 2, 1, kRealComb, 2>(a, nz, s); break;
      case kRC32x32_w8: st = launch_stage_t<kF16x2, 64, 128, 64, 2, 2, 2, kRealComb, 2>(a, nz, s); break;
      case kR64x64_h: st = launch_stage_t<kF16x2, 64, 64, 32, 2, 2, 1, kRealPadded, 2>(a, nz, s); break;
      default: set_last_error("tile configuration %d is not available to the parity form", cfg); return kInvalidValue;
    }
  } else if (prec == kBF16x3) {
    // 64 x 64 tiles with 64-deep k-tiles (one workgroup per CU) on grids of 128 such tiles or
    // more: twice as fast as the 8-wave 64 x 32 shape on stage 1 at 1024^2 -> 512^2
    const int cfg = a.tile_cfg == kTileAuto ? (tiles64 >= 128 ? k64x64_w2x2 : k64x32_w8) : a.tile_cfg;
    switch (cfg) {
      case k64x64_w2x2: st = launch_stage_t<kBF16x3, 64, 64, 64, 2, 2, 1, kRealPadded, 2>(a, nz, s); break;
      case k64x32_w2x1: st = launch_stage_t<kBF16x3, 64, 32, 64, 2, 1, 2, kRealPadded, 2>(a, nz, s); break;
      case k128x64: st = launch_stage_t<kBF16x3, 128, 64, 32, 2, 2, 1, kRealPadded, 2>(a, nz, s); break;
      case k64x32_w8: st = launch_stage_t<kBF16x3, 64, 32, 128, 2, 1, 4, kRealPadded, 2>(a, nz, s); break;
      // 32-deep k-tiles: 64 x 64 tiles on 4 waves (61 KB of LDS, two workgroups per CU) or on 8
      // waves, two per output tile each on half of every k-tile
      case kR64x64_h: st = launch_stage_t<kBF16x3, 64, 64, 32, 2, 2, 1, kRealPadded, 2>(a, nz, s); break;
      case k64x64_w8: st = launch_stage_t<kBF16x3, 64, 64, 32, 2, 2, 2, kRealPadded, 2>(a, nz, s); break;
      case kRC32x32_h: st = launch_stage_t<kBF16x3, 64, 128, 32, 2, 2, 1, kRealComb, 2>(a, nz, s); break;
      // two waves per item on 32-deep k-tiles (64-deep tiles of three arrays exceed the LDS)
      case kRC32x32_w8: st = launch_stage_t<kBF16x3, 64, 128, 32, 2, 2, 2, kRealComb, 2>(a, nz, s); break;
      default: set_last_error("tile configuration %d is not available to the parity form", cfg); return kInvalidValue;
    }
  } else {
    set_last_error("the parity form runs the split precisions only");
    return kInvalidValue;
  }
  if (st) return st;
  return check_hip(hipGetLastError(), "mft_stage_kernel (parity) launch");
}

int launch_parity_split(int prec, const float* in, int nx, int ny, const ParAxis& ax, const ParAxis& ay,
                        int64_t batch, SplitOut o, int64_t ld_out, float* row_scale, const double* nyq_coef,
                        double* nyq_out, hipStream_t s) {
  HCIPY_REQUIRE(ld_out <= 128 * 16, "parity split of %d-sample half rows (at most 2048)", ax.h);
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const int64_t nb = std::min<int64_t>(65535, batch - b0);
    const float2* ib = reinterpret_cast<const float2*>(in) + b0 * (int64_t)nx * ny;
    const int64_t rows = b0 * 4 * 2 * (int64_t)ay.h;  // split rows of the fields before this chunk
    SplitOut ob = o;
    for (int i = 0; i < 3; ++i)
      if (ob.p[i]) ob.p[i] = static_cast<char*>(ob.p[i]) + (size_t)rows * ld_out * 2;
    // more threads per y' for more loads in flight: XPT * TPR covers the padded half row
#define HCIPY_PSPLIT(FMT, XPT, TPR)                                                                          \
  hipLaunchKernelGGL((parity_split_kernel<FMT, XPT, TPR>), dim3((unsigned)ceil_div(ay.h, 256 / TPR), (unsigned)nb), \
                     dim3(256), 0, s, ib, nx, ny, ax, ay, ob, ld_out, row_scale + rows, nyq_coef,            \
                     nyq_out != nullptr ? nyq_out + b0 * 4 * 2 * (int64_t)ay.h : nullptr)
#define HCIPY_PSPLIT_ALL(FMT)                             \
  if (ld_out <= 256) HCIPY_PSPLIT(FMT, 2, 128);           \
  else if (ld_out <= 512) HCIPY_PSPLIT(FMT, 2, 256);      \
  else if (ld_out <= 1024) HCIPY_PSPLIT(FMT, 4, 256);     \
  else HCIPY_PSPLIT(FMT, 8, 256);
    if (prec == kF16x2) {
      HCIPY_PSPLIT_ALL(kSplitF16x2)
    } else {
      HCIPY_PSPLIT_ALL(kSplitBF16x3)
    }
#undef HCIPY_PSPLIT_ALL
#undef HCIPY_PSPLIT
  }
  return check_hip(hipGetLastError(), "parity_split_kernel launch");
}

int launch_parity_split_items(int prec, const SplitItems& items, int n, int nx, int ny, const ParAxis& ax,
                              const ParAxis& ay, int64_t ld_out, hipStream_t s) {
  HCIPY_REQUIRE(n >= 1 && n <= kSplitItemsMax, "%d items in one parity split launch", n);
  HCIPY_REQUIRE(ld_out <= 128 * 16, "parity split of %d-sample half rows (at most 2048)", ax.h);
#define HCIPY_PSPLIT(FMT, XPT, TPR)                                                                            \
  hipLaunchKernelGGL((parity_split_items_kernel<FMT, XPT, TPR>), dim3((unsigned)ceil_div(ay.h, 256 / TPR), (unsigned)n), \
                     dim3(256), 0, s, items, nx, ny, ax, ay, ld_out)
#define HCIPY_PSPLIT_ALL(FMT)                             \
  if (ld_out <= 256) HCIPY_PSPLIT(FMT, 2, 128);           \
  else if (ld_out <= 512) HCIPY_PSPLIT(FMT, 2, 256);      \
  else if (ld_out <= 1024) HCIPY_PSPLIT(FMT, 4, 256);     \
  else HCIPY_PSPLIT(FMT, 8, 256);
  if (prec == kF16x2) {
    HCIPY_PSPLIT_ALL(kSplitF16x2)
  } else {
    HCIPY_PSPLIT_ALL(kSplitBF16x3)
  }
#undef HCIPY_PSPLIT_ALL
#undef HCIPY_PSPLIT
  return check_hip(hipGetLastError(), "parity_split_items_kernel launch");
}

int launch_gen_parity(const double* c, int nc, const double* k, int nk, int prec, SplitOut out, int64_t lda,
                      int rows_pad, hipStream_t s) {
  const int64_t total = 2 * (int64_t)rows_pad * lda;
  const int blocks = (int)std::min<int64_t>(ceil_div(total, 256), 65535);
  hipLaunchKernelGGL(gen_parity_kernel, dim3(blocks), dim3(256), 0, s, c, nc, k, nk, prec, out, lda, rows_pad);
  return check_hip(hipGetLastError(), "gen_parity_kernel launch");
}

int launch_parity_combine(const double* H, const ParAxis& au, const ParAxis& av, int nu, int nv, double sign, double ar,
                          double ai, int out_mode, void* out, int64_t batch, hipStream_t s) {
  const int64_t n = batch * (int64_t)au.h * av.h;
  if (n <= 0) return kOk;
  HCIPY_REQUIRE(ceil_div(n, 256) <= INT32_MAX, "parity recombination of %lld points", (long long)n);
  hipLaunchKernelGGL(parity_combine_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(H), au, av, nu, nv, sign, ar, ai, out_mode, out, batch);
  return check_hip(hipGetLastError(), "parity_combine_kernel launch");
}

// Split tasks per field of a fused execution: the split's thread layout for half rows of ld_out
// reals (two y' per task where a team of 128 threads covers the row).
int fused_split_teams(int64_t ld_out) { return ld_out <= 512 ? 2 : 1; }

bool fused_supported(int prec, int64_t ld_out) { return prec == kF16x2 && ld_out <= 128 * 16; }

int launch_parity_fused(const FusedArgs& f, int grid, hipStream_t s) {
  using T = _Float16;
  constexpr int VEC = 16 / (int)sizeof(T);
  constexpr size_t lds1 = 2 * (size_t)2 * (64 + 64) * (64 + VEC) * sizeof(T);
  constexpr size_t lds2 = 2 * (size_t)2 * (64 + 128) * (64 + VEC) * sizeof(T);
  constexpr size_t lds = lds1 > lds2 ? lds1 : lds2;
  static_assert(lds <= 150 * 1024, "LDS budget of one CU (with the static arrays of the three task kinds)");
  HCIPY_REQUIRE(grid >= 8 && grid % 8 == 0, "fused parity launch of %d workgroups (a positive multiple of 8)", grid);
  auto launch = [&](void (*kern)(FusedArgs)) -> int {
    HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), lds, s, f);
    return check_hip(hipGetLastError(), "parity_fused_kernel launch");
  };
  HCIPY_REQUIRE(fused_supported(kF16x2, f.ld_out), "fused parity form of %lld-real half rows", (long long)f.ld_out);
  if (f.ld_out <= 256) return launch(parity_fused_kernel<2, 128>);
  if (f.ld_out <= 512) return launch(parity_fused_kernel<4, 128>);
  if (f.ld_out <= 1024) return launch(parity_fused_kernel<4, 256>);
  return launch(parity_fused_kernel<8, 256>);
}

}  // namespace hcipy
