// MFMA kernels for one stage of the separable matrix Fourier transform.
//
// A 2-D MFT is  O = alpha * M1 . E . M2  (reference: hcipy/fourier/matrix_fourier_transform.py:171-172,
// backward :219-220).  Each of the two complex contractions is run here as ONE real NT-GEMM
//
//     C[m][n] = sum_k  A[m][k] * B[n][k]            (A: M x K, B: N x K, both K-contiguous)
//
// where
//   * B is the data operand exactly as it sits in HBM: rows of interleaved complex numbers
//     (re, im, re, im, ...), i.e. K = 2 * (complex contraction length);
//   * A is the constant Fourier matrix, expanded once per plan into 2x2 real blocks so that row
//     2c gives Re and row 2c+1 gives Im of the complex result:
//         A[2c  ][2k] =  Re m(c,k)   A[2c  ][2k+1] = -Im m(c,k)
//         A[2c+1][2k] =  Im m(c,k)   A[2c+1][2k+1] =  Re m(c,k)
//   * the MFMA accumulator layout puts rows (2c, 2c+1) of one column in the same lane, so the
//     epilogue holds complete complex values and writes them interleaved, coalesced along n.
//
// Four arithmetic variants share the tiling:
//   kF64    : v_mfma_f64_16x16x4_f64   (complex128 fields; parity with the FP64 reference)
//   kF32    : v_mfma_f32_32x32x2_f32   (plain FP32; single-level FP32 accumulation)
//   kF16x2  : v_mfma_f32_32x32x16_f16  on hi/lo FP16 splits of both operands:
//               x = hi + lo * 2^-11,  A.B ~= Ah.Bh + 2^-11 (Ah.Bl + Al.Bh)
//             The Ah.Bh accumulator is flushed into FP64 every kFlushK real k; the cross
//             accumulators are added (times 2^-11) at the end.  Data rows carry a power-of-two
//             scale so that every split value lies in [-1, 1].
//   kBF16x3 : v_mfma_f32_32x32x16_bf16 on three-way BF16 splits of both operands:
//               x = b0 + b1 + b2 (each the bfloat16 rounding of the remainder, |x - sum| <= 2^-27 |x|)
//               A.B ~= A0.B0 + (A0.B1 + A1.B0) + (A0.B2 + A1.B1 + A2.B0)   (six MFMAs per k-vector)
//             A0.B0 is flushed into FP64 like the FP16 hi.hi products; the two cross sums
//             (2^-9 and 2^-18 of it) run in FP32 over the whole contraction.
#pragma once

#include "common.hpp"

namespace hcipy {

enum Prec : int { kF64 = 0, kF32 = 1, kF16x2 = 2, kBF16x3 = 3 };

// True for the split precisions (data rows are split before stage 1, the intermediate is
// written split, the operand arrays are 2 or 3 per matrix).
inline bool is_split(int prec) { return prec == kF16x2 || prec == kBF16x3; }
inline int split_arrays(int prec) { return prec == kF16x2 ? 2 : (prec == kBF16x3 ? 3 : 1); }

// Output formats written by the epilogue.
enum OutMode : int {
  kOutC128 = 0,   // interleaved complex128
  kOutC64 = 1,    // interleaved complex64
  kOutSplit = 2,  // interleaved split arrays (FP16 hi/lo or BF16 b0/b1/b2) feeding the next split stage
  kOutI32 = 3,    // |z|^2 as float32, evaluated from the FP64 result
  kOutI64 = 4,    // |z|^2 as float64
  kOutI32Base = 5,  // |z + base|^2 as float32 (base: complex128 array of the output's shape)
  kOutI64Base = 6,  // |z + base|^2 as float64
  // Real-constant (parity) stages, one real value per accumulator:
  kOutSplitT = 7,   // split arrays, element (m, n) at row 2m + (n & 1), column n >> 1 (re / im rows)
  kOutR64 = 8,      // float64, element (m, n) at m * ldo + n
};

template <int P> struct PrecTraits;
template <> struct PrecTraits<kF64> {
  using T = double;
  static constexpr int MT = 16;      // MFMA tile edge
  static constexpr int GROUPS = 4;   // lane groups along k inside one MFMA
  static constexpr int KPL = 4;      // consecutive k values one lane reads per k-vector
  static constexpr int KVEC = GROUPS * KPL;
  static constexpr int NACC = 4;     // accumulator registers per lane per MFMA tile
  static constexpr bool SPLIT = false;
  static constexpr int NARR = 1;     // operand arrays per matrix
};
template <> struct PrecTraits<kF32> {
  using T = float;
  static constexpr int MT = 32;
  static constexpr int GROUPS = 2;
  static constexpr int KPL = 4;
  static constexpr int KVEC = GROUPS * KPL;
  static constexpr int NACC = 16;
  static constexpr bool SPLIT = false;
  static constexpr int NARR = 1;
};
template <> struct PrecTraits<kF16x2> {
  using T = _Float16;
  static constexpr int MT = 32;
  static constexpr int GROUPS = 2;
  static constexpr int KPL = 8;
  static constexpr int KVEC = GROUPS * KPL;
  static constexpr int NACC = 16;
  static constexpr bool SPLIT = true;
  static constexpr int NARR = 2;
};
template <> struct PrecTraits<kBF16x3> {
  using T = __bf16;
  static constexpr int MT = 32;
  static constexpr int GROUPS = 2;
  static constexpr int KPL = 8;
  static constexpr int KVEC = GROUPS * KPL;
  static constexpr int NACC = 16;
  static constexpr bool SPLIT = true;
  static constexpr int NARR = 3;
};

// Scale of the lo part of a split value: x = hi + lo * kLoScale.
constexpr double kLoScale = 1.0 / 2048.0;
constexpr float kLoInv = 2048.0f;
// Real k values accumulated in FP32 by the hi.hi MFMAs between flushes into FP64.
constexpr int kFlushK = 128;

// One axis of a grid in the parity form: h magnitudes P[j] >= 0; the sample at +P[j] is grid
// index plus[j], the one at -P[j] is minus[j] (-1: no such sample, e.g. the origin's mirror or
// the unpaired first sample of an FFT-style grid).  Device tables.
struct ParAxis {
  const int* plus;
  const int* minus;
  const double* P;
  int h;
  // Regular grids: plus[j] = p0 + j where that is < n, minus[j] = m0 - j for j >= mlo where that
  // is >= 0 (affine != 0; the tables then need not be read).
  int affine, p0, m0, mlo, n;
};

__device__ inline int par_plus(const ParAxis& a, int j) {
  if (a.affine) return a.p0 + j < a.n ? a.p0 + j : -1;
  return a.plus[j];
}
__device__ inline int par_minus(const ParAxis& a, int j) {
  if (a.affine) return (j >= a.mlo && a.m0 - j >= 0) ? a.m0 - j : -1;
  return a.minus[j];
}

struct StageArgs {
  // Constant operand (expanded Fourier matrix), padded so that tile loads never leave it.
  const void* a;
  const void* a_lo;
  const void* a_lo2;            // third split array (kBF16x3)
  int64_t lda;
  // Data operand.
  const void* b;
  const void* b_lo;
  const void* b_lo2;
  int64_t ldb;
  int64_t b_zstride;            // elements between batch items along gridDim.z
  // Problem size: M = 2 * (number of constant rows), N = data rows, K = real contraction length.
  int M, N, K;
  // Output.
  void* out;
  void* out_lo;
  void* out_lo2;
  int64_t ldo;                  // complex elements between consecutive constant rows
  int64_t out_zstride;          // complex elements between batch items along gridDim.z
  int out_mode;
  const void* base;             // complex128 field added before |.|^2 (kOutI32Base / kOutI64Base)
  int nc_layout;                // 1: out[n * ldo + c] instead of out[c * ldo + n]
  // Folding of several fields into the N dimension: n = item * fold + y is written to
  // out[item * fold_stride + c * ldo + y].  fold = 0 disables folding.
  int fold;
  int64_t fold_stride;
  // Scales of the split (kF16x2) path.
  const float* row_scale;       // data row n was multiplied by row_scale[n]   (or nullptr)
  const float* z_scale;         // data of batch item z was multiplied by z_scale[z] (or nullptr)
  float* t_out;                 // split output: per folded item scale, written here (or nullptr)
  int t_log2k;                  // ceil(log2(K)): bound |C| <= K / min(row_scale)
  double alpha_re, alpha_im;
  int tile_cfg;                 // kTileAuto or one of TileCfg
  // Per batch slice z: the constant operand is plane (z >> a_zshift) & 1 of a two-plane array
  // (a_zstride elements apart; 0: one plane), row scales and split-output scales of slice z start
  // rs_zstride / t_zstride entries further on.
  int64_t a_zstride;
  int a_zshift;
  int64_t rs_zstride;
  int t_zstride;
  // Real split output (kOutSplitT) with one more output row than the product has (row nyq_row),
  // stored by the workgroups of the first row tile with the product rows' scales and split:
  //  - nyq set: from the FP64 values nyq[z * nyq_zstride + n] (unscaled results, formed by the
  //    kernel that split the data);
  //  - nyq_coef set: formed here, sum_k c[k] B[n][k] with the constant row c = nyq_coef +
  //    ((z >> a_zshift) & 1) * K (K floats per plane), from the data tiles the workgroup stages.
  const double* nyq;
  int64_t nyq_zstride;
  const float* nyq_coef;
  int nyq_row;
  // Parity stage 2 with the recombination in its epilogue (tile shapes kRC*): batch slice z is one
  // field whose four parity items are b_zstride apart; M = rows of the output axis cv, N = 2 x
  // magnitudes of cu; out_mode is the final output kind, the field's outputs are c_nv x c_nu at
  // out + z * c_nv * c_nu; z_scale holds four scales per field; c_sign = -1 forward, +1 backward.
  ParAxis cu, cv;
  int c_nu, c_nv;
  double c_sign;
};

// Workgroup tile shapes of the stage kernel (256 threads = 4 waves in every shape):
//   k64x64_w2x2 : 64 x 64 tile, 2 x 2 waves of 32 x 32
//   k64x32_w2x1 : 64 x 32 tile, 2 x 1 waves of 32 x 32, k-vectors split over 2 waves
//   k64x64_k4   : 64 x 64 tile, every wave the whole tile on a quarter of the k-vectors
//   k64x32_k4   : 64 x 32 tile, every wave the whole tile on a quarter of the k-vectors
//   k128x64     : 128 x 64 tile, 2 x 2 waves of 64 x 32, half-depth k-tiles (fewer bytes per output)
//   k64x128     : 64 x 128 tile, 2 x 2 waves of 32 x 64, half-depth k-tiles
//   k64x64_w8   : 64 x 64 tile, 8 waves (2 per SIMD): 2 x 2 waves of 32 x 32, k-vectors split over 2
//   k64x32_w8   : 64 x 32 tile, 8 waves: 2 x 1 waves of 32 x 32, k-vectors split over 4
// Split-FP16 only, constant kept compact (32 complex rows = 64 expanded rows per tile):
//   kC64x64_w4  : 32 x 64 compact tile, 4 waves: 1 x 2 waves, k-vectors split over 2
//   kC64x64_w8  : 32 x 64 compact tile, 8 waves: 1 x 2 waves, k-vectors split over 4
//   kC64x32_w4  : 32 x 32 compact tile, 4 waves, k-vectors split over 4
//   kC128x32_w8 : 64 x 32 compact tile, 8 waves: 2 x 1 waves, k-vectors split over 4
// Split-FP16 only, compact constant, operands loaded straight into LDS (three LDS stages):
//   kD64x64_w4 / kD64x64_w8 : 32 x 64 compact tile, 4 / 8 waves (k-vectors split over 2 / 4)
//   kD64x32_w4 / kD64x32_w8 : 32 x 32 compact tile, 4 / 8 waves (k-vectors split over 4 / 8)
enum TileCfg : int {
  kTileAuto = -1,
  k64x64_w2x2 = 0,
  k64x32_w2x1 = 1,
  k64x64_k4 = 2,
  k64x32_k4 = 3,
  k128x64 = 4,
  k64x128 = 5,
  k64x64_w8 = 6,
  k64x32_w8 = 7,
  kC64x64_w4 = 8,
  kC64x64_w8 = 9,
  kC64x32_w4 = 10,
  kC128x32_w8 = 11,
  kD64x64_w4 = 12,
  kD64x64_w8 = 13,
  kD64x32_w4 = 14,
  kD64x32_w8 = 15,
  // Parity form only (real constant), 64-deep k-tiles: 74 KB of LDS, two workgroups per CU
  kR64x64_w4 = 16,  // 64 x 64 tile, 2 x 2 waves of 32 x 32
  kR64x32_w4 = 17,  // 64 x 32 tile, 2 x 1 waves of 32 x 32, k-vectors split over 2
  kR32x64_w4 = 18,  // 32 x 64 tile, 1 x 2 waves of 32 x 32, k-vectors split over 2
  // Parity form, stage 2 only: 32 output rows x 32 (u', re / im) columns of all four parity items
  // per workgroup (wave w: item w), recombined into the output quadrants in the epilogue
  kRC32x32_w4 = 19,  // 64-deep k-tiles (split-FP16 only)
  kRC32x32_h = 20,   // 32-deep k-tiles
  kRC32x32_w8 = 21,  // 64-deep k-tiles, 8 waves: two per item, each on half of every k-tile (split-FP16 only)
  // Parity form, split-FP16: 64 x 64 tile, 2 x 2 waves of 32 x 32, 32-deep k-tiles (41 KB of LDS),
  // a candidate for short contractions
  kR64x64_h = 22,
  // Complex form, split-FP16, grouped launches only: as k128x64 with 32-deep k-tiles (61 KB of
  // LDS, two workgroups per CU instead of one) for launches of very many short contractions
  k128x64_h = 23,
  kNumTileCfg = 24,
};

// Up to three split arrays (device pointers), passed by value to the split and matrix kernels.
struct SplitOut {
  void* p[3];
};

// Parity splits of several single fields of one geometry in one launch (grouped executions whose
// items read different inputs): input, split arrays and row scales per item, passed by value.
constexpr int kSplitItemsMax = 16;
struct SplitItems {
  const void* in[kSplitItemsMax];
  SplitOut o[kSplitItemsMax];
  float* rs[kSplitItemsMax];
};

// DISABLED (capi.hip kFusedEnabled): launches of this kernel did not finish on MI355X, and plans
// never take it; the code is kept for the investigation of that failure.
// One execution of the parity form as a single persistent launch (parity_fused_kernel): the
// parity split, stage 1 and the recombining stage 2 run as tasks of one kernel.  Tasks are
// numbered level by level (every split task, then every stage-1 tile, then every stage-2 tile,
// each level in the order its own launch would dispatch them), and workgroup b takes the tasks
// with linear index = b mod 8 within each level from queue b mod 8, in order.  A task waits only
// for tasks of lower levels, which are taken before it, so the launch finishes whatever number
// of its workgroups are resident; the waits are bounded anyway (a timeout sets an error flag).
//
// Sync block (plan-owned device memory, one 128-byte line per counter, all zero between
// executions: the last workgroup to finish clears them):
//   line q (0..7)       next task of queue q
//   line 8 + q          split tasks of queue q finished
//   line 16             workgroups finished
//   line 17             timeouts (sticky; read and cleared by the host)
//   line 18 + r         stage-1 tiles of row tile r finished (all items, fields and column tiles)
constexpr int kSyncLine = 32;  // unsigned per counter line
constexpr int kSyncFixedLines = 18;
struct FusedArgs {
  const void* in;   // c64 fields [batch][ny][nx]
  int nx, ny;
  ParAxis ax, ay;   // input axes (stage-1 and stage-2 contractions)
  SplitOut fo;      // split components, ld_out reals per row
  int64_t ld_out;
  float* row_scale;
  const double* nyq_coef;  // extra stage-1 row formed by the split (or nullptr)
  double* nyq_out;
  int nblk;         // split tasks per field
  StageArgs s1;     // stage 1 over 4 * batch slices
  StageArgs s2;     // recombining stage 2 over batch slices
  int64_t n0, n1, n2;  // tasks of each level over all fields
  unsigned* sync;
};

// True if `cfg` is instantiated for precision `prec`.
bool tile_cfg_supported(int prec, int cfg);

}  // namespace hcipy
