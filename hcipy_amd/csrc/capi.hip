// C-ABI of the library: error state and the matrix-Fourier-transform plan.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/hcipy_amd.h"
#include "mft_gemm.hpp"
#include "pokes.hpp"

namespace hcipy {

static thread_local char g_last_error[512] = "";

void set_last_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return kOk;
  set_last_error("%s: %s", what, hipGetErrorString(e));
  return kHipError;
}

int refuse_under_capture(hipStream_t stream, const char* what) {
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  HCIPY_HIP(hipStreamIsCapturing(stream, &capture));
  HCIPY_REQUIRE(capture == hipStreamCaptureStatusNone,
                "%s: the stream is capturing a graph; create the plan before the capture starts", what);
  return kOk;
}

int launch_stage(int prec, const StageArgs& a, int nz, int b_mode, hipStream_t s);
int launch_stage_grouped(int prec, int cfg, int b_mode, const StageArgs* table, const int64_t* tile_start,
                         const int* tile_item, int n, int64_t tiles, hipStream_t s);
int grouped_tile_cfg(int prec, int stage, int max_k, int64_t tiles64);
bool grouped_cfg_supported(int prec, int cfg);
void grouped_tile_shape(int cfg, int* bm, int* bn);
int grouped_real_cfg(int prec, int stage, int64_t tiles64);
void real_tile_shape(int cfg, int* bm, int* bn);
int launch_stage_grouped_real(int prec, int cfg, const StageArgs* table, const int64_t* tile_start,
                              const int* tile_item, int n, int64_t tiles, hipStream_t s, int nyq_k);
constexpr int kModePadded = 0, kModeAligned = 1, kModeUnaligned = 2;  // BMode of mft_gemm.hip
int launch_split_rows(int prec, const float* in, int64_t ld_in, int K, int64_t rows, SplitOut o, int64_t ld_out,
                      float* row_scale, hipStream_t s);
int launch_gen_matrix(const double* a, int nc, const double* b, int nk, double sign, int prec, SplitOut out,
                      int64_t lda, int rows_pad, hipStream_t s);
bool real_cfg_supported(int prec, int cfg, int stage);
bool is_comb_cfg(int cfg);
int launch_stage_real(int prec, const StageArgs& a, int nz, hipStream_t s);
int launch_parity_split(int prec, const float* in, int nx, int ny, const ParAxis& ax, const ParAxis& ay,
                        int64_t batch, SplitOut o, int64_t ld_out, float* row_scale, const double* nyq_coef,
                        double* nyq_out, hipStream_t s);
int launch_gen_parity(const double* c, int nc, const double* k, int nk, int prec, SplitOut out, int64_t lda,
                      int rows_pad, hipStream_t s);
int launch_parity_split_items(int prec, const SplitItems& items, int n, int nx, int ny, const ParAxis& ax,
                              const ParAxis& ay, int64_t ld_out, hipStream_t s);
int launch_parity_combine(const double* H, const ParAxis& au, const ParAxis& av, int nu, int nv, double sign, double ar,
                          double ai, int out_mode, void* out, int64_t batch, hipStream_t s);
int launch_parity_fused(const FusedArgs& f, int grid, hipStream_t s);
int fused_split_teams(int64_t ld_out);
bool fused_supported(int prec, int64_t ld_out);

namespace {

constexpr int64_t kAlign = 256;
// Padded operand rows (constant matrices, split data, intermediate) hold a multiple of this many
// elements, the largest k-tile of the stage kernels, so full-vector loads of the last k-tile
// stay inside the row.
constexpr int64_t kRowPad = 128;
int64_t padded_row(int64_t k) { return round_up(k, kRowPad); }
constexpr int kMaxGridZ = 65535;
// The persistent single-launch parity execution (parity_fused_kernel) did not finish on MI355X:
// two test launches (1024^2 -> 512^2, and 256^2 -> 128^2 with every dependency wait bounded to
// 10 ms) ran until the test's time limit.  Its ISA showed a task loop whose exit the compiler took
// as divergent; the loop is rewritten with a uniform skeleton that has not run on a GPU.  Off until
// it has; plans never take it.
constexpr bool kFusedEnabled = false;

size_t elem_size(int prec) { return prec == kF64 ? 8 : (prec == kF32 ? 4 : 2); }
int n_arrays(int prec) { return split_arrays(prec); }

int ilog2_ceil(int64_t v) {
  int e = 0;
  while ((int64_t(1) << e) < v) ++e;
  return e;
}

// One compact Fourier matrix (nc rows of 2 * nk interleaved reals, padded).
struct FMatrix {
  void* hi = nullptr;
  void* lo = nullptr;
  void* lo2 = nullptr;
  int nc = 0, nk = 0;
  int64_t lda = 0;
  int rows_pad = 0;
  size_t bytes() const { return hi == nullptr ? 0 : (size_t)rows_pad * lda; }
};

// Real parity matrices of one stage: plane 0 cos, plane 1 sin, rows_pad x lda reals each.
struct PMatrix {
  void* p[3] = {nullptr, nullptr, nullptr};
  int64_t lda = 0;
  int rows_pad = 0;
  int64_t plane() const { return (int64_t)rows_pad * lda; }
};

struct Plan {
  int ndim = 0, prec = 0, device = 0;
  int64_t nx = 1, ny = 1, nu = 1, nv = 1;
  FMatrix f1, f2, b1, b2;  // forward stage 1/2, backward stage 1/2
  double* coords = nullptr;
  size_t matrix_bytes = 0;
  int tiles[2][2] = {{kTileAuto, kTileAuto}, {kTileAuto, kTileAuto}};  // [direction][stage]
  // Parity form (grids symmetric about 0, split precisions): available, and in use for single
  // executions.
  bool parity_ok = false, parity = false;
  PMatrix pf1, pf2, pb1, pb2;  // forward stage 1/2, backward stage 1/2
  ParAxis px{}, py{}, pu{}, pv{};  // axis tables (device: plus, minus, P)
  void* ptables = nullptr;
  // Per direction: the last magnitude of the stage-1 output axis has one sample only (the -n/2
  // sample of an FFT-style axis) and dropping its row saves a row tile.  Stage 1 then runs one
  // row fewer; that row (the cosines, then sines, of that magnitude times the contraction
  // magnitudes, against the data) is formed in FP64 by the parity split of a single execution
  // (nyq_coef64) or, in grouped launches, where one split may feed items with different output
  // axes, by stage 1's first row tile (nyq_coef32).
  bool nyq[2] = {false, false};
  double* nyq_coef64[2] = {nullptr, nullptr};
  float* nyq_coef32[2] = {nullptr, nullptr};
  // Single executions of the parity form as one persistent launch (parity_fused_kernel): on by
  // default where available.  `sync` is the launch's counter block (FusedArgs in mft_gemm.hpp),
  // zero between executions; executions of one plan must therefore not overlap in time.
  bool fused = false;
  unsigned* sync = nullptr;
  int cus = 0;  // compute units of the plan's device (persistent grid size)
};

int make_matrix(Plan& pl, FMatrix& m, const double* a, int nc, const double* b, int nk, double sign,
                hipStream_t s) {
  m.nc = nc;
  m.nk = nk;
  // compact complex rows (the stage kernel expands them); padded so tile loads stay inside
  m.lda = padded_row(2 * (int64_t)nk);
  m.rows_pad = (int)round_up((int64_t)nc, 64);
  const size_t bytes = (size_t)m.rows_pad * m.lda * elem_size(pl.prec);
  void** arrs[3] = {&m.hi, &m.lo, &m.lo2};
  for (int i = 0; i < n_arrays(pl.prec); ++i) {
    HCIPY_HIP(hipMalloc(arrs[i], bytes));
    pl.matrix_bytes += bytes;
  }
  return launch_gen_matrix(a, nc, b, nk, sign, pl.prec, SplitOut{{m.hi, m.lo, m.lo2}}, m.lda, m.rows_pad, s);
}

void free_matrix(FMatrix& m) {
  if (m.hi) (void)hipFree(m.hi);
  if (m.lo) (void)hipFree(m.lo);
  if (m.lo2) (void)hipFree(m.lo2);
  m.hi = m.lo = m.lo2 = nullptr;
}

// Parity matrices of a stage with nc rows (positive half of the output coordinates c) and nk
// contraction values (positive half of k); rows padded to whole 128-row tiles.
int make_parity(Plan& pl, PMatrix& m, const double* c, int nc, const double* k, int nk, hipStream_t s) {
  m.lda = padded_row(nk);
  m.rows_pad = (int)round_up((int64_t)nc, 128);
  const size_t bytes = (size_t)2 * m.plane() * elem_size(pl.prec);
  for (int i = 0; i < n_arrays(pl.prec); ++i) {
    HCIPY_HIP(hipMalloc(&m.p[i], bytes));
    pl.matrix_bytes += bytes;
  }
  return launch_gen_parity(c, nc, k, nk, pl.prec, SplitOut{{m.p[0], m.p[1], m.p[2]}}, m.lda, m.rows_pad, s);
}

void free_parity(PMatrix& m) {
  for (auto& q : m.p) {
    if (q) (void)hipFree(q);
    q = nullptr;
  }
}

// Parity description of one axis (host): magnitudes P[j] >= 0 with the indices of the samples
// at +P[j] / -P[j] (-1 if absent), from pairing x[i] with the x[k] = -x[i] (to 1e-12 of max |x|).
// False when the axis is not symmetric enough to gain: more than n / 2 + 1 magnitudes, or
// repeated coordinates.
struct HostAxis {
  std::vector<double> P;
  std::vector<int> plus, minus;
};

bool parity_axis(const double* x, int64_t n, HostAxis& a) {
  if (n < 2) return false;
  std::vector<int> idx((size_t)n);
  for (int64_t i = 0; i < n; ++i) idx[(size_t)i] = (int)i;
  std::sort(idx.begin(), idx.end(), [&](int i, int k) { return x[i] < x[k]; });
  double mx = 0.0;
  for (int64_t i = 0; i < n; ++i) mx = std::max(mx, std::fabs(x[i]));
  if (!(mx > 0.0) || !(mx < INFINITY)) return false;
  const double tol = 1e-12 * mx;
  for (int64_t i = 1; i < n; ++i)
    if (x[idx[(size_t)i]] - x[idx[(size_t)i - 1]] <= tol) return false;
  a = HostAxis{};
  int64_t lo = 0, hi = n - 1;
  while (lo <= hi) {
    const double s = x[idx[(size_t)lo]] + x[idx[(size_t)hi]];
    if (lo == hi) {
      const double v = x[idx[(size_t)lo]];
      if (std::fabs(v) <= tol) {
        a.P.push_back(0.0);
        a.plus.push_back(idx[(size_t)lo]);
        a.minus.push_back(-1);
      } else if (v > 0) {
        a.P.push_back(v);
        a.plus.push_back(idx[(size_t)lo]);
        a.minus.push_back(-1);
      } else {
        a.P.push_back(-v);
        a.plus.push_back(-1);
        a.minus.push_back(idx[(size_t)lo]);
      }
      break;
    }
    if (std::fabs(s) <= tol) {  // a mirror pair
      a.P.push_back(x[idx[(size_t)hi]]);
      a.plus.push_back(idx[(size_t)hi]);
      a.minus.push_back(idx[(size_t)lo]);
      ++lo;
      --hi;
    } else if (s < 0) {  // the most negative sample has no mirror
      a.P.push_back(-x[idx[(size_t)lo]]);
      a.plus.push_back(-1);
      a.minus.push_back(idx[(size_t)lo]);
      ++lo;
    } else {
      a.P.push_back(x[idx[(size_t)hi]]);
      a.plus.push_back(idx[(size_t)hi]);
      a.minus.push_back(-1);
      --hi;
    }
  }
  // smallest magnitude first: on regular grids plus[] then ascends and minus[] descends by one
  std::reverse(a.P.begin(), a.P.end());
  std::reverse(a.plus.begin(), a.plus.end());
  std::reverse(a.minus.begin(), a.minus.end());
  return (int64_t)a.P.size() <= n / 2 + 1;
}

// Whether the stage-1 rows of output axis `a` end in a magnitude with one sample only whose row
// alone fills a row tile of 32 (the Nyquist sample -n/2 d of an FFT-style axis of even n).
bool nyq_row_worth(const HostAxis& a) {
  const size_t h = a.P.size();
  if (h < 33 || (h - 1) % 32 != 0) return false;
  return (a.plus.back() < 0) != (a.minus.back() < 0) && a.P.back() > 0.0;
}

// Geometry of one direction of a 2-D (or 1-D) transform.
struct Dir {
  const FMatrix* m1;  // stage 1 constant (rows c1, contraction k1)
  const FMatrix* m2;  // stage 2 constant (rows c2, contraction n1), 2-D only
  int64_t n1, k1, c1, c2;
};

Dir direction(const Plan& pl, int dir) {
  Dir d{};
  if (dir == HCIPY_FORWARD) {
    d.m1 = &pl.f1;
    d.m2 = &pl.f2;
    d.n1 = pl.ny;
    d.k1 = pl.nx;
    d.c1 = pl.nu;
    d.c2 = pl.nv;
  } else {
    d.m1 = &pl.b1;
    d.m2 = &pl.b2;
    d.n1 = pl.nv;
    d.k1 = pl.nu;
    d.c1 = pl.nx;
    d.c2 = pl.ny;
  }
  return d;
}

// Workspace carve-up for one direction.
struct Ws {
  size_t split_hi = 0, split_lo = 0, split_lo2 = 0, row_scale = 0, t = 0, z_hi = 0, z_lo = 0, z_lo2 = 0, total = 0;
  int64_t ld1 = 0, n1_pad = 0;
};

Ws workspace_layout(const Plan& pl, const Dir& d, int64_t batch) {
  Ws w;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (size_t)round_up((int64_t)bytes, kAlign);
    return at;
  };
  const int64_t rows1 = batch * (pl.ndim == 2 ? d.n1 : 1);
  if (is_split(pl.prec)) {
    w.ld1 = padded_row(2 * d.k1);
    w.split_hi = take((size_t)rows1 * w.ld1 * 2);
    w.split_lo = take((size_t)rows1 * w.ld1 * 2);
    if (pl.prec == kBF16x3) w.split_lo2 = take((size_t)rows1 * w.ld1 * 2);
    w.row_scale = take((size_t)rows1 * 4);
    w.t = take((size_t)batch * 4);
  }
  if (pl.ndim == 2) {
    w.n1_pad = padded_row(2 * d.n1) / 2;  // complex elements per intermediate row (k-tile multiple)
    const size_t zelems = (size_t)batch * d.c1 * 2 * w.n1_pad;  // reals
    w.z_hi = take(zelems * elem_size(pl.prec));
    if (is_split(pl.prec)) w.z_lo = take(zelems * 2);
    if (pl.prec == kBF16x3) w.z_lo2 = take(zelems * 2);
  }
  w.total = off;
  return w;
}

// Workspace of the parity form of one direction: parity split F (4 items per field, rows
// (y', re / im), x' contiguous), its row scales and per-item scales, stage-1 result G (rows
// (u', re / im), y' contiguous) and the FP64 products H (rows v', complex u').  H is written by
// stage 2, after stage 1 has read the last of F: the two share one region.
struct ParWs {
  size_t f[3] = {0, 0, 0}, rs = 0, t = 0, g[3] = {0, 0, 0}, h = 0, nyq = 0, total = 0;
  int64_t ldf = 0, ldg = 0;
};

ParWs parity_layout(const Plan& pl, const Dir& d, int64_t batch) {
  ParWs w;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (size_t)round_up((int64_t)bytes, kAlign);
    return at;
  };
  const bool fw = d.m1 == &pl.f1;
  const int64_t hk = (fw ? pl.px : pl.pu).h, hn = (fw ? pl.py : pl.pv).h;
  const int64_t hc1 = (fw ? pl.pu : pl.px).h, hc2 = (fw ? pl.pv : pl.py).h;
  const int na = n_arrays(pl.prec);
  w.ldf = padded_row(hk);
  const size_t fbytes = (size_t)round_up((int64_t)batch * 4 * 2 * hn * w.ldf * 2, kAlign);
  const size_t hbytes = (size_t)batch * 4 * hc2 * 2 * hc1 * 8;
  const size_t shared = take(std::max((size_t)na * fbytes, hbytes));
  for (int i = 0; i < na; ++i) w.f[i] = shared + (size_t)i * fbytes;
  w.h = shared;
  w.rs = take((size_t)batch * 4 * 2 * hn * 4);
  w.t = take((size_t)batch * 4 * 4);
  w.ldg = padded_row(hn);
  for (int i = 0; i < na; ++i) w.g[i] = take((size_t)batch * 4 * 2 * hc1 * w.ldg * 2);
  if (pl.nyq[fw ? 0 : 1]) w.nyq = take((size_t)batch * 4 * 2 * hn * sizeof(double));
  w.total = off;
  return w;
}

// Workspace bytes of one direction as a single execution runs it.
size_t direction_ws_bytes(const Plan& pl, int dir, int64_t batch) {
  const Dir d = direction(pl, dir);
  return pl.parity ? parity_layout(pl, d, batch).total : workspace_layout(pl, d, batch).total;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// ev (optional, 4 events): recorded before the split, before stage 1, before stage 2, after stage 2.
int out_mode_for(const Plan& pl, int output) {
  switch (output) {
    case HCIPY_OUTPUT_FIELD: return pl.prec == kF64 ? kOutC128 : kOutC64;
    case HCIPY_OUTPUT_FIELD_C128: return kOutC128;
    case HCIPY_OUTPUT_INTENSITY: return pl.prec == kF64 ? kOutI64 : kOutI32;
    case HCIPY_OUTPUT_INTENSITY_F64: return kOutI64;
    default: return -1;
  }
}

size_t out_elem_bytes(int mode) {
  switch (mode) {
    case kOutC128: return 16;
    case kOutC64: return 8;
    case kOutI32: return 4;
    default: return 8;
  }
}

// Stage arguments of one direction, gathered for a grouped launch instead of being launched.
struct Collected {
  StageArgs a1{}, a2{};
  int b_mode1 = 0;
  int64_t nz2 = 0;  // batch slices of stage 2 (0: 1-D plan, no stage 2)
  int64_t nz1 = 1;  // batch slices of stage 1 (the parity form: four items per field)
};

// Runs one direction in the parity form: parity split, stage 1 (real half-size constants, four
// items per field), stage 2 (likewise, FP64 products), recombination into the four quadrants.
// `only` >= 0 launches one step (0 split, 1 stage 1, 2 stage 2 and recombination).
int run_parity(const Plan& pl, int dir, const void* in, void* out, int64_t batch, double ar, double ai, char* ws,
               hipStream_t s, hipEvent_t* ev, int output, int only) {
  auto mark = [&](int i) -> int {
    if (ev != nullptr) return check_hip(hipEventRecord(ev[i], s), "hipEventRecord");
    return kOk;
  };
  const Dir d = direction(pl, dir);
  const bool fw = dir == HCIPY_FORWARD;
  const PMatrix& p1 = fw ? pl.pf1 : pl.pb1;
  const PMatrix& p2 = fw ? pl.pf2 : pl.pb2;
  const ParAxis& ak = fw ? pl.px : pl.pu;   // input x (contraction of stage 1)
  const ParAxis& an = fw ? pl.py : pl.pv;   // input y (contraction of stage 2)
  const ParAxis& ac1 = fw ? pl.pu : pl.px;  // output x
  const ParAxis& ac2 = fw ? pl.pv : pl.py;  // output y
  const int out_mode = out_mode_for(pl, output);
  HCIPY_REQUIRE(out_mode >= 0, "unknown output kind %d", output);
  const ParWs w = parity_layout(pl, d, batch);
  const int64_t hn = an.h, hk = ak.h, hc1 = ac1.h, hc2 = ac2.h;
  const int na = n_arrays(pl.prec);
  float* rs = reinterpret_cast<float*>(ws + w.rs);
  float* t = reinterpret_cast<float*>(ws + w.t);
  double* H = reinterpret_cast<double*>(ws + w.h);
  const bool nyq = pl.nyq[fw ? 0 : 1];
  double* nyq_vals = nyq ? reinterpret_cast<double*>(ws + w.nyq) : nullptr;
  // stage arguments for the fields b0 ... (stage 1: four items per field; stage 2: the
  // recombining shape cfg2, or FP64 products otherwise)
  auto stage1_args = [&](int64_t b0) {
    StageArgs a{};
    a.tile_cfg = pl.tiles[dir][0];
    a.a = p1.p[0];
    a.a_lo = p1.p[1];
    a.a_lo2 = p1.p[2];
    a.lda = p1.lda;
    a.a_zstride = p1.plane();
    a.a_zshift = 0;
    const int64_t fz = 2 * hn * w.ldf;  // elements per item of F
    a.b = ws + w.f[0] + (size_t)b0 * 4 * fz * 2;
    a.b_lo = na > 1 ? ws + w.f[1] + (size_t)b0 * 4 * fz * 2 : nullptr;
    a.b_lo2 = na > 2 ? ws + w.f[2] + (size_t)b0 * 4 * fz * 2 : nullptr;
    a.ldb = w.ldf;
    a.b_zstride = fz;
    a.M = (int)(nyq ? hc1 - 1 : hc1);
    if (nyq) {
      a.nyq = nyq_vals + b0 * 4 * 2 * hn;
      a.nyq_zstride = 2 * hn;
      a.nyq_row = (int)(hc1 - 1);
    }
    a.N = (int)(2 * hn);
    a.K = (int)hk;
    a.row_scale = rs + b0 * 4 * 2 * hn;
    a.rs_zstride = 2 * hn;
    a.out_mode = kOutSplitT;
    const int64_t gz = 2 * hc1 * w.ldg;  // elements per item of G
    a.out = ws + w.g[0] + (size_t)b0 * 4 * gz * 2;
    a.out_lo = na > 1 ? ws + w.g[1] + (size_t)b0 * 4 * gz * 2 : nullptr;
    a.out_lo2 = na > 2 ? ws + w.g[2] + (size_t)b0 * 4 * gz * 2 : nullptr;
    a.ldo = w.ldg;
    a.out_zstride = gz;
    a.fold = (int)(2 * hn);
    a.t_out = t + b0 * 4;
    a.t_zstride = 1;
    a.t_log2k = ilog2_ceil(hk);
    a.alpha_re = 1.0;
    return a;
  };
  // default stage-2 shape: the recombining one, two waves per parity item (1024^2 -> 512^2 split-BF16:
  // 39.5 us per propagation against 41.2 with one wave per item)
  const int cfg2 = pl.tiles[dir][1] != kTileAuto ? pl.tiles[dir][1] : kRC32x32_w8;
  auto stage2_comb_args = [&](int64_t b0) {
    const size_t ob = out_elem_bytes(out_mode) * (size_t)d.c1 * d.c2;
    StageArgs a{};
    a.tile_cfg = cfg2;
    a.a = p2.p[0];
    a.a_lo = p2.p[1];
    a.a_lo2 = p2.p[2];
    a.lda = p2.lda;
    a.a_zstride = p2.plane();
    const int64_t gz = 2 * hc1 * w.ldg;
    a.b = ws + w.g[0] + (size_t)b0 * 4 * gz * 2;
    a.b_lo = na > 1 ? ws + w.g[1] + (size_t)b0 * 4 * gz * 2 : nullptr;
    a.b_lo2 = na > 2 ? ws + w.g[2] + (size_t)b0 * 4 * gz * 2 : nullptr;
    a.ldb = w.ldg;
    a.b_zstride = gz;
    a.M = (int)hc2;
    a.N = (int)(2 * hc1);
    a.K = (int)hn;
    a.z_scale = t + b0 * 4;
    a.out_mode = out_mode;
    a.out = static_cast<char*>(out) + (size_t)b0 * ob;
    a.cu = ac1;
    a.cv = ac2;
    a.c_nu = (int)d.c1;
    a.c_nv = (int)d.c2;
    a.c_sign = fw ? -1.0 : 1.0;
    a.alpha_re = ar;
    a.alpha_im = ai;
    return a;
  };
  // One persistent launch for a whole execution, where the plan's tile shapes are the ones the
  // fused kernel runs (64 x 64 stage-1 tiles with 64-deep k-tiles, recombining stage 2).
  const int64_t m1 = nyq ? hc1 - 1 : hc1;
  const int64_t tiles64 = ceil_div(m1, 64) * ceil_div(2 * hn, 64) * 4 * batch;
  const int cfg1 = pl.tiles[dir][0] != kTileAuto ? pl.tiles[dir][0] : (tiles64 >= 128 ? kR64x64_w4 : k64x32_w8);
  if (only < 0 && ev == nullptr && pl.fused && pl.sync != nullptr && fused_supported(pl.prec, w.ldf) &&
      cfg1 == kR64x64_w4 && cfg2 == kRC32x32_w4 && m1 > 0 && ceil_div(m1, 64) <= ceil_div(std::max(pl.pu.h, pl.px.h), 64)) {
    FusedArgs f{};
    f.in = in;
    f.nx = (int)d.k1;
    f.ny = (int)d.n1;
    f.ax = ak;
    f.ay = an;
    f.fo = SplitOut{{ws + w.f[0], na > 1 ? ws + w.f[1] : nullptr, na > 2 ? ws + w.f[2] : nullptr}};
    f.ld_out = w.ldf;
    f.row_scale = rs;
    f.nyq_coef = nyq ? pl.nyq_coef64[fw ? 0 : 1] : nullptr;
    f.nyq_out = nyq_vals;
    f.nblk = (int)ceil_div(hn, fused_split_teams(w.ldf));
    f.s1 = stage1_args(0);
    f.s2 = stage2_comb_args(0);
    const int64_t t1 = ceil_div(m1, 64) * ceil_div(2 * hn, 64);
    const int64_t t2 = ceil_div(hc2, 32) * ceil_div(2 * hc1, 32);
    f.n0 = batch * f.nblk;
    f.n1 = 4 * batch * t1;
    f.n2 = batch * t2;
    HCIPY_REQUIRE(f.n1 < INT32_MAX && 4 * batch * ceil_div(2 * hn, 64) < UINT32_MAX, "fused execution of %lld fields",
                  (long long)batch);
    f.sync = pl.sync;
    // one workgroup per compute unit (the stage-2 tiles take most of a CU's LDS), fewer when
    // the busiest queue has fewer tasks
    const int64_t per_queue = std::max(ceil_div(f.n0, 8), std::max(ceil_div(f.n1, 8), ceil_div(f.n2, 8)));
    const int grid = (int)std::max<int64_t>(8, std::min<int64_t>(pl.cus / 8, per_queue) * 8);
    return launch_parity_fused(f, grid, s);
  }
  int st = mark(0);
  if (!st && (only < 0 || only == 0)) {
    SplitOut fo{{ws + w.f[0], na > 1 ? ws + w.f[1] : nullptr, na > 2 ? ws + w.f[2] : nullptr}};
    st = launch_parity_split(pl.prec, static_cast<const float*>(in), (int)d.k1, (int)d.n1, ak, an, batch, fo, w.ldf, rs,
                             nyq ? pl.nyq_coef64[fw ? 0 : 1] : nullptr, nyq_vals, s);
  }
  if (!st) st = mark(1);
  if (st) return st;
  constexpr int64_t kChunk = kMaxGridZ / 4;  // fields per launch (four grid.z slices each)
  for (int64_t b0 = 0; b0 < batch && (only < 0 || only == 1); b0 += kChunk) {
    const int nb = (int)std::min<int64_t>(kChunk, batch - b0);
    const StageArgs a = stage1_args(b0);
    if ((st = launch_stage_real(pl.prec, a, 4 * nb, s))) return st;
  }
  if ((st = mark(2))) return st;
  if (only >= 0 && only != 2) return mark(3);
  // default stage-2 shape of the parity form: the recombining one
  if (is_comb_cfg(cfg2)) {
    // stage 2 with the recombination in its epilogue: one workgroup computes all four items of a
    // tile and writes the output quadrants
    for (int64_t b0 = 0; b0 < batch; b0 += kChunk) {
      const int nb = (int)std::min<int64_t>(kChunk, batch - b0);
      const StageArgs a = stage2_comb_args(b0);
      if ((st = launch_stage_real(pl.prec, a, nb, s))) return st;
    }
    return mark(3);
  }
  for (int64_t b0 = 0; b0 < batch; b0 += kChunk) {
    const int nb = (int)std::min<int64_t>(kChunk, batch - b0);
    StageArgs a{};
    a.tile_cfg = pl.tiles[dir][1];
    a.a = p2.p[0];
    a.a_lo = p2.p[1];
    a.a_lo2 = p2.p[2];
    a.lda = p2.lda;
    a.a_zstride = p2.plane();
    a.a_zshift = 1;
    const int64_t gz = 2 * hc1 * w.ldg;
    a.b = ws + w.g[0] + (size_t)b0 * 4 * gz * 2;
    a.b_lo = na > 1 ? ws + w.g[1] + (size_t)b0 * 4 * gz * 2 : nullptr;
    a.b_lo2 = na > 2 ? ws + w.g[2] + (size_t)b0 * 4 * gz * 2 : nullptr;
    a.ldb = w.ldg;
    a.b_zstride = gz;
    a.M = (int)hc2;
    a.N = (int)(2 * hc1);
    a.K = (int)hn;
    a.z_scale = t + b0 * 4;
    a.out_mode = kOutR64;
    a.out = H + b0 * 4 * hc2 * 2 * hc1;
    a.ldo = 2 * hc1;
    a.out_zstride = hc2 * 2 * hc1;
    a.alpha_re = 1.0;
    if ((st = launch_stage_real(pl.prec, a, 4 * nb, s))) return st;
  }
  st = launch_parity_combine(H, ac1, ac2, (int)d.c1, (int)d.c2, fw ? -1.0 : 1.0, ar, ai, out_mode, out, batch, s);
  if (st) return st;
  return mark(3);
}

// A parity split of one single field that a grouped execution launches together with the splits
// of its other items of the same geometry (launch_parity_split_items).
struct SplitJob {
  int prec, nx, ny;
  ParAxis ax, ay;
  int64_t ld;
  const void* in;
  SplitOut fo;
  float* rs;
};

// Batches up to this size are split with the other items' fields (one job per field); larger
// ones keep a launch of their own, which already has a workgroup row per field.
constexpr int64_t kSplitJobBatchMax = 4;

// Stage arguments of one direction in the parity form for a grouped launch (stage 1 with the
// real constant, recombining stage 2).  The parity split is launched here (recorded in `jobs`
// instead, one job per field, for batches up to kSplitJobBatchMax when `jobs` is given) unless
// `shared` names an earlier item that split the same input with the same axis pairing; stage 1
// then reads that item's split data and row scales.
int collect_parity(const Plan& pl, int dir, const void* in, void* out, int64_t batch, double ar, double ai, char* ws,
                   hipStream_t s, int output, Collected* col, const Collected* shared,
                   std::vector<SplitJob>* jobs = nullptr) {
  const Dir d = direction(pl, dir);
  const bool fw = dir == HCIPY_FORWARD;
  const PMatrix& p1 = fw ? pl.pf1 : pl.pb1;
  const PMatrix& p2 = fw ? pl.pf2 : pl.pb2;
  const ParAxis& ak = fw ? pl.px : pl.pu;
  const ParAxis& an = fw ? pl.py : pl.pv;
  const ParAxis& ac1 = fw ? pl.pu : pl.px;
  const ParAxis& ac2 = fw ? pl.pv : pl.py;
  const int out_mode = out_mode_for(pl, output);
  HCIPY_REQUIRE(out_mode >= 0, "unknown output kind %d", output);
  const ParWs w = parity_layout(pl, d, batch);
  const int64_t hn = an.h, hk = ak.h, hc1 = ac1.h, hc2 = ac2.h;
  const int na = n_arrays(pl.prec);
  StageArgs a{};
  if (shared != nullptr) {
    a.b = shared->a1.b;
    a.b_lo = shared->a1.b_lo;
    a.b_lo2 = shared->a1.b_lo2;
    a.row_scale = shared->a1.row_scale;
  } else {
    SplitOut fo{{ws + w.f[0], na > 1 ? ws + w.f[1] : nullptr, na > 2 ? ws + w.f[2] : nullptr}};
    float* rs = reinterpret_cast<float*>(ws + w.rs);
    if (jobs != nullptr && batch <= kSplitJobBatchMax) {
      // field b: input b nx ny samples on, split rows and row scales 8 hn rows on
      for (int64_t b = 0; b < batch; ++b) {
        const int64_t rows = b * 8 * hn;
        SplitOut fb = fo;
        for (int q = 0; q < 3; ++q)
          if (fb.p[q]) fb.p[q] = static_cast<char*>(fb.p[q]) + (size_t)rows * w.ldf * 2;
        jobs->push_back(SplitJob{pl.prec, (int)d.k1, (int)d.n1, ak, an, w.ldf,
                                 static_cast<const float2*>(in) + b * d.k1 * d.n1, fb, rs + rows});
      }
    } else if (int st = launch_parity_split(pl.prec, static_cast<const float*>(in), (int)d.k1, (int)d.n1, ak, an, batch,
                                            fo, w.ldf, rs, nullptr, nullptr, s)) {
      return st;
    }
    a.b = fo.p[0];
    a.b_lo = fo.p[1];
    a.b_lo2 = fo.p[2];
    a.row_scale = rs;
  }
  float* t = reinterpret_cast<float*>(ws + w.t);
  a.a = p1.p[0];
  a.a_lo = p1.p[1];
  a.a_lo2 = p1.p[2];
  a.lda = p1.lda;
  a.a_zstride = p1.plane();
  a.a_zshift = 0;
  a.ldb = w.ldf;
  a.b_zstride = 2 * hn * w.ldf;
  const bool nyq = pl.nyq[fw ? 0 : 1];
  a.M = (int)(nyq ? hc1 - 1 : hc1);
  if (nyq) {
    a.nyq_coef = pl.nyq_coef32[fw ? 0 : 1];
    a.nyq_row = (int)(hc1 - 1);
  }
  a.N = (int)(2 * hn);
  a.K = (int)hk;
  a.rs_zstride = 2 * hn;
  a.out_mode = kOutSplitT;
  const int64_t gz = 2 * hc1 * w.ldg;
  a.out = ws + w.g[0];
  a.out_lo = na > 1 ? ws + w.g[1] : nullptr;
  a.out_lo2 = na > 2 ? ws + w.g[2] : nullptr;
  a.ldo = w.ldg;
  a.out_zstride = gz;
  a.fold = (int)(2 * hn);
  a.t_out = t;
  a.t_zstride = 1;
  a.t_log2k = ilog2_ceil(hk);
  a.alpha_re = 1.0;
  col->a1 = a;
  col->nz1 = 4 * batch;
  StageArgs c{};
  c.a = p2.p[0];
  c.a_lo = p2.p[1];
  c.a_lo2 = p2.p[2];
  c.lda = p2.lda;
  c.a_zstride = p2.plane();
  c.b = a.out;
  c.b_lo = a.out_lo;
  c.b_lo2 = a.out_lo2;
  c.ldb = w.ldg;
  c.b_zstride = gz;
  c.M = (int)hc2;
  c.N = (int)(2 * hc1);
  c.K = (int)hn;
  c.z_scale = t;
  c.out_mode = out_mode;
  c.out = out;
  c.cu = ac1;
  c.cv = ac2;
  c.c_nu = (int)d.c1;
  c.c_nv = (int)d.c2;
  c.c_sign = fw ? -1.0 : 1.0;
  c.alpha_re = ar;
  c.alpha_im = ai;
  col->a2 = c;
  col->nz2 = batch;
  return kOk;
}

// Same pairing of the input axes (the parity split depends on the grid indices of the mirror
// samples only).
bool same_pairing(const ParAxis& a, const ParAxis& b) {
  if (a.h != b.h || a.n != b.n || a.affine != b.affine) return false;
  if (a.affine) return a.p0 == b.p0 && a.m0 == b.m0 && a.mlo == b.mlo;
  return a.plus == b.plus && a.minus == b.minus;
}

// Runs one direction: data split (split-FP16 only), stage 1, stage 2.  `only` >= 0 launches just
// that kernel (0 split, 1 stage 1, 2 stage 2) on a workspace a full run has filled (timing).
// With `col`, the split is launched and the stage arguments are returned instead of launched.
// With `shared_split` (split-FP16, `col` set), the split is not launched: stage 1 reads the split
// data and row scales that `shared_split->a1` points at (an earlier item of the same input).
int run_direction(const Plan& pl, int dir, const void* in, void* out, int64_t batch, double ar, double ai,
                  char* ws, hipStream_t s, hipEvent_t* ev = nullptr, int output = HCIPY_OUTPUT_FIELD, int only = -1,
                  Collected* col = nullptr, const Collected* shared_split = nullptr) {
  if (pl.parity && col == nullptr && shared_split == nullptr)
    return run_parity(pl, dir, in, out, batch, ar, ai, ws, s, ev, output, only);
  auto mark = [&](int i) -> int {
    if (ev != nullptr) return check_hip(hipEventRecord(ev[i], s), "hipEventRecord");
    return kOk;
  };
  const Dir d = direction(pl, dir);
  const Ws w = workspace_layout(pl, d, batch);
  const int out_mode = out_mode_for(pl, output);
  HCIPY_REQUIRE(out_mode >= 0, "unknown output kind %d", output);
  const size_t es = elem_size(pl.prec);
  const int64_t rows1 = batch * (pl.ndim == 2 ? d.n1 : 1);
  HCIPY_REQUIRE(rows1 <= INT32_MAX / 2, "batch too large (%lld data rows)", (long long)rows1);

  // Stage-1 data operand.
  StageArgs a1{};
  a1.tile_cfg = pl.tiles[dir][0];
  int b_mode = kModePadded;
  float* row_scale = nullptr;
  float* t = nullptr;
  if (int st0 = mark(0)) return st0;
  if (is_split(pl.prec)) {
    void* hi = ws + w.split_hi;
    void* lo = ws + w.split_lo;
    void* lo2 = pl.prec == kBF16x3 ? ws + w.split_lo2 : nullptr;
    row_scale = reinterpret_cast<float*>(ws + w.row_scale);
    t = reinterpret_cast<float*>(ws + w.t);
    if (shared_split != nullptr) {
      hi = const_cast<void*>(shared_split->a1.b);
      lo = const_cast<void*>(shared_split->a1.b_lo);
      lo2 = const_cast<void*>(shared_split->a1.b_lo2);
      row_scale = const_cast<float*>(shared_split->a1.row_scale);
    } else if (only < 0 || only == 0) {
      int st = launch_split_rows(pl.prec, static_cast<const float*>(in), 2 * d.k1, (int)(2 * d.k1), rows1,
                                 SplitOut{{hi, lo, lo2}}, w.ld1, row_scale, s);
      if (st != kOk) return st;
    }
    a1.b = hi;
    a1.b_lo = lo;
    a1.b_lo2 = lo2;
    a1.ldb = w.ld1;
    a1.row_scale = row_scale;
  } else {
    a1.b = in;
    a1.ldb = 2 * d.k1;
    b_mode = (aligned16(in) && ((2 * d.k1 * (int64_t)es) % 16 == 0)) ? kModeAligned : kModeUnaligned;
  }
  a1.a = d.m1->hi;
  a1.a_lo = d.m1->lo;
  a1.a_lo2 = d.m1->lo2;
  a1.lda = d.m1->lda;
  a1.M = (int)(2 * d.c1);
  a1.N = (int)rows1;
  a1.K = (int)(2 * d.k1);
  a1.t_log2k = ilog2_ceil(a1.K);
  if (int st0 = mark(1)) return st0;

  if (pl.ndim == 1) {
    a1.out = out;
    a1.ldo = d.c1;
    a1.nc_layout = 1;
    a1.out_mode = out_mode;
    a1.alpha_re = ar;
    a1.alpha_im = ai;
    if (col != nullptr) {
      col->a1 = a1;
      col->b_mode1 = b_mode;
      col->nz2 = 0;
      return kOk;
    }
    int st1 = (only < 0 || only == 1) ? launch_stage(pl.prec, a1, 1, b_mode, s) : kOk;
    if (!st1) st1 = mark(2);
    if (!st1) st1 = mark(3);
    return st1;
  }

  // Stage 1 writes the intermediate, one (c1 x n1_pad) block per field.
  a1.out = ws + w.z_hi;
  a1.out_lo = is_split(pl.prec) ? ws + w.z_lo : nullptr;
  a1.out_lo2 = pl.prec == kBF16x3 ? ws + w.z_lo2 : nullptr;
  a1.ldo = w.n1_pad;
  a1.fold = (int)d.n1;
  a1.fold_stride = d.c1 * w.n1_pad;
  a1.out_mode = is_split(pl.prec) ? kOutSplit : (pl.prec == kF64 ? kOutC128 : kOutC64);
  a1.t_out = t;
  a1.alpha_re = 1.0;
  a1.alpha_im = 0.0;
  int st = (only < 0 || only == 1) && col == nullptr ? launch_stage(pl.prec, a1, 1, b_mode, s) : kOk;
  if (st != kOk) return st;
  if ((st = mark(2))) return st;

  // Stage 2: one grid.z slice per field.
  StageArgs a2{};
  a2.tile_cfg = pl.tiles[dir][1];
  a2.a = d.m2->hi;
  a2.a_lo = d.m2->lo;
  a2.a_lo2 = d.m2->lo2;
  a2.lda = d.m2->lda;
  a2.b = ws + w.z_hi;
  a2.b_lo = is_split(pl.prec) ? ws + w.z_lo : nullptr;
  a2.b_lo2 = pl.prec == kBF16x3 ? ws + w.z_lo2 : nullptr;
  a2.ldb = 2 * w.n1_pad;
  a2.b_zstride = d.c1 * 2 * w.n1_pad;
  a2.M = (int)(2 * d.c2);
  a2.N = (int)d.c1;
  a2.K = (int)(2 * d.n1);
  a2.out_mode = out_mode;
  a2.ldo = d.c1;
  a2.out_zstride = d.c2 * d.c1;
  a2.alpha_re = ar;
  a2.alpha_im = ai;
  if (col != nullptr) {
    col->a1 = a1;
    col->b_mode1 = b_mode;
    a2.out = out;
    a2.z_scale = t;
    col->a2 = a2;
    col->nz2 = batch;
    return kOk;
  }
  const size_t out_es = out_elem_bytes(out_mode);
  for (int64_t b0 = 0; b0 < batch && (only < 0 || only == 2); b0 += kMaxGridZ) {
    const int nz = (int)std::min<int64_t>(kMaxGridZ, batch - b0);
    StageArgs a = a2;
    a.b = static_cast<const char*>(a2.b) + b0 * a2.b_zstride * es;
    if (a2.b_lo) a.b_lo = static_cast<const char*>(a2.b_lo) + b0 * a2.b_zstride * es;
    if (a2.b_lo2) a.b_lo2 = static_cast<const char*>(a2.b_lo2) + b0 * a2.b_zstride * es;
    a.out = static_cast<char*>(out) + b0 * a2.out_zstride * out_es;
    a.z_scale = t ? t + b0 : nullptr;
    st = launch_stage(pl.prec, a, nz, kModePadded, s);
    if (st != kOk) return st;
  }
  return mark(3);
}

}  // namespace
}  // namespace hcipy

using namespace hcipy;

struct hcipy_mft_plan_s {
  Plan p;
};

// Table of stage arguments for grouped launches: device copy and pinned host mirror.
// Layout: StageArgs[2 * cap] | int64 tile starts[2 * (cap + 4)].
// Launches recorded under stream capture copy their table from a snapshot of their own (the copy
// node reads the host buffer at every replay, so a shared mirror would hand every replay the
// last table written): `snaps` holds `snap_count` tables; captured calls take them in turn until
// hcipy_mft_group_release_captures hands them all back.
struct hcipy_mft_group_s {
  int cap = 0;
  int device = 0;
  size_t bytes = 0;
  char* dev = nullptr;
  char* host = nullptr;
  char* snaps = nullptr;
  int snap_count = 0;
  int snap_used = 0;
  hipEvent_t copied = nullptr;
  bool pending = false;
};

// Workspace of hcipy_mft_execute_noalloc: a file-scope device array, one copy per device that
// loads the library, held until the process ends.  Host code never uses its name as an address:
// each device's copy is looked up once with hipGetSymbolAddress (with that device current) and
// only that pointer is handed to kernels.  Every call on one device shares that device's copy, so
// such calls must be ordered on one stream (overlapping calls race).  The stage kernels and the
// split write every workspace element they later read, so nothing needs clearing between calls.
static __device__ __align__(256) unsigned char scratch[HCIPY_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on first use per device).
int scratch_on_device(int device, void** ptr) {
  HCIPY_REQUIRE(device >= 0 && device < kMaxDevices, "device %d beyond the %d the library tracks", device,
                kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device] == nullptr) {
    void* p = nullptr;
    HCIPY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    HCIPY_REQUIRE(p != nullptr, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  return kOk;
}

// Tables kept for launches recorded under stream capture (at most 64 MiB per group).
constexpr size_t kGroupCaptureSlots = 64;

size_t group_table_bytes(int cap) {
  return round_up((int64_t)(2 * (size_t)cap * sizeof(StageArgs)), 64) + 2 * ((size_t)cap + 4) * sizeof(int64_t);
}

// One grouped launch: `n` consecutive entries of a stage-argument table starting at `first`, with
// their tile prefix sums at `start_off` of the tile-start array.
struct GroupedLaunch {
  int stage, b_mode, cfg, first, n;
  int64_t start_off, tiles;
  bool real = false;  // parity-form launch (launch_stage_grouped_real)
  int nyq_k = 0;      // longest contraction of the items that form an extra stage-1 row
};

// Parity-form grouped launches: stage 1 of every item, then the recombining stage 2 of every item.
void build_parity_launches(int prec, const std::vector<Collected>& col, StageArgs* hargs, int64_t* hstart,
                           std::vector<GroupedLaunch>& launches) {
  int na = 0;
  int64_t ns = 0;
  for (int stage = 0; stage < 2; ++stage) {
    int64_t tiles64 = 0;
    for (const Collected& c : col) {
      const StageArgs& a = stage == 0 ? c.a1 : c.a2;
      tiles64 += ceil_div(a.M, 64) * ceil_div(a.N, 64) * (stage == 0 ? c.nz1 : c.nz2);
    }
    const int cfg = grouped_real_cfg(prec, stage, tiles64);
    int bm, bn;
    real_tile_shape(cfg, &bm, &bn);
    GroupedLaunch L{stage, 0, cfg, na, 0, ns, 0};
    L.real = true;
    int64_t acc = 0;
    for (const Collected& c : col) {
      const StageArgs& a = stage == 0 ? c.a1 : c.a2;
      hargs[na] = a;
      hargs[na].tile_cfg = cfg;
      ++na;
      if (a.nyq_coef != nullptr) L.nyq_k = std::max(L.nyq_k, a.K);
      hstart[ns++] = acc;
      acc += ceil_div(a.M, bm) * ceil_div(a.N, bn) * (stage == 0 ? c.nz1 : c.nz2);
      ++L.n;
    }
    hstart[ns++] = acc;
    L.tiles = acc;
    launches.push_back(L);
  }
}

// Writes the stage arguments of `col` to hargs (stage-1 entries ordered by data mode, one launch
// per mode, then the stage-2 entries) and their tile prefix sums to hstart; appends the launches.
// hargs holds 2 * col.size() entries, hstart 2 * col.size() + 4.
void build_grouped_launches(int prec, const std::vector<Collected>& col, StageArgs* hargs, int64_t* hstart,
                            std::vector<GroupedLaunch>& launches, const int* force_cfg = nullptr) {
  const int n_items = (int)col.size();
  int na = 0;
  int64_t ns = 0;
  for (int stage = 0; stage < 2; ++stage) {
    for (int mode = 0; mode < 3; ++mode) {
      if (stage == 1 && mode != kModePadded) continue;
      auto member = [&](const Collected& c) { return stage == 0 ? c.b_mode1 == mode : c.nz2 != 0; };
      // tile shape of this launch from its longest contraction and its size in 64x64 tiles
      int max_k = 0;
      int64_t tiles64 = 0;
      for (int i = 0; i < n_items; ++i) {
        const Collected& c = col[(size_t)i];
        if (!member(c)) continue;
        const StageArgs& a = stage == 0 ? c.a1 : c.a2;
        max_k = std::max(max_k, a.K);
        tiles64 += ceil_div(a.M, 64) * ceil_div(a.N, 64) * (stage == 0 ? 1 : c.nz2);
      }
      const int cfg = (force_cfg != nullptr && force_cfg[stage] >= 0) ? force_cfg[stage]
                                                                      : grouped_tile_cfg(prec, stage, max_k, tiles64);
      int bm, bn;
      grouped_tile_shape(cfg, &bm, &bn);
      GroupedLaunch L{stage, mode, cfg, na, 0, ns, 0};
      int64_t acc = 0;
      for (int i = 0; i < n_items; ++i) {
        const Collected& c = col[(size_t)i];
        if (!member(c)) continue;
        const StageArgs& a = stage == 0 ? c.a1 : c.a2;
        const int64_t nz = stage == 0 ? 1 : c.nz2;
        hargs[na++] = a;
        hstart[ns++] = acc;
        acc += ceil_div(a.M, bm) * ceil_div(a.N, bn) * nz;
        ++L.n;
      }
      if (L.n == 0) continue;
      hstart[ns++] = acc;
      L.tiles = acc;
      launches.push_back(L);
    }
  }
}
}  // namespace

// Poke set: the base transform of every item and the windowed transform of every single-actuator
// poke, as fixed grouped launches over tables built once at creation (no allocation or copy per
// execution, so executions can be captured).  One device allocation holds the workspaces, the
// base focal fields, the window descriptors and the tables.
struct hcipy_poke_set_s {
  int prec = 0, device = 0;
  char* mem = nullptr;
  size_t mem_bytes = 0;
  int64_t total_rows = 0;
  int max_wx = 0;
  int n_items = 0;
  int64_t n_pokes = 0;
  double flops_windowed = 0.0, flops_full = 0.0;  // real flops (8 per complex MAC) per execution
  void* win_items = nullptr;
  struct Item {
    const Plan* pl;
    const void* field;
    char* ws;        // base transform workspace
    double2* base;   // base focal field (complex128, nv x nu)
    void* base_out;  // base intensity (float32 / float64) or NULL
    int64_t n_focal;
  };
  std::vector<Item> items;
  const StageArgs* base_tab = nullptr;
  const int64_t* base_start = nullptr;
  std::vector<GroupedLaunch> base_launches;
  const StageArgs* poke_tab = nullptr;
  const int64_t* poke_start = nullptr;
  std::vector<GroupedLaunch> poke_launches;
  // item of every workgroup of each poke launch (one dependent load instead of a binary search
  // over tens of thousands of tile starts)
  int* map_mem = nullptr;
  size_t map_cap = 0;  // entries of map_mem (the tile count of the smallest tile shape)
  std::vector<const int*> poke_maps;
  std::vector<Collected> pcol;  // stage arguments of every poke (host copy, for re-tiling)
  StageArgs* ptab_dev = nullptr;
  int64_t* pst_dev = nullptr;
  int cfg[2] = {-1, -1};  // tile shapes of the poke launches (-1: size heuristic)
};

namespace {
// Builds the poke launches with tile shapes `cfg` and uploads their table, tile starts and
// tile-to-item maps.
int upload_poke_tables(hcipy_poke_set_s* h, const int* cfg) {
  const size_t n = h->pcol.size();
  std::vector<StageArgs> tab(2 * n);
  std::vector<int64_t> start(2 * n + 4);
  std::vector<GroupedLaunch> launches;
  build_grouped_launches(h->prec, h->pcol, tab.data(), start.data(), launches, cfg);
  int64_t total_tiles = 0;
  for (const GroupedLaunch& L : launches) total_tiles += L.tiles;
  HCIPY_REQUIRE((size_t)total_tiles <= h->map_cap, "tile map of %lld entries exceeds %zu", (long long)total_tiles,
                h->map_cap);
  std::vector<int> map((size_t)total_tiles);
  std::vector<const int*> maps;
  int64_t off_t = 0;
  for (const GroupedLaunch& L : launches) {
    maps.push_back(h->map_mem + off_t);
    for (int i = 0; i < L.n; ++i)
      for (int64_t t = start[(size_t)(L.start_off + i)]; t < start[(size_t)(L.start_off + i + 1)]; ++t)
        map[(size_t)(off_t + t)] = i;
    off_t += L.tiles;
  }
  HCIPY_HIP(hipMemcpy(h->ptab_dev, tab.data(), 2 * n * sizeof(StageArgs), hipMemcpyHostToDevice));
  HCIPY_HIP(hipMemcpy(h->pst_dev, start.data(), (2 * n + 4) * sizeof(int64_t), hipMemcpyHostToDevice));
  HCIPY_HIP(hipMemcpy(h->map_mem, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice));
  h->poke_launches = launches;
  h->poke_maps = maps;
  h->poke_tab = h->ptab_dev;
  h->poke_start = h->pst_dev;
  return kOk;
}
}  // namespace

namespace {
__global__ __launch_bounds__(256) void c128_intensity_kernel(const double2* __restrict__ z, void* __restrict__ out,
                                                             int f64, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = z[i];
    const double I = v.x * v.x + v.y * v.y;
    if (f64) static_cast<double*>(out)[i] = I;
    else static_cast<float*>(out)[i] = (float)I;
  }
}

// Window of one poke along one axis: [lo, hi] the support pixels (hi < lo when empty), n the grid
// size.  The window is w pixels from an origin that is a multiple of 8 (16-byte aligned matrix
// columns), within the matrix's padded contraction length round_up(n, 64).
int64_t window_origin(int64_t lo, int64_t hi, int64_t n, int64_t w) {
  const int64_t lim = round_up(n, 64);
  if (hi < lo || w >= lim) return 0;
  return std::min(lo / 8 * 8, lim - w);
}

// Workspace of one item's pokes (bytes from the item's base), each array indexed by poke:
// windows [poke][wy][2 wx], row scales [poke][wy], intermediate scales [poke], intermediates
// [poke][nu][2 wy] (one array per operand array of the precision).
struct PokeWs {
  size_t win[3] = {0, 0, 0}, row_scale = 0, t = 0, z[3] = {0, 0, 0}, total = 0;
  size_t win_pitch = 0, z_pitch = 0;  // bytes per poke
};
PokeWs poke_ws(int prec, int64_t n_pokes, int64_t wx, int64_t wy, int64_t nu) {
  PokeWs p;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (size_t)round_up((int64_t)bytes, 256);
    return at;
  };
  const int na = n_arrays(prec);
  const size_t es = elem_size(prec);
  p.win_pitch = (size_t)wy * 2 * wx * es;
  p.z_pitch = (size_t)nu * 2 * wy * es;
  for (int q = 0; q < na; ++q) p.win[q] = take(p.win_pitch * (size_t)n_pokes);
  if (is_split(prec)) {
    p.row_scale = take((size_t)n_pokes * wy * 4);
    p.t = take((size_t)n_pokes * 4);
  }
  for (int q = 0; q < na; ++q) p.z[q] = take(p.z_pitch * (size_t)n_pokes);
  p.total = off;
  return p;
}
}  // namespace

extern "C" {

int hcipy_abi_version(void) { return 1; }

const char* hcipy_get_last_error(void) { return g_last_error; }

int hcipy_mft_plan_create(hcipy_mft_plan* plan, int ndim, int precision, const double* x_in, int64_t nx,
                          const double* y_in, int64_t ny, const double* u_out, int64_t nu, const double* v_out,
                          int64_t nv, void* stream) {
  HCIPY_REQUIRE(plan != nullptr, "plan pointer is NULL");
  *plan = nullptr;
  HCIPY_REQUIRE(ndim == 1 || ndim == 2, "The input_grid must be one- or two-dimensional (got ndim=%d).", ndim);
  HCIPY_REQUIRE(precision == kF64 || precision == kF32 || precision == kF16x2 || precision == kBF16x3,
                "unknown precision %d", precision);
  HCIPY_REQUIRE(nx > 0 && nu > 0 && x_in != nullptr && u_out != nullptr, "empty grid");
  if (ndim == 1) {
    HCIPY_REQUIRE(ny == 1 && nv == 1, "1-D transform requires ny == nv == 1");
  } else {
    HCIPY_REQUIRE(ny > 0 && nv > 0 && y_in != nullptr && v_out != nullptr, "empty grid");
  }
  const int64_t lim = (int64_t)1 << 24;
  HCIPY_REQUIRE(nx < lim && ny < lim && nu < lim && nv < lim, "grid dimension too large");
  if (int st = refuse_under_capture(static_cast<hipStream_t>(stream), "hcipy_mft_plan_create")) return st;

  hcipy_mft_plan_s* h = new (std::nothrow) hcipy_mft_plan_s();
  HCIPY_REQUIRE(h != nullptr, "out of host memory");
  Plan& pl = h->p;
  pl.ndim = ndim;
  pl.prec = precision;
  pl.nx = nx;
  pl.ny = ny;
  pl.nu = nu;
  pl.nv = nv;
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto fail = [&](int st) {
    hcipy_mft_plan_destroy(h);
    return st;
  };
  if (hipGetDevice(&pl.device) != hipSuccess) return fail(check_hip(hipGetLastError(), "hipGetDevice"));

  const int64_t ncoord = nx + ny + nu + nv;
  if (int st = check_hip(hipMalloc(&pl.coords, ncoord * sizeof(double)), "hipMalloc(coords)")) return fail(st);
  double* dx = pl.coords;
  double* dy = dx + nx;
  double* du = dy + ny;
  double* dv = du + nu;
  int st = check_hip(hipMemcpyAsync(dx, x_in, nx * sizeof(double), hipMemcpyHostToDevice, s), "copy x");
  if (!st) st = check_hip(hipMemcpyAsync(du, u_out, nu * sizeof(double), hipMemcpyHostToDevice, s), "copy u");
  if (!st && ndim == 2) {
    st = check_hip(hipMemcpyAsync(dy, y_in, ny * sizeof(double), hipMemcpyHostToDevice, s), "copy y");
    if (!st) st = check_hip(hipMemcpyAsync(dv, v_out, nv * sizeof(double), hipMemcpyHostToDevice, s), "copy v");
  }
  if (st) return fail(st);

  // forward stage 1: rows u, contraction x: exp(-i x u)   (M2^T)
  // forward stage 2: rows v, contraction y: exp(-i v y)   (M1)
  // backward stage 1: rows x, contraction u: exp(+i x u)  (M2^H)
  // backward stage 2: rows y, contraction v: exp(+i v y)  (M1^H)
  st = make_matrix(pl, pl.f1, du, (int)nu, dx, (int)nx, -1.0, s);
  if (!st) st = make_matrix(pl, pl.b1, dx, (int)nx, du, (int)nu, +1.0, s);
  if (!st && ndim == 2) {
    st = make_matrix(pl, pl.f2, dv, (int)nv, dy, (int)ny, -1.0, s);
    if (!st) st = make_matrix(pl, pl.b2, dy, (int)ny, dv, (int)nv, +1.0, s);
  }
  // parity form: 2-D split precisions between grids symmetric about 0 on every axis
  HostAxis hx, hy, hu, hv;
  pl.parity_ok = !st && ndim == 2 && is_split(precision) && parity_axis(x_in, nx, hx) && parity_axis(y_in, ny, hy) &&
                 parity_axis(u_out, nu, hu) && parity_axis(v_out, nv, hv);
  if (pl.parity_ok) {
    // one device block: per axis plus[h], minus[h] (int), then per axis P[h] (double)
    const HostAxis* ha[4] = {&hx, &hy, &hu, &hv};
    ParAxis* pa[4] = {&pl.px, &pl.py, &pl.pu, &pl.pv};
    size_t ints = 0, dbls = 0;
    for (const HostAxis* a : ha) {
      ints += 2 * a->P.size();
      dbls += a->P.size();
    }
    const size_t ibytes = round_up((int64_t)(ints * sizeof(int)), 16);
    std::vector<char> host(ibytes + dbls * sizeof(double));
    int* hi = reinterpret_cast<int*>(host.data());
    double* hd = reinterpret_cast<double*>(host.data() + ibytes);
    st = check_hip(hipMalloc(&pl.ptables, host.size()), "hipMalloc(parity tables)");
    if (!st) {
      size_t io = 0, dof = 0;
      for (int i = 0; i < 4; ++i) {
        const HostAxis& a = *ha[i];
        const size_t h = a.P.size();
        std::copy(a.plus.begin(), a.plus.end(), hi + io);
        std::copy(a.minus.begin(), a.minus.end(), hi + io + h);
        std::copy(a.P.begin(), a.P.end(), hd + dof);
        pa[i]->plus = static_cast<const int*>(pl.ptables) + io;
        pa[i]->minus = static_cast<const int*>(pl.ptables) + io + h;
        pa[i]->P = reinterpret_cast<const double*>(static_cast<char*>(pl.ptables) + ibytes) + dof;
        pa[i]->h = (int)h;
        pa[i]->n = (int)(i == 0 ? nx : i == 1 ? ny : i == 2 ? nu : nv);
        // affine description of the tables when it reproduces them exactly
        int p0 = 0, m0 = 0, mlo = 0;
        for (size_t j = 0; j < h; ++j)
          if (a.plus[j] >= 0) {
            p0 = a.plus[j] - (int)j;
            break;
          }
        for (size_t j = 0; j < h; ++j)
          if (a.minus[j] >= 0) {
            m0 = a.minus[j] + (int)j;
            mlo = (int)j;
            break;
          }
        bool aff = true;
        for (size_t j = 0; j < h && aff; ++j) {
          const int pj = p0 + (int)j < pa[i]->n ? p0 + (int)j : -1;
          const int mj = ((int)j >= mlo && m0 - (int)j >= 0) ? m0 - (int)j : -1;
          aff = (pj == a.plus[j] || (a.plus[j] < 0 && pj < 0)) && mj == a.minus[j] && p0 + (int)j >= 0;
        }
        pa[i]->affine = aff ? 1 : 0;
        pa[i]->p0 = p0;
        pa[i]->m0 = m0;
        pa[i]->mlo = mlo;
        io += 2 * h;
        dof += h;
      }
      st = check_hip(hipMemcpy(pl.ptables, host.data(), host.size(), hipMemcpyHostToDevice), "copy parity tables");
    }
    if (!st) st = make_parity(pl, pl.pf1, pl.pu.P, pl.pu.h, pl.px.P, pl.px.h, s);
    if (!st) st = make_parity(pl, pl.pf2, pl.pv.P, pl.pv.h, pl.py.P, pl.py.h, s);
    if (!st) st = make_parity(pl, pl.pb1, pl.px.P, pl.px.h, pl.pu.P, pl.pu.h, s);
    if (!st) st = make_parity(pl, pl.pb2, pl.py.P, pl.py.h, pl.pv.P, pl.pv.h, s);
    // Nyquist rows of stage 1: forward output u (contraction x), backward output x (contraction u)
    const HostAxis* c1[2] = {&hu, &hx};
    const HostAxis* kk[2] = {&hx, &hu};
    for (int dir = 0; dir < 2 && !st; ++dir) {
      if (!nyq_row_worth(*c1[dir]) || kk[dir]->P.size() > 4096) continue;
      const size_t hk = kk[dir]->P.size();
      std::vector<double> cd(2 * hk);
      std::vector<float> cf(2 * hk);
      const double pn = c1[dir]->P.back();
      for (size_t j = 0; j < hk; ++j) {
        cd[j] = std::cos(pn * kk[dir]->P[j]);
        cd[hk + j] = std::sin(pn * kk[dir]->P[j]);
      }
      for (size_t j = 0; j < 2 * hk; ++j) cf[j] = (float)cd[j];
      st = check_hip(hipMalloc(&pl.nyq_coef64[dir], cd.size() * sizeof(double)), "hipMalloc(nyquist row)");
      if (!st) st = check_hip(hipMalloc(&pl.nyq_coef32[dir], cf.size() * sizeof(float)), "hipMalloc(nyquist row)");
      if (!st) {
        pl.matrix_bytes += cd.size() * sizeof(double) + cf.size() * sizeof(float);
        st = check_hip(hipMemcpy(pl.nyq_coef64[dir], cd.data(), cd.size() * sizeof(double), hipMemcpyHostToDevice),
                       "copy nyquist row");
      }
      if (!st)
        st = check_hip(hipMemcpy(pl.nyq_coef32[dir], cf.data(), cf.size() * sizeof(float), hipMemcpyHostToDevice),
                       "copy nyquist row");
      if (!st) pl.nyq[dir] = true;
    }
    pl.parity = true;
    // The fused execution (parity_fused_kernel) is disabled: its launches did not finish on MI355X
    // (kFusedEnabled).  With it enabled the plan would hold its counter block: fixed lines plus one
    // per stage-1 row tile (64 rows of the larger of the two stage-1 output axes).
    if (!st && precision == kF16x2 && kFusedEnabled) {
      const int row_tiles = (int)ceil_div(std::max(pl.pu.h, pl.px.h), 64);
      const size_t bytes = (size_t)(kSyncFixedLines + row_tiles) * kSyncLine * sizeof(unsigned);
      st = check_hip(hipMalloc(&pl.sync, bytes), "hipMalloc(fused counters)");
      if (!st) st = check_hip(hipMemset(pl.sync, 0, bytes), "clear fused counters");
      if (!st) st = check_hip(hipDeviceGetAttribute(&pl.cus, hipDeviceAttributeMultiprocessorCount, pl.device),
                              "hipDeviceGetAttribute(compute units)");
      if (!st) pl.fused = pl.cus >= 8;
    }
  }
  if (!st) st = check_hip(hipStreamSynchronize(s), "plan creation");
  if (st) return fail(st);
  *plan = h;
  return kOk;
}

int hcipy_mft_plan_destroy(hcipy_mft_plan plan) {
  if (plan == nullptr) return kOk;
  Plan& pl = plan->p;
  free_matrix(pl.f1);
  free_matrix(pl.f2);
  free_matrix(pl.b1);
  free_matrix(pl.b2);
  free_parity(pl.pf1);
  free_parity(pl.pf2);
  free_parity(pl.pb1);
  free_parity(pl.pb2);
  if (pl.ptables) (void)hipFree(pl.ptables);
  if (pl.sync) (void)hipFree(pl.sync);
  for (double* c : pl.nyq_coef64)
    if (c) (void)hipFree(c);
  for (float* c : pl.nyq_coef32)
    if (c) (void)hipFree(c);
  if (pl.coords) (void)hipFree(pl.coords);
  delete plan;
  return kOk;
}

int hcipy_mft_workspace_size(hcipy_mft_plan plan, int64_t batch, size_t* bytes) {
  HCIPY_REQUIRE(plan != nullptr && bytes != nullptr, "NULL argument");
  HCIPY_REQUIRE(batch >= 0, "negative batch");
  const Plan& pl = plan->p;
  const Ws f = workspace_layout(pl, direction(pl, HCIPY_FORWARD), batch);
  const Ws b = workspace_layout(pl, direction(pl, HCIPY_BACKWARD), batch);
  *bytes = std::max(f.total, b.total);
  // enough for the parity form too, so a workspace stays valid when it is switched on or off
  if (pl.parity_ok) {
    *bytes = std::max(*bytes, parity_layout(pl, direction(pl, HCIPY_FORWARD), batch).total);
    *bytes = std::max(*bytes, parity_layout(pl, direction(pl, HCIPY_BACKWARD), batch).total);
  }
  return kOk;
}

int hcipy_mft_plan_set_parity(hcipy_mft_plan plan, int enable) {
  HCIPY_REQUIRE(plan != nullptr, "plan is NULL");
  Plan& pl = plan->p;
  HCIPY_REQUIRE(!enable || pl.parity_ok,
                "the parity form needs a two-dimensional split-precision plan between grids whose axes are all "
                "symmetric about 0 (every sample but at most two paired with its mirror image)");
  if (pl.parity != (enable != 0)) {
    pl.parity = enable != 0;
    for (auto& d : pl.tiles) d[0] = d[1] = kTileAuto;  // tile shapes of the other form do not carry over
  }
  return kOk;
}

int hcipy_mft_plan_set_fused(hcipy_mft_plan plan, int enable) {
  HCIPY_REQUIRE(plan != nullptr, "plan is NULL");
  Plan& pl = plan->p;
  HCIPY_REQUIRE(!enable || kFusedEnabled,
                "the fused parity execution is disabled in this build: its persistent launches did not finish "
                "on MI355X and the rewritten task loop has not run on a GPU; single executions run as three "
                "launches");
  HCIPY_REQUIRE(!enable || (pl.sync != nullptr && pl.cus >= 8),
                "the fused parity execution needs a split-FP16 plan in which the parity form is available");
  pl.fused = enable != 0;
  return kOk;
}

int hcipy_mft_plan_get_fused(hcipy_mft_plan plan, int* available, int* enabled, unsigned* timeouts) {
  HCIPY_REQUIRE(plan != nullptr, "plan is NULL");
  const Plan& pl = plan->p;
  if (available) *available = (pl.sync != nullptr && pl.cus >= 8) ? 1 : 0;
  if (enabled) *enabled = pl.fused ? 1 : 0;
  if (timeouts) {
    *timeouts = 0;
    if (pl.sync != nullptr) {
      unsigned* line = pl.sync + 17 * kSyncLine;
      HCIPY_HIP(hipDeviceSynchronize());
      HCIPY_HIP(hipMemcpy(timeouts, line, sizeof(unsigned), hipMemcpyDeviceToHost));
      if (*timeouts != 0) HCIPY_HIP(hipMemset(line, 0, sizeof(unsigned)));
    }
  }
  return kOk;
}

int hcipy_mft_plan_get_parity(hcipy_mft_plan plan, int* available, int* enabled) {
  HCIPY_REQUIRE(plan != nullptr, "plan is NULL");
  if (available) *available = plan->p.parity_ok ? 1 : 0;
  if (enabled) *enabled = plan->p.parity ? 1 : 0;
  return kOk;
}

int hcipy_mft_plan_bytes(hcipy_mft_plan plan, size_t* bytes) {
  HCIPY_REQUIRE(plan != nullptr && bytes != nullptr, "NULL argument");
  *bytes = plan->p.matrix_bytes;
  return kOk;
}

int hcipy_mft_plan_set_tiles(hcipy_mft_plan plan, int direction_, int stage, int tile_cfg) {
  HCIPY_REQUIRE(plan != nullptr, "plan is NULL");
  HCIPY_REQUIRE(direction_ == HCIPY_FORWARD || direction_ == HCIPY_BACKWARD, "unknown direction %d", direction_);
  HCIPY_REQUIRE(stage == 0 || stage == 1, "stage must be 0 or 1, got %d", stage);
  HCIPY_REQUIRE(tile_cfg == HCIPY_TILES_AUTO || (plan->p.parity ? real_cfg_supported(plan->p.prec, tile_cfg, stage)
                                                                 : tile_cfg_supported(plan->p.prec, tile_cfg)),
                "tile configuration %d is not available for precision %d%s", tile_cfg, plan->p.prec,
                plan->p.parity ? " in the parity form" : "");
  plan->p.tiles[direction_][stage] = tile_cfg;
  return kOk;
}

int hcipy_mft_plan_get_tiles(hcipy_mft_plan plan, int direction_, int stage, int* tile_cfg) {
  HCIPY_REQUIRE(plan != nullptr && tile_cfg != nullptr, "NULL argument");
  HCIPY_REQUIRE(direction_ == HCIPY_FORWARD || direction_ == HCIPY_BACKWARD, "unknown direction %d", direction_);
  HCIPY_REQUIRE(stage == 0 || stage == 1, "stage must be 0 or 1, got %d", stage);
  *tile_cfg = plan->p.tiles[direction_][stage];
  return kOk;
}

int hcipy_mft_execute(hcipy_mft_plan plan, int direction_, const void* in, void* out, int64_t batch,
                      double alpha_re, double alpha_im, void* workspace, size_t workspace_bytes, void* stream) {
  return hcipy_mft_execute_ex(plan, direction_, in, out, batch, alpha_re, alpha_im, HCIPY_OUTPUT_FIELD, workspace,
                              workspace_bytes, stream);
}

int hcipy_mft_execute_ex(hcipy_mft_plan plan, int direction_, const void* in, void* out, int64_t batch,
                         double alpha_re, double alpha_im, int output, void* workspace, size_t workspace_bytes,
                         void* stream) {
  HCIPY_REQUIRE(plan != nullptr, "plan is NULL");
  HCIPY_REQUIRE(direction_ == HCIPY_FORWARD || direction_ == HCIPY_BACKWARD, "bad direction %d", direction_);
  HCIPY_REQUIRE(batch >= 0, "negative batch");
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(in != nullptr && out != nullptr, "NULL data pointer");
  const Plan& pl = plan->p;
  const size_t need = direction_ws_bytes(pl, direction_, batch);
  HCIPY_REQUIRE(workspace_bytes >= need, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
  HCIPY_REQUIRE(need == 0 || workspace != nullptr, "workspace is NULL");
  HCIPY_REQUIRE(aligned16(workspace), "workspace must be 16-byte aligned");
  int cur = -1;
  HCIPY_HIP(hipGetDevice(&cur));
  HCIPY_REQUIRE(cur == pl.device, "plan was created on device %d but device %d is current", pl.device, cur);
  return run_direction(pl, direction_, in, out, batch, alpha_re, alpha_im, static_cast<char*>(workspace),
                       static_cast<hipStream_t>(stream), nullptr, output);
}

int hcipy_mft_execute_noalloc(hcipy_mft_plan plan, int direction_, const void* in, void* out, int64_t batch,
                              double alpha_re, double alpha_im, int output, void* stream) {
  HCIPY_REQUIRE(plan != nullptr, "plan is NULL");
  HCIPY_REQUIRE(direction_ == HCIPY_FORWARD || direction_ == HCIPY_BACKWARD, "bad direction %d", direction_);
  HCIPY_REQUIRE(batch >= 0, "negative batch");
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(in != nullptr && out != nullptr, "NULL data pointer");
  const Plan& pl = plan->p;
  const size_t need = direction_ws_bytes(pl, direction_, batch);
  if (need > (size_t)HCIPY_NOALLOC_SCRATCH_BYTES) {
    set_last_error("hcipy_mft_execute_noalloc: %lld fields need %zu bytes of workspace, above the %d-byte static "
                   "scratch; use hcipy_mft_execute_ex with a caller workspace or a smaller batch",
                   (long long)batch, need, (int)HCIPY_NOALLOC_SCRATCH_BYTES);
    return kOutOfRange;
  }
  int cur = -1;
  HCIPY_HIP(hipGetDevice(&cur));
  HCIPY_REQUIRE(cur == pl.device, "plan was created on device %d but device %d is current", pl.device, cur);
  void* ws = nullptr;
  if (int st = scratch_on_device(cur, &ws)) return st;
  return run_direction(pl, direction_, in, out, batch, alpha_re, alpha_im, static_cast<char*>(ws),
                       static_cast<hipStream_t>(stream), nullptr, output);
}

int hcipy_mft_profile(hcipy_mft_plan plan, int direction_, const void* in, void* out, int64_t batch, double alpha_re,
                      double alpha_im, int output, void* workspace, size_t workspace_bytes, int repeats,
                      float* kernel_ms, void* stream) {
  HCIPY_REQUIRE(plan != nullptr && kernel_ms != nullptr && repeats > 0, "bad argument");
  HCIPY_REQUIRE(direction_ == HCIPY_FORWARD || direction_ == HCIPY_BACKWARD, "bad direction %d", direction_);
  HCIPY_REQUIRE(batch > 0 && in != nullptr && out != nullptr, "bad data arguments");
  const Plan& pl = plan->p;
  const size_t need = direction_ws_bytes(pl, direction_, batch);
  HCIPY_REQUIRE(workspace_bytes >= need, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
  HCIPY_REQUIRE(aligned16(workspace), "workspace must be 16-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t ev[2];
  int created = 0;
  int st = kOk;
  for (; created < 2 && !st; ++created) st = check_hip(hipEventCreate(&ev[created]), "hipEventCreate");
  // one full run fills the workspace, then each kernel is timed over `repeats` back-to-back launches
  if (!st)
    st = run_direction(pl, direction_, in, out, batch, alpha_re, alpha_im, static_cast<char*>(workspace), s, nullptr,
                       output);
  for (int k = 0; k < 3 && !st; ++k) {
    st = check_hip(hipEventRecord(ev[0], s), "hipEventRecord");
    for (int r = 0; r < repeats && !st; ++r)
      st = run_direction(pl, direction_, in, out, batch, alpha_re, alpha_im, static_cast<char*>(workspace), s,
                         nullptr, output, k);
    if (!st) st = check_hip(hipEventRecord(ev[1], s), "hipEventRecord");
    if (!st) st = check_hip(hipEventSynchronize(ev[1]), "hipEventSynchronize");
    float ms = 0.0f;
    if (!st) st = check_hip(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime");
    kernel_ms[k] = ms / (float)repeats;
  }
  for (int i = 0; i < created; ++i) (void)hipEventDestroy(ev[i]);
  return st;
}

int hcipy_mft_profile_pipeline(hcipy_mft_plan plan, int direction_, const void* in, void* out, int64_t batch,
                               double alpha_re, double alpha_im, int output, void* workspace,
                               size_t workspace_bytes, int repeats, float* kernel_ms, void* stream) {
  HCIPY_REQUIRE(plan != nullptr && kernel_ms != nullptr && repeats > 0, "bad argument");
  HCIPY_REQUIRE(direction_ == HCIPY_FORWARD || direction_ == HCIPY_BACKWARD, "bad direction %d", direction_);
  HCIPY_REQUIRE(batch > 0 && in != nullptr && out != nullptr, "bad data arguments");
  const Plan& pl = plan->p;
  const size_t need = direction_ws_bytes(pl, direction_, batch);
  HCIPY_REQUIRE(workspace_bytes >= need, "workspace too small: %zu < %zu bytes", workspace_bytes, need);
  HCIPY_REQUIRE(aligned16(workspace), "workspace must be 16-byte aligned");
  if (int st = refuse_under_capture(static_cast<hipStream_t>(stream), "hcipy_mft_profile_pipeline")) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t ev[2];
  int created = 0;
  int st = kOk;
  for (; created < 2 && !st; ++created) st = check_hip(hipEventCreate(&ev[created]), "hipEventCreate");
  if (!st)
    st = run_direction(pl, direction_, in, out, batch, alpha_re, alpha_im, static_cast<char*>(workspace), s, nullptr,
                       output);
  // prefix k: `repeats` executions of steps 0..k back to back between two events; step k's time
  // is the difference of consecutive prefixes
  float prev = 0.0f;
  for (int k = 0; k < 3 && !st; ++k) {
    st = check_hip(hipEventRecord(ev[0], s), "hipEventRecord");
    for (int r = 0; r < repeats && !st; ++r)
      for (int j = 0; j <= k && !st; ++j)
        st = run_direction(pl, direction_, in, out, batch, alpha_re, alpha_im, static_cast<char*>(workspace), s,
                           nullptr, output, j);
    if (!st) st = check_hip(hipEventRecord(ev[1], s), "hipEventRecord");
    if (!st) st = check_hip(hipEventSynchronize(ev[1]), "hipEventSynchronize");
    float ms = 0.0f;
    if (!st) st = check_hip(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime");
    kernel_ms[k] = (ms - prev) / (float)repeats;
    prev = ms;
  }
  for (int i = 0; i < created; ++i) (void)hipEventDestroy(ev[i]);
  return st;
}

int hcipy_mft_group_create(hcipy_mft_group* group, int capacity) {
  HCIPY_REQUIRE(group != nullptr && capacity > 0 && capacity <= (1 << 20), "bad argument");
  *group = nullptr;
  hcipy_mft_group g = new (std::nothrow) hcipy_mft_group_s;
  HCIPY_REQUIRE(g != nullptr, "out of host memory");
  g->cap = capacity;
  g->bytes = group_table_bytes(capacity);
  int st = check_hip(hipGetDevice(&g->device), "hipGetDevice");
  if (!st) st = check_hip(hipMalloc(&g->dev, g->bytes), "hipMalloc(group table)");
  if (!st) st = check_hip(hipHostMalloc(&g->host, g->bytes), "hipHostMalloc(group table)");
  if (!st) {
    g->snap_count = (int)std::max<size_t>(1, std::min<size_t>(kGroupCaptureSlots, (size_t(64) << 20) / g->bytes));
    st = check_hip(hipHostMalloc(&g->snaps, g->bytes * (size_t)g->snap_count), "hipHostMalloc(group snapshots)");
  }
  if (!st) st = check_hip(hipEventCreateWithFlags(&g->copied, hipEventDisableTiming), "hipEventCreate");
  if (st) {
    hcipy_mft_group_destroy(g);
    return st;
  }
  *group = g;
  return kOk;
}

int hcipy_mft_group_destroy(hcipy_mft_group group) {
  if (group == nullptr) return kOk;
  if (group->copied) {
    (void)hipEventSynchronize(group->copied);
    (void)hipEventDestroy(group->copied);
  }
  if (group->dev) (void)hipFree(group->dev);
  if (group->host) (void)hipHostFree(group->host);
  if (group->snaps) (void)hipHostFree(group->snaps);
  delete group;
  return kOk;
}

int hcipy_mft_group_release_captures(hcipy_mft_group group, int* held) {
  HCIPY_REQUIRE(group != nullptr, "NULL group");
  if (held != nullptr) *held = group->snap_used;
  group->snap_used = 0;
  return kOk;
}

int hcipy_mft_execute_grouped(hcipy_mft_group group, const hcipy_mft_group_item* items, int n_items, void* stream) {
  HCIPY_REQUIRE(group != nullptr && items != nullptr, "NULL argument");
  HCIPY_REQUIRE(n_items > 0 && n_items <= group->cap, "%d items for a group of capacity %d", n_items, group->cap);
  hipStream_t s = static_cast<hipStream_t>(stream);
  int cur = 0;
  HCIPY_HIP(hipGetDevice(&cur));
  HCIPY_REQUIRE(cur == group->device, "the group was created on device %d, current device is %d", group->device, cur);
  const int prec = items[0].plan != nullptr ? items[0].plan->p.prec : -1;
  std::vector<Collected> col((size_t)n_items);
  // the parity form when every item's plan runs it (2-D split plans between symmetric grids)
  bool parity = true;
  for (int i = 0; i < n_items && parity; ++i) parity = items[i].plan != nullptr && items[i].plan->p.parity;
  std::vector<SplitJob> jobs;
  for (int i = 0; i < n_items && parity; ++i) {
    const hcipy_mft_group_item& it = items[i];
    HCIPY_REQUIRE(it.in != nullptr && it.out != nullptr && it.batch > 0, "bad item %d", i);
    const Plan& pl = it.plan->p;
    HCIPY_REQUIRE(pl.prec == prec, "item %d: all items of a grouped launch share one precision", i);
    HCIPY_REQUIRE(pl.device == cur, "item %d: plan lives on device %d", i, pl.device);
    HCIPY_REQUIRE(it.direction == HCIPY_FORWARD || it.direction == HCIPY_BACKWARD, "item %d: bad direction", i);
    const Dir di = direction(pl, it.direction);
    HCIPY_REQUIRE(it.workspace_bytes >= parity_layout(pl, di, it.batch).total && aligned16(it.workspace),
                  "item %d: workspace too small", i);
    const bool fw = it.direction == HCIPY_FORWARD;
    const Collected* shared = nullptr;
    for (int j = 0; j < i && shared == nullptr; ++j) {
      const hcipy_mft_group_item& o = items[j];
      const Plan& pj = o.plan->p;
      const bool fj = o.direction == HCIPY_FORWARD;
      if (o.in == it.in && o.batch == it.batch && same_pairing(fw ? pl.px : pl.pu, fj ? pj.px : pj.pu) &&
          same_pairing(fw ? pl.py : pl.pv, fj ? pj.py : pj.pv) &&
          parity_layout(pj, direction(pj, o.direction), o.batch).ldf == parity_layout(pl, di, it.batch).ldf)
        shared = &col[(size_t)j];
    }
    const int st = collect_parity(pl, it.direction, it.in, it.out, it.batch, it.alpha_re, it.alpha_im,
                                  static_cast<char*>(it.workspace), s, it.output, &col[(size_t)i], shared, &jobs);
    if (st != kOk) return st;
  }
  // The recorded splits of single fields, in launches of up to kSplitItemsMax items of one
  // geometry (for example one per wavelength of a broadband source): all are on the stream before
  // the stage-1 launch that reads them.
  for (size_t i = 0; i < jobs.size(); ++i) {
    if (jobs[i].in == nullptr) continue;  // taken by an earlier launch
    const SplitJob& g = jobs[i];
    SplitItems si{};
    int n = 0;
    for (size_t j = i; j < jobs.size() && n < kSplitItemsMax; ++j) {
      SplitJob& o = jobs[j];
      if (o.in == nullptr || o.prec != g.prec || o.nx != g.nx || o.ny != g.ny || o.ld != g.ld ||
          !same_pairing(o.ax, g.ax) || !same_pairing(o.ay, g.ay))
        continue;
      si.in[n] = o.in;
      si.o[n] = o.fo;
      si.rs[n] = o.rs;
      ++n;
      if (j != i) o.in = nullptr;
    }
    const int st = launch_parity_split_items(g.prec, si, n, g.nx, g.ny, g.ax, g.ay, g.ld, s);
    jobs[i].in = nullptr;
    if (st != kOk) return st;
  }
  for (int i = 0; i < n_items && !parity; ++i) {
    const hcipy_mft_group_item& it = items[i];
    HCIPY_REQUIRE(it.plan != nullptr && it.in != nullptr && it.out != nullptr && it.batch > 0, "bad item %d", i);
    const Plan& pl = it.plan->p;
    HCIPY_REQUIRE(pl.prec == prec, "item %d: all items of a grouped launch share one precision", i);
    HCIPY_REQUIRE(pl.device == cur, "item %d: plan lives on device %d", i, pl.device);
    HCIPY_REQUIRE(it.direction == HCIPY_FORWARD || it.direction == HCIPY_BACKWARD, "item %d: bad direction", i);
    const Dir di = direction(pl, it.direction);
    const Ws w = workspace_layout(pl, di, it.batch);
    HCIPY_REQUIRE(it.workspace_bytes >= w.total && aligned16(it.workspace), "item %d: workspace too small", i);
    // The data split (split-FP16) is launched here, per distinct input: an item that reads the
    // same input as an earlier item, with the same rows, contraction length and split row pitch
    // (for example one pupil field sent to several focal grids), reuses that item's split.
    const Collected* shared = nullptr;
    if (is_split(prec)) {
      for (int j = 0; j < i && shared == nullptr; ++j) {
        const hcipy_mft_group_item& o = items[j];
        const Plan& pj = o.plan->p;
        const Dir dj = direction(pj, o.direction);
        const int64_t rows_i = it.batch * (pl.ndim == 2 ? di.n1 : 1);
        const int64_t rows_j = o.batch * (pj.ndim == 2 ? dj.n1 : 1);
        if (o.in == it.in && rows_i == rows_j && dj.k1 == di.k1 && workspace_layout(pj, dj, o.batch).ld1 == w.ld1)
          shared = &col[(size_t)j];
      }
    }
    const int st = run_direction(pl, it.direction, it.in, it.out, it.batch, it.alpha_re, it.alpha_im,
                                 static_cast<char*>(it.workspace), s, nullptr, it.output, -1, &col[(size_t)i], shared);
    if (st != kOk) return st;
  }
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  HCIPY_HIP(hipStreamIsCapturing(s, &capture));
  char* src = group->host;
  if (capture != hipStreamCaptureStatusNone) {
    HCIPY_REQUIRE(group->snap_used < group->snap_count,
                  "all %d capture snapshots of this group are taken; call hcipy_mft_group_release_captures once "
                  "the graphs that recorded this group are no longer replayed, or use another group",
                  group->snap_count);
    src = group->snaps + (size_t)group->snap_used * group->bytes;
    ++group->snap_used;
  } else if (group->pending && hipEventQuery(group->copied) == hipErrorNotReady) {
    // the mirror may still be read by the previous call's copy: wait before rewriting it
    HCIPY_HIP(hipEventSynchronize(group->copied));
  }
  // Table: stage-1 items ordered by data mode (one launch per mode), then stage-2 items.
  StageArgs* hargs = reinterpret_cast<StageArgs*>(src);
  int64_t* hstart = reinterpret_cast<int64_t*>(src + round_up((int64_t)(2 * (size_t)group->cap * sizeof(StageArgs)), 64));
  std::vector<GroupedLaunch> launches;
  if (parity) build_parity_launches(prec, col, hargs, hstart, launches);
  else build_grouped_launches(prec, col, hargs, hstart, launches);
  HCIPY_HIP(hipMemcpyAsync(group->dev, src, group->bytes, hipMemcpyHostToDevice, s));
  if (capture == hipStreamCaptureStatusNone) {
    HCIPY_HIP(hipEventRecord(group->copied, s));
    group->pending = true;
  }
  const StageArgs* dargs = reinterpret_cast<const StageArgs*>(group->dev);
  const int64_t* dstart = reinterpret_cast<const int64_t*>(group->dev + round_up((int64_t)(2 * (size_t)group->cap * sizeof(StageArgs)), 64));
  for (const GroupedLaunch& L : launches) {
    const int st = L.real ? launch_stage_grouped_real(prec, L.cfg, dargs + L.first, dstart + L.start_off, nullptr, L.n,
                                                      L.tiles, s, L.nyq_k)
                          : launch_stage_grouped(prec, L.cfg, L.b_mode, dargs + L.first, dstart + L.start_off, nullptr,
                                                 L.n, L.tiles, s);
    if (st != kOk) return st;
  }
  return kOk;
}

int hcipy_mft_autotune(hcipy_mft_plan plan, int direction_, const void* in, void* out, int64_t batch, int output,
                       void* workspace, size_t workspace_bytes, int repeats, int* chosen, void* stream) {
  HCIPY_REQUIRE(plan != nullptr && repeats > 0, "bad argument");
  HCIPY_REQUIRE(direction_ == HCIPY_FORWARD || direction_ == HCIPY_BACKWARD, "bad direction %d", direction_);
  Plan& pl = plan->p;
  const int nstage = pl.ndim == 2 ? 2 : 1;
  for (int stage = 0; stage < nstage; ++stage) {
    int best = kTileAuto;
    float best_ms = 0.0f;
    for (int cfg = 0; cfg < kNumTileCfg; ++cfg) {
      if (!(pl.parity ? real_cfg_supported(pl.prec, cfg, stage) : tile_cfg_supported(pl.prec, cfg))) continue;
      pl.tiles[direction_][stage] = cfg;
      float ms[3] = {0.0f, 0.0f, 0.0f};
      const int st = hcipy_mft_profile(plan, direction_, in, out, batch, 1.0, 0.0, output, workspace, workspace_bytes,
                                       repeats, ms, stream);
      if (st != kOk) {
        pl.tiles[direction_][stage] = kTileAuto;
        return st;
      }
      if (best == kTileAuto || ms[1 + stage] < best_ms) {
        best = cfg;
        best_ms = ms[1 + stage];
      }
    }
    pl.tiles[direction_][stage] = best;
    if (chosen != nullptr) chosen[stage] = best;
  }
  if (chosen != nullptr && nstage == 1) chosen[1] = kTileAuto;
  return kOk;
}

int hcipy_poke_set_create(hcipy_poke_set* set, const hcipy_poke_item* items, int n_items, void* stream) {
  HCIPY_REQUIRE(set != nullptr && items != nullptr && n_items > 0, "bad argument");
  *set = nullptr;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (int st = refuse_under_capture(s, "hcipy_poke_set_create")) return st;
  int cur = 0;
  HCIPY_HIP(hipGetDevice(&cur));
  const int prec = items[0].plan != nullptr ? items[0].plan->p.prec : -1;
  // geometry of every item: window size and per-actuator origins
  struct Geo {
    int64_t wx, wy, n_act2, nu, nv;
    std::vector<int> org;
  };
  std::vector<Geo> geo((size_t)n_items);
  int64_t total_pokes = 0;
  for (int k = 0; k < n_items; ++k) {
    const hcipy_poke_item& it = items[k];
    HCIPY_REQUIRE(it.plan != nullptr && it.field != nullptr && it.poke_out != nullptr, "item %d: NULL pointer", k);
    HCIPY_REQUIRE(it.output == HCIPY_POKE_INTENSITY || it.output == HCIPY_POKE_FIELD_CHANGE, "item %d: output %d", k,
                  it.output);
    const Plan& pl = it.plan->p;
    HCIPY_REQUIRE(pl.prec == prec, "item %d: all items of a poke set share one precision", k);
    HCIPY_REQUIRE(pl.ndim == 2 && pl.device == cur, "item %d: needs a 2-D plan on the current device", k);
    if (int st = check_gaussian_desc(&it.dm)) return st;
    HCIPY_REQUIRE(it.dm.nx == pl.nx && it.dm.ny == pl.ny, "item %d: influence functions on a %lld x %lld grid, plan "
                  "input %lld x %lld", k, (long long)it.dm.nx, (long long)it.dm.ny, (long long)pl.nx, (long long)pl.ny);
    Geo& g = geo[(size_t)k];
    g.n_act2 = it.dm.n_act * it.dm.n_act;
    g.nu = pl.nu;
    g.nv = pl.nv;
    int64_t need_x = 1, need_y = 1;
    std::vector<int64_t> box((size_t)g.n_act2 * 4);
    for (int64_t a = 0; a < g.n_act2; ++a) {
      int64_t* b = &box[(size_t)a * 4];
      gaussian_support_window(it.dm, a, &b[0], &b[1], &b[2], &b[3]);
      need_x = std::max(need_x, b[1] - b[0] + 1);
      need_y = std::max(need_y, b[3] - b[2] + 1);
    }
    g.wx = std::min(round_up(need_x + 8, 64), round_up(pl.nx, 64));
    g.wy = std::min(round_up(need_y + 8, 64), round_up(pl.ny, 64));
    HCIPY_REQUIRE(g.wx <= kPokeWindowMax && g.wy <= kPokeWindowMax,
                  "item %d: influence-function windows of %lld x %lld pixels exceed %d", k, (long long)g.wx,
                  (long long)g.wy, kPokeWindowMax);
    g.org.resize((size_t)g.n_act2 * 2);
    for (int64_t a = 0; a < g.n_act2; ++a) {
      const int64_t* b = &box[(size_t)a * 4];
      g.org[(size_t)a * 2] = (int)window_origin(b[0], b[1], pl.nx, g.wx);
      g.org[(size_t)a * 2 + 1] = (int)window_origin(b[2], b[3], pl.ny, g.wy);
    }
    total_pokes += g.n_act2;
  }
  HCIPY_REQUIRE(total_pokes <= (int64_t)1 << 24, "too many pokes (%lld)", (long long)total_pokes);

  // memory layout
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (size_t)round_up((int64_t)bytes, 256);
    return at;
  };
  std::vector<size_t> o_base_ws((size_t)n_items), o_base((size_t)n_items), o_org((size_t)n_items), o_poke((size_t)n_items);
  std::vector<PokeWs> pws((size_t)n_items);
  for (int k = 0; k < n_items; ++k) {
    const Plan& pl = items[k].plan->p;
    const Geo& g = geo[(size_t)k];
    o_base_ws[(size_t)k] = take(std::max<size_t>(workspace_layout(pl, direction(pl, HCIPY_FORWARD), 1).total, 256));
    o_base[(size_t)k] = take((size_t)g.nu * g.nv * 16);
    o_org[(size_t)k] = take(g.org.size() * sizeof(int));
    pws[(size_t)k] = poke_ws(prec, g.n_act2, g.wx, g.wy, g.nu);
    o_poke[(size_t)k] = take(pws[(size_t)k].total);
  }
  const size_t o_win = take((size_t)poke_window_bytes(n_items));
  const size_t o_btab = take(2 * (size_t)n_items * sizeof(StageArgs));
  const size_t o_bst = take((2 * (size_t)n_items + 4) * sizeof(int64_t));
  const size_t o_ptab = take(2 * (size_t)total_pokes * sizeof(StageArgs));
  const size_t o_pst = take((2 * (size_t)total_pokes + 4) * sizeof(int64_t));

  hcipy_poke_set_s* h = new (std::nothrow) hcipy_poke_set_s();
  HCIPY_REQUIRE(h != nullptr, "out of host memory");
  h->prec = prec;
  h->device = cur;
  h->n_items = n_items;
  h->mem_bytes = off;
  auto fail = [&](int st) {
    hcipy_poke_set_destroy(h);
    return st;
  };
  if (int st = check_hip(hipMalloc(&h->mem, off), "hipMalloc(poke set)")) return fail(st);
  char* mem = h->mem;

  std::vector<Collected> bcol((size_t)n_items), pcol((size_t)total_pokes);
  std::vector<PokeWinItem> win((size_t)n_items);
  const size_t es = elem_size(prec);
  const int out_mode = prec == kF64 ? kOutI64Base : kOutI32Base;
  int64_t row0 = 0, pk = 0;
  for (int k = 0; k < n_items; ++k) {
    const hcipy_poke_item& it = items[k];
    const Plan& pl = it.plan->p;
    const Geo& g = geo[(size_t)k];
    hcipy_poke_set_s::Item bi{&pl, it.field, mem + o_base_ws[(size_t)k], reinterpret_cast<double2*>(mem + o_base[(size_t)k]),
                              it.base_out, g.nu * g.nv};
    h->items.push_back(bi);
    // base transform: stage arguments only (its split is launched by each execution)
    int st = run_direction(pl, HCIPY_FORWARD, it.field, bi.base, 1, it.alpha_re, it.alpha_im, bi.ws, s, nullptr,
                           HCIPY_OUTPUT_FIELD_C128, 1, &bcol[(size_t)k], nullptr);
    if (st) return fail(st);
    int* org_dev = reinterpret_cast<int*>(mem + o_org[(size_t)k]);
    if ((st = check_hip(hipMemcpy(org_dev, g.org.data(), g.org.size() * sizeof(int), hipMemcpyHostToDevice),
                        "copy poke window origins")))
      return fail(st);
    const PokeWs& w = pws[(size_t)k];
    char* p0 = mem + o_poke[(size_t)k];
    PokeWinItem& wi = win[(size_t)k];
    wi.dm = it.dm;
    wi.field = it.field;
    wi.phase_scale = it.phase_scale;
    wi.prec = prec;
    wi.wx = (int)g.wx;
    wi.wy = (int)g.wy;
    wi.row0 = row0;
    wi.org = org_dev;
    wi.row_scale = is_split(prec) ? reinterpret_cast<float*>(p0 + w.row_scale) : nullptr;
    // the window kernel writes actuator a's rows at out[q] + a * wy * 2 wx elements
    for (int q = 0; q < 3; ++q) wi.out[q] = q < n_arrays(prec) ? p0 + w.win[q] : nullptr;
    row0 += g.n_act2 * g.wy;
    const FMatrix* m1 = &pl.f1;
    const FMatrix* m2 = &pl.f2;
    for (int64_t a = 0; a < g.n_act2; ++a, ++pk) {
      char* win_a[3], *z_a[3];
      for (int q = 0; q < 3; ++q) {
        win_a[q] = q < n_arrays(prec) ? p0 + w.win[q] + (size_t)a * w.win_pitch : nullptr;
        z_a[q] = q < n_arrays(prec) ? p0 + w.z[q] + (size_t)a * w.z_pitch : nullptr;
      }
      const float* rs_a = is_split(prec) ? reinterpret_cast<const float*>(p0 + w.row_scale) + a * g.wy : nullptr;
      float* t_a = is_split(prec) ? reinterpret_cast<float*>(p0 + w.t) + a : nullptr;
      const int64_t x0 = g.org[(size_t)a * 2], y0 = g.org[(size_t)a * 2 + 1];
      Collected& c = pcol[(size_t)pk];
      StageArgs& a1 = c.a1;
      a1.a = static_cast<const char*>(m1->hi) + 2 * x0 * es;
      a1.a_lo = m1->lo ? static_cast<const char*>(m1->lo) + 2 * x0 * es : nullptr;
      a1.a_lo2 = m1->lo2 ? static_cast<const char*>(m1->lo2) + 2 * x0 * es : nullptr;
      a1.lda = m1->lda;
      a1.b = win_a[0];
      a1.b_lo = win_a[1];
      a1.b_lo2 = win_a[2];
      a1.ldb = 2 * g.wx;
      a1.M = (int)(2 * g.nu);
      a1.N = (int)g.wy;
      a1.K = (int)(2 * g.wx);
      a1.t_log2k = ilog2_ceil(a1.K);
      a1.out = z_a[0];
      a1.out_lo = z_a[1];
      a1.out_lo2 = z_a[2];
      a1.ldo = g.wy;
      a1.fold = (int)g.wy;
      a1.fold_stride = g.nu * g.wy;
      a1.out_mode = is_split(prec) ? kOutSplit : (prec == kF64 ? kOutC128 : kOutC64);
      a1.row_scale = rs_a;
      a1.t_out = t_a;
      a1.alpha_re = 1.0;
      a1.alpha_im = 0.0;
      a1.tile_cfg = kTileAuto;
      StageArgs& a2 = c.a2;
      a2.a = static_cast<const char*>(m2->hi) + 2 * y0 * es;
      a2.a_lo = m2->lo ? static_cast<const char*>(m2->lo) + 2 * y0 * es : nullptr;
      a2.a_lo2 = m2->lo2 ? static_cast<const char*>(m2->lo2) + 2 * y0 * es : nullptr;
      a2.lda = m2->lda;
      a2.b = z_a[0];
      a2.b_lo = a1.out_lo;
      a2.b_lo2 = a1.out_lo2;
      a2.ldb = 2 * g.wy;
      a2.b_zstride = g.nu * 2 * g.wy;
      a2.M = (int)(2 * g.nv);
      a2.N = (int)g.nu;
      a2.K = (int)(2 * g.wy);
      const bool field_change = it.output == HCIPY_POKE_FIELD_CHANGE;
      const size_t out_es = field_change ? (prec == kF64 ? 16 : 8) : (prec == kF64 ? 8 : 4);
      a2.out = static_cast<char*>(it.poke_out) + (size_t)a * g.nu * g.nv * out_es;
      a2.ldo = g.nu;
      a2.out_zstride = g.nu * g.nv;
      a2.out_mode = field_change ? (prec == kF64 ? kOutC128 : kOutC64) : out_mode;
      a2.base = field_change ? nullptr : bi.base;
      a2.z_scale = t_a;
      a2.alpha_re = it.alpha_re;
      a2.alpha_im = it.alpha_im;
      a2.tile_cfg = kTileAuto;
      c.b_mode1 = kModePadded;
      c.nz2 = 1;
    }
  }
  h->total_rows = row0;
  h->n_pokes = total_pokes;
  for (const Geo& g : geo) h->max_wx = std::max(h->max_wx, (int)g.wx);
  for (int k = 0; k < n_items; ++k) {
    const Plan& pl = items[k].plan->p;
    const Geo& g = geo[(size_t)k];
    const double full = 8.0 * ((double)pl.ny * pl.nx * pl.nu + (double)pl.nv * pl.ny * pl.nu);
    const double win = 8.0 * ((double)g.wy * g.wx * g.nu + (double)g.nv * g.wy * g.nu);
    h->flops_full += full * (double)(g.n_act2 + 1);
    h->flops_windowed += full + win * (double)g.n_act2;
  }
  std::vector<char> host((size_t)poke_window_bytes(n_items));
  h->win_items = mem + o_win;
  int st = upload_poke_items(win.data(), n_items, host.data(), h->win_items, s);
  if (!st) st = check_hip(hipStreamSynchronize(s), "poke set descriptors");
  if (st) return fail(st);
  // tables
  std::vector<StageArgs> tab(2 * (size_t)std::max<int64_t>(total_pokes, n_items));
  std::vector<int64_t> start(2 * (size_t)std::max<int64_t>(total_pokes, n_items) + 4);
  build_grouped_launches(prec, bcol, tab.data(), start.data(), h->base_launches);
  if ((st = check_hip(hipMemcpy(mem + o_btab, tab.data(), 2 * (size_t)n_items * sizeof(StageArgs), hipMemcpyHostToDevice),
                      "copy base table")))
    return fail(st);
  if ((st = check_hip(hipMemcpy(mem + o_bst, start.data(), (2 * (size_t)n_items + 4) * sizeof(int64_t),
                                hipMemcpyHostToDevice), "copy base tile starts")))
    return fail(st);
  // tile maps sized for the smallest tile shape of any grouped launch (64 x 32)
  for (const Collected& c : pcol)
    h->map_cap += (size_t)(ceil_div(c.a1.M, 64) * ceil_div(c.a1.N, 32) + ceil_div(c.a2.M, 64) * ceil_div(c.a2.N, 32));
  if ((st = check_hip(hipMalloc(&h->map_mem, std::max<size_t>(h->map_cap, 1) * sizeof(int)), "hipMalloc(tile map)")))
    return fail(st);
  h->pcol = std::move(pcol);
  h->ptab_dev = reinterpret_cast<StageArgs*>(mem + o_ptab);
  h->pst_dev = reinterpret_cast<int64_t*>(mem + o_pst);
  if ((st = upload_poke_tables(h, h->cfg))) return fail(st);
  h->base_tab = reinterpret_cast<const StageArgs*>(mem + o_btab);
  h->base_start = reinterpret_cast<const int64_t*>(mem + o_bst);
  *set = h;
  return kOk;
}

int hcipy_poke_set_bytes(hcipy_poke_set set, size_t* bytes) {
  HCIPY_REQUIRE(set != nullptr && bytes != nullptr, "NULL argument");
  *bytes = set->mem_bytes;
  return kOk;
}

int hcipy_poke_set_info(hcipy_poke_set set, int64_t* n_pokes, double* flops_windowed, double* flops_full) {
  HCIPY_REQUIRE(set != nullptr, "NULL poke set");
  if (n_pokes) *n_pokes = set->n_pokes;
  if (flops_windowed) *flops_windowed = set->flops_windowed;
  if (flops_full) *flops_full = set->flops_full;
  return kOk;
}

int hcipy_poke_set_execute(hcipy_poke_set set, void* stream) {
  HCIPY_REQUIRE(set != nullptr, "NULL poke set");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int cur = 0;
  HCIPY_HIP(hipGetDevice(&cur));
  HCIPY_REQUIRE(cur == set->device, "the poke set was created on device %d, current device is %d", set->device, cur);
  const int prec = set->prec;
  if (is_split(prec)) {
    for (const auto& it : set->items) {
      const Plan& pl = *it.pl;
      const Ws w = workspace_layout(pl, direction(pl, HCIPY_FORWARD), 1);
      SplitOut o{{it.ws + w.split_hi, it.ws + w.split_lo, prec == kBF16x3 ? it.ws + w.split_lo2 : nullptr}};
      if (int st = launch_split_rows(prec, static_cast<const float*>(it.field), 2 * pl.nx, (int)(2 * pl.nx), pl.ny, o,
                                     w.ld1, reinterpret_cast<float*>(it.ws + w.row_scale), s))
        return st;
    }
  }
  if (int st = launch_poke_windows(set->win_items, set->n_items, set->total_rows, set->max_wx, s)) return st;
  for (const GroupedLaunch& L : set->base_launches)
    if (int st = launch_stage_grouped(prec, L.cfg, L.b_mode, set->base_tab + L.first, set->base_start + L.start_off,
                                      nullptr, L.n, L.tiles, s))
      return st;
  for (const GroupedLaunch& L : set->poke_launches)
    if (int st = launch_stage_grouped(prec, L.cfg, L.b_mode, set->poke_tab + L.first, set->poke_start + L.start_off,
                                      set->poke_maps[&L - set->poke_launches.data()], L.n, L.tiles, s))
      return st;
  for (const auto& it : set->items) {
    if (it.base_out == nullptr) continue;
    hipLaunchKernelGGL(c128_intensity_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(it.n_focal, 256), 4096)),
                       dim3(256), 0, s, it.base, it.base_out, prec == kF64 ? 1 : 0, it.n_focal);
    HCIPY_HIP(hipGetLastError());
  }
  return kOk;
}

int hcipy_poke_set_autotune(hcipy_poke_set set, int repeats, int* chosen, void* stream) {
  HCIPY_REQUIRE(set != nullptr && repeats > 0, "bad argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (int st = refuse_under_capture(s, "hcipy_poke_set_autotune")) return st;
  hipEvent_t ev[2];
  HCIPY_HIP(hipEventCreate(&ev[0]));
  if (int st = check_hip(hipEventCreate(&ev[1]), "hipEventCreate")) {
    (void)hipEventDestroy(ev[0]);
    return st;
  }
  int st = kOk;
  for (int stage = 0; stage < 2 && !st; ++stage) {
    int best = -1;
    float best_ms = 0.0f;
    for (int cfg = 0; cfg < kNumTileCfg && !st; ++cfg) {
      if (!grouped_cfg_supported(set->prec, cfg)) continue;
      int trial[2] = {set->cfg[0], set->cfg[1]};
      trial[stage] = cfg;
      if ((st = upload_poke_tables(set, trial))) break;
      if ((st = hcipy_poke_set_execute(set, stream))) break;  // warm-up
      if ((st = check_hip(hipEventRecord(ev[0], s), "hipEventRecord"))) break;
      for (int r = 0; r < repeats && !st; ++r) st = hcipy_poke_set_execute(set, stream);
      if (!st) st = check_hip(hipEventRecord(ev[1], s), "hipEventRecord");
      if (!st) st = check_hip(hipEventSynchronize(ev[1]), "hipEventSynchronize");
      float ms = 0.0f;
      if (!st) st = check_hip(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime");
      if (!st && (best < 0 || ms < best_ms)) {
        best = cfg;
        best_ms = ms;
      }
    }
    if (!st) set->cfg[stage] = best;
  }
  if (!st) st = upload_poke_tables(set, set->cfg);
  if (!st && chosen != nullptr) {
    chosen[0] = set->cfg[0];
    chosen[1] = set->cfg[1];
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  return st;
}

int hcipy_poke_set_set_tiles(hcipy_poke_set set, int stage1_tiles, int stage2_tiles) {
  HCIPY_REQUIRE(set != nullptr, "NULL poke set");
  const int cfg[2] = {stage1_tiles, stage2_tiles};
  for (int stage = 0; stage < 2; ++stage)
    HCIPY_REQUIRE(cfg[stage] == kTileAuto || grouped_cfg_supported(set->prec, cfg[stage]),
                  "tile configuration %d is not available to poke sets of this precision", cfg[stage]);
  if (int st = upload_poke_tables(set, cfg)) return st;
  set->cfg[0] = cfg[0];
  set->cfg[1] = cfg[1];
  return kOk;
}

int hcipy_poke_set_destroy(hcipy_poke_set set) {
  if (set == nullptr) return kOk;
  if (set->mem) (void)hipFree(set->mem);
  if (set->map_mem) (void)hipFree(set->map_mem);
  delete set;
  return kOk;
}

}  // extern "C"
