// C-ABI of the library: error state and the matrix-Fourier-transform plan.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/hcipy_amd.h"
#include "mft_gemm.hpp"

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
int launch_stage_grouped(int prec, int cfg, int b_mode, const StageArgs* table, const int64_t* tile_start, int n,
                         int64_t tiles, hipStream_t s);
int grouped_tile_cfg(int prec, int stage, int max_k, int64_t tiles64);
void grouped_tile_shape(int cfg, int* bm, int* bn);
constexpr int kModePadded = 0, kModeAligned = 1, kModeUnaligned = 2;  // BMode of mft_gemm.hip
int launch_split_rows(int prec, const float* in, int64_t ld_in, int K, int64_t rows, SplitOut o, int64_t ld_out,
                      float* row_scale, hipStream_t s);
int launch_gen_matrix(const double* a, int nc, const double* b, int nk, double sign, int prec, SplitOut out,
                      int64_t lda, int rows_pad, hipStream_t s);

namespace {

constexpr int64_t kAlign = 256;
// Padded operand rows (constant matrices, split data, intermediate) hold a multiple of this many
// elements, the largest k-tile of the stage kernels, so full-vector loads of the last k-tile
// stay inside the row.
constexpr int64_t kRowPad = 128;
int64_t padded_row(int64_t k) { return round_up(k, kRowPad); }
constexpr int kMaxGridZ = 65535;

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

struct Plan {
  int ndim = 0, prec = 0, device = 0;
  int64_t nx = 1, ny = 1, nu = 1, nv = 1;
  FMatrix f1, f2, b1, b2;  // forward stage 1/2, backward stage 1/2
  double* coords = nullptr;
  size_t matrix_bytes = 0;
  int tiles[2][2] = {{kTileAuto, kTileAuto}, {kTileAuto, kTileAuto}};  // [direction][stage]
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
};

// Runs one direction: data split (split-FP16 only), stage 1, stage 2.  `only` >= 0 launches just
// that kernel (0 split, 1 stage 1, 2 stage 2) on a workspace a full run has filled (timing).
// With `col`, the split is launched and the stage arguments are returned instead of launched.
// With `shared_split` (split-FP16, `col` set), the split is not launched: stage 1 reads the split
// data and row scales that `shared_split->a1` points at (an earlier item of the same input).
int run_direction(const Plan& pl, int dir, const void* in, void* out, int64_t batch, double ar, double ai,
                  char* ws, hipStream_t s, hipEvent_t* ev = nullptr, int output = HCIPY_OUTPUT_FIELD, int only = -1,
                  Collected* col = nullptr, const Collected* shared_split = nullptr) {
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
  HCIPY_REQUIRE(tile_cfg == HCIPY_TILES_AUTO || tile_cfg_supported(plan->p.prec, tile_cfg),
                "tile configuration %d is not available for precision %d", tile_cfg, plan->p.prec);
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
  const Ws w = workspace_layout(pl, direction(pl, direction_), batch);
  HCIPY_REQUIRE(workspace_bytes >= w.total, "workspace too small: %zu < %zu bytes", workspace_bytes, w.total);
  HCIPY_REQUIRE(w.total == 0 || workspace != nullptr, "workspace is NULL");
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
  const Ws w = workspace_layout(pl, direction(pl, direction_), batch);
  if (w.total > (size_t)HCIPY_NOALLOC_SCRATCH_BYTES) {
    set_last_error("hcipy_mft_execute_noalloc: %lld fields need %zu bytes of workspace, above the %d-byte static "
                   "scratch; use hcipy_mft_execute_ex with a caller workspace or a smaller batch",
                   (long long)batch, w.total, (int)HCIPY_NOALLOC_SCRATCH_BYTES);
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
  const Ws w = workspace_layout(pl, direction(pl, direction_), batch);
  HCIPY_REQUIRE(workspace_bytes >= w.total, "workspace too small: %zu < %zu bytes", workspace_bytes, w.total);
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
  for (int i = 0; i < n_items; ++i) {
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
  struct Launch {
    int stage, b_mode, cfg, first, n;
    int64_t start_off, tiles;
  };
  std::vector<Launch> launches;
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
      const int cfg = grouped_tile_cfg(prec, stage, max_k, tiles64);
      int bm, bn;
      grouped_tile_shape(cfg, &bm, &bn);
      Launch L{stage, mode, cfg, na, 0, ns, 0};
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
  HCIPY_HIP(hipMemcpyAsync(group->dev, src, group->bytes, hipMemcpyHostToDevice, s));
  if (capture == hipStreamCaptureStatusNone) {
    HCIPY_HIP(hipEventRecord(group->copied, s));
    group->pending = true;
  }
  const StageArgs* dargs = reinterpret_cast<const StageArgs*>(group->dev);
  const int64_t* dstart = reinterpret_cast<const int64_t*>(group->dev + round_up((int64_t)(2 * (size_t)group->cap * sizeof(StageArgs)), 64));
  for (const Launch& L : launches) {
    const int st = launch_stage_grouped(prec, L.cfg, L.b_mode, dargs + L.first, dstart + L.start_off, L.n, L.tiles, s);
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
      if (!tile_cfg_supported(pl.prec, cfg)) continue;
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

}  // extern "C"
