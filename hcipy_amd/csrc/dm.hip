// Deformable mirror with Gaussian or tabulated (Xinetics) influence functions, evaluated without
// the influence matrix.
//
// The reference builds every influence function as a sparse pupil-grid vector and forms the
// surface as a CSR matrix-vector product (hcipy/optics/deformable_mirror.py:43-93,
// hcipy/mode_basis/gaussian_pokes.py:5-54, hcipy/mode_basis/mode_basis.py:239-258).  At a 1024^2
// pupil with 32 x 32 actuators that matrix holds ~3e7 non-zeros.  Here each pupil pixel sums the
// few actuators whose truncated Gaussian reaches it, in FP64, and the adjoint (gradient with
// respect to the actuators) reduces, per actuator, the surface gradient over the pixel window
// its support can reach.  Both read only the actuators / the gradient and write one array, so
// they are bound by the HBM write of the surface (8 bytes per pixel) or by FP64 exp throughput.
//
// Influence function of actuator a at pupil point p (same steps as the reference):
//   q = R(-z_tilt) . diag(1/cos(y_tilt), 1/cos(x_tilt)) . (p + dither)
//   r2 = |q - m_a|^2
//   g = exp(-r2 / (2 sigma^2)) - exp(-radius^2 / (2 sigma^2))   if r2 <= radius^2, else 0
//   IF_a(p) = mean over the oversampling dithers of g, divided by cos(x_tilt) cos(y_tilt)
// with dithers ((k + 0.5) / ovs - 0.5) * delta per axis (evaluate_supersampled,
// hcipy/field/util.py:606-698).
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "pokes.hpp"
#include "../../include/hcipy_amd.h"

namespace hcipy {

namespace {

struct DmGeom {
  int64_t nx, ny, n_act;
  double x0, y0, dx, dy;
  double ax0, ay0, pitch;
  double sigma, radius;
  double inv_two_sigma2, offset, r2max;
  double sx, sy;    // 1/cos(y_tilt), 1/cos(x_tilt)
  double cz, sz;    // cos, sin of z_tilt
  double norm;      // 1 / (ovs^2 cos(x_tilt) cos(y_tilt))
  int ovs;
};

DmGeom make_geom(const hcipy_dm_gaussian_desc& d) {
  DmGeom g;
  g.nx = d.nx;
  g.ny = d.ny;
  g.n_act = d.n_act;
  g.x0 = d.x0;
  g.y0 = d.y0;
  g.dx = d.dx;
  g.dy = d.dy;
  g.ax0 = d.act_x0;
  g.ay0 = d.act_y0;
  g.pitch = d.pitch;
  g.sigma = d.sigma;
  g.radius = d.radius;
  g.inv_two_sigma2 = 0.5 / (d.sigma * d.sigma);
  g.offset = std::exp(-0.5 * (d.radius / d.sigma) * (d.radius / d.sigma));
  g.r2max = d.radius * d.radius;
  g.sx = 1.0 / std::cos(d.y_tilt);
  g.sy = 1.0 / std::cos(d.x_tilt);
  g.cz = std::cos(d.z_tilt);
  g.sz = std::sin(d.z_tilt);
  g.ovs = (int)d.oversampling;
  g.norm = 1.0 / ((double)g.ovs * g.ovs * std::cos(d.x_tilt) * std::cos(d.y_tilt));
  return g;
}

// Transformed evaluation point of pupil pixel (i, j) and dither (u, v).
__device__ inline void eval_point(const DmGeom& g, int64_t i, int64_t j, int u, int v, double& qx, double& qy) {
  const double fx = ((u + 0.5) / g.ovs - 0.5) * g.dx;
  const double fy = ((v + 0.5) / g.ovs - 0.5) * g.dy;
  const double px = (g.x0 + (double)i * g.dx + fx) * g.sx;
  const double py = (g.y0 + (double)j * g.dy + fy) * g.sy;
  // rotated(-z_tilt): [[cos z, sin z], [-sin z, cos z]] applied to the point
  qx = g.cz * px + g.sz * py;
  qy = -g.sz * px + g.cz * py;
}

__device__ inline double poke(const DmGeom& g, double qx, double qy, double mx, double my) {
  const double ddx = qx - mx, ddy = qy - my;
  const double r2 = ddx * ddx + ddy * ddy;
  return (r2 <= g.r2max) ? (exp(-r2 * g.inv_two_sigma2) - g.offset) : 0.0;
}

// Rows per workgroup of the surface kernels: a workgroup is 256 pixel columns (one per thread)
// by kSurfRows rows, launched on a 2-D grid so that no thread divides a linear pixel index.  Four
// rows give a 1024^2 pupil four resident waves per SIMD (eight rows: two, too few to hide the LDS
// and FP64 latencies of the separable kernel).
constexpr int kSurfRows = 4;

// surface[b][pixel] = sum_a w[b][a] IF_a(pixel)
__global__ __launch_bounds__(256) void dm_surface_kernel(DmGeom g, const double* __restrict__ w,
                                                         double* __restrict__ surface, int64_t b_off) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.nx) return;
  const int64_t npix = g.nx * g.ny;
  const int64_t b = blockIdx.z + b_off;
  const double* wb = w + b * g.n_act * g.n_act;
  const int64_t jb = (int64_t)blockIdx.y * kSurfRows;
  const int rows = (int)min((int64_t)kSurfRows, g.ny - jb);
  for (int r = 0; r < rows; ++r) {
    const int64_t j = jb + r;
    double acc = 0.0;
    for (int v = 0; v < g.ovs; ++v)
      for (int u = 0; u < g.ovs; ++u) {
        double qx, qy;
        eval_point(g, i, j, u, v, qx, qy);
        // actuators within radius of q on the (untilted) actuator grid
        const int64_t ia0 = max((int64_t)0, (int64_t)ceil((qx - g.radius - g.ax0) / g.pitch));
        const int64_t ia1 = min(g.n_act - 1, (int64_t)floor((qx + g.radius - g.ax0) / g.pitch));
        const int64_t ja0 = max((int64_t)0, (int64_t)ceil((qy - g.radius - g.ay0) / g.pitch));
        const int64_t ja1 = min(g.n_act - 1, (int64_t)floor((qy + g.radius - g.ay0) / g.pitch));
        for (int64_t ja = ja0; ja <= ja1; ++ja) {
          const double my = g.ay0 + (double)ja * g.pitch;
          for (int64_t ia = ia0; ia <= ia1; ++ia) {
            const double wa = wb[ja * g.n_act + ia];
            if (wa != 0.0) acc += wa * poke(g, qx, qy, g.ax0 + (double)ia * g.pitch, my);
          }
        }
      }
    surface[b * npix + j * g.nx + i] = acc * g.norm;
  }
}

// Without a z tilt the transformed point is (px sx, py sy) and the Gaussian factorises,
//   exp(-(ddx^2 + ddy^2) / (2 sigma^2)) = exp(-ddx^2 / (2 sigma^2)) exp(-ddy^2 / (2 sigma^2)),
// so the factors of a pixel column (per x dither and actuator column in reach) and of a pixel row
// (per y dither and actuator row in reach) are computed once and every (pixel, actuator) pair needs
// multiply-adds only; the support test compares the same ddx^2 + ddy^2 as the general kernels.  At
// most kSepMax actuators per axis may lie within reach of a point (a support radius of up to 3
// pitches, the reference's default cutoff; else the general kernels run).
constexpr int kSepMax = 8;
constexpr int kMaxOvs = 16;  // the descriptor's oversampling limit

bool separable(const DmGeom& g) { return g.sz == 0.0 && g.cz == 1.0 && 2.0 * g.radius / g.pitch + 2.0 <= kSepMax; }

// Actuator weights of a surface workgroup's reach staged in LDS (when they fit; else they are read
// from global memory): a wave's 64 columns share a few actuator columns, which LDS serves as
// broadcasts where vector loads would issue 64 addresses each.
constexpr int kSurfWLds = 1024;

// Thread = pixel column i of the workgroup's kSurfRows rows.  The row factors of those rows are
// tabulated in LDS by the workgroup; the column factors stay in registers over the rows.
__global__ __launch_bounds__(256) void dm_surface_sep_kernel(DmGeom g, const double* __restrict__ w,
                                                             double* __restrict__ surface, int64_t b_off) {
  __shared__ double s_ey[kSurfRows * kMaxOvs * kSepMax];
  __shared__ double s_dy2[kSurfRows * kMaxOvs * kSepMax];
  __shared__ int s_ja0[kSurfRows * kMaxOvs];
  __shared__ int s_nja[kSurfRows * kMaxOvs];
  __shared__ double s_w[kSurfWLds];
  const int ovs = g.ovs;
  const int64_t b = blockIdx.z + b_off;
  const int64_t jb = (int64_t)blockIdx.y * kSurfRows;
  const int rows = (int)min((int64_t)kSurfRows, g.ny - jb);
  // entry (r, v, k): pixel row jb + r, y dither v, k-th actuator row in reach
  for (int t = threadIdx.x; t < rows * ovs * kSepMax; t += 256) {
    const int k = t % kSepMax, rv = t / kSepMax;
    const int r = rv / ovs, v = rv - r * ovs;
    const double fy = ((v + 0.5) / ovs - 0.5) * g.dy;
    const double qy = (g.y0 + (double)(jb + r) * g.dy + fy) * g.sy;
    const int64_t ja0 = max((int64_t)0, (int64_t)ceil((qy - g.radius - g.ay0) / g.pitch));
    const int64_t ja1 = min(g.n_act - 1, (int64_t)floor((qy + g.radius - g.ay0) / g.pitch));
    const double d = qy - (g.ay0 + (double)(ja0 + k) * g.pitch);
    s_dy2[t] = d * d;
    s_ey[t] = (ja0 + k <= ja1) ? exp(-d * d * g.inv_two_sigma2) : 0.0;
    if (k == 0) {
      s_ja0[rv] = (int)ja0;
      s_nja[rv] = (int)(ja1 - ja0 + 1);
    }
  }
  __syncthreads();
  const double* wb = w + b * g.n_act * g.n_act;
  // actuators in reach of the workgroup: q grows with the pixel index and the dither, so the
  // first column's first dither and the last column's last dither bound the actuator columns,
  // the first and last table rows the actuator rows
  const int64_t ib = (int64_t)blockIdx.x * 256;
  const int64_t ie = min(ib + 255, g.nx - 1);
  const double qx_lo = (g.x0 + (double)ib * g.dx + (0.5 / ovs - 0.5) * g.dx) * g.sx;
  const double qx_hi = (g.x0 + (double)ie * g.dx + ((ovs - 0.5) / ovs - 0.5) * g.dx) * g.sx;
  const int ia_lo = (int)max((int64_t)0, (int64_t)ceil((qx_lo - g.radius - g.ax0) / g.pitch));
  const int ia_hi = (int)min(g.n_act - 1, (int64_t)floor((qx_hi + g.radius - g.ax0) / g.pitch));
  const int ja_lo = max(s_ja0[0], 0);
  const int ja_hi = min(s_ja0[rows * ovs - 1] + s_nja[rows * ovs - 1] - 1, (int)g.n_act - 1);
  const int ldw = ia_hi - ia_lo + 1;
  const bool w_lds = ldw > 0 && ja_hi >= ja_lo && ldw * (ja_hi - ja_lo + 1) <= kSurfWLds;
  if (w_lds) {
    for (int t = threadIdx.x; t < ldw * (ja_hi - ja_lo + 1); t += 256) {
      const int rr = t / ldw, cc = t - rr * ldw;
      s_w[t] = wb[(int64_t)(ja_lo + rr) * g.n_act + ia_lo + cc];
    }
  }
  __syncthreads();
  const int64_t i = ib + threadIdx.x;
  if (i >= g.nx) return;
  // two partial sums per row (even and odd actuator columns): independent FP64 chains
  double acc[kSurfRows][2];
#pragma unroll
  for (int r = 0; r < kSurfRows; ++r) acc[r][0] = acc[r][1] = 0.0;
  for (int u = 0; u < ovs; ++u) {
    const double fx = ((u + 0.5) / ovs - 0.5) * g.dx;
    const double qx = (g.x0 + (double)i * g.dx + fx) * g.sx;
    const int64_t ia0 = max((int64_t)0, (int64_t)ceil((qx - g.radius - g.ax0) / g.pitch));
    const int64_t ia1 = min(g.n_act - 1, (int64_t)floor((qx + g.radius - g.ax0) / g.pitch));
    const int nia = (int)(ia1 - ia0 + 1);
    double ex[kSepMax], dx2[kSepMax];
#pragma unroll
    for (int c = 0; c < kSepMax; ++c) {
      const double d = qx - (g.ax0 + (double)(ia0 + c) * g.pitch);
      dx2[c] = d * d;
      ex[c] = 0.0;
      if (c < nia) ex[c] = exp(-dx2[c] * g.inv_two_sigma2);
    }
#pragma unroll
    for (int r = 0; r < kSurfRows; ++r) {
      if (r >= rows) break;
      for (int v = 0; v < ovs; ++v) {
        const int rv = r * ovs + v;
        const int ja0 = s_ja0[rv], nja = s_nja[rv];
        for (int k = 0; k < nja; ++k) {
          const double eyk = s_ey[rv * kSepMax + k], dyk = s_dy2[rv * kSepMax + k];
          if (w_lds) {
            const double* wr = s_w + (ja0 + k - ja_lo) * ldw + (int)(ia0 - ia_lo);
#pragma unroll
            for (int c = 0; c < kSepMax; ++c) {
              if (c < nia && dx2[c] + dyk <= g.r2max) acc[r][c & 1] += wr[c] * (ex[c] * eyk - g.offset);
            }
          } else {
            const double* wr = wb + (int64_t)(ja0 + k) * g.n_act + ia0;
#pragma unroll
            for (int c = 0; c < kSepMax; ++c) {
              if (c < nia && dx2[c] + dyk <= g.r2max) acc[r][c & 1] += wr[c] * (ex[c] * eyk - g.offset);
            }
          }
        }
      }
    }
  }
  const int64_t npix = g.nx * g.ny;
#pragma unroll
  for (int r = 0; r < kSurfRows; ++r)
    if (r < rows) surface[b * npix + (jb + r) * g.nx + i] = (acc[r][0] + acc[r][1]) * g.norm;
}

// Actuator of adjoint workgroup h.  Workgroups are dispatched round-robin over the 8 XCDs, each
// with its own L2; XCD x takes the x-th contiguous eighth of the actuators (whole actuator rows
// when n % 8 == 0), so the windows its workgroups read cover a band of the pupil that its L2 can
// hold, instead of the whole gradient.
__device__ inline int64_t adjoint_actuator(int64_t h, int64_t n) {
  const int64_t q = n >> 3, r = n & 7, x = h & 7;
  return x * q + min(x, r) + (h >> 3);
}

// grad_w[b][a] = sum_pixel grad_s[b][pixel] IF_a(pixel): one workgroup per (actuator, batch item),
// a fixed thread -> pixel assignment (wave w takes window rows w, w + 4, ..., lane l the columns
// l, l + 64, ...) and a fixed reduction tree (deterministic).
__global__ __launch_bounds__(256) void dm_adjoint_kernel(DmGeom g, const double* __restrict__ gs,
                                                         double* __restrict__ gw, int64_t b_off) {
  __shared__ double red[4];
  const int64_t a = adjoint_actuator(blockIdx.x, g.n_act * g.n_act);
  const int64_t b = blockIdx.y + b_off;
  const int64_t npix = g.nx * g.ny;
  const double* gsb = gs + b * npix;
  const double mx = g.ax0 + (double)(a % g.n_act) * g.pitch;
  const double my = g.ay0 + (double)(a / g.n_act) * g.pitch;
  // pixel window: the support disc mapped back to the pupil (the inverse transform does not
  // lengthen vectors), widened by one pixel for the dithers
  const double cx = (g.cz * mx - g.sz * my) / g.sx;
  const double cy = (g.sz * mx + g.cz * my) / g.sy;
  const int64_t i0 = max((int64_t)0, (int64_t)floor((cx - g.radius - g.x0) / g.dx) - 1);
  const int64_t i1 = min(g.nx - 1, (int64_t)ceil((cx + g.radius - g.x0) / g.dx) + 1);
  const int64_t j0 = max((int64_t)0, (int64_t)floor((cy - g.radius - g.y0) / g.dy) - 1);
  const int64_t j1 = min(g.ny - 1, (int64_t)ceil((cy + g.radius - g.y0) / g.dy) + 1);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc = 0.0;
  if (i0 <= i1 && j0 <= j1) {
    const int wi = (int)(i1 - i0 + 1), hj = (int)(j1 - j0 + 1);
    for (int jj = wave; jj < hj; jj += 4) {
      const int64_t j = j0 + jj;
      const double* row = gsb + j * g.nx + i0;
      for (int ii = lane; ii < wi; ii += 64) {
        double val = 0.0;
        for (int v = 0; v < g.ovs; ++v)
          for (int u = 0; u < g.ovs; ++u) {
            double qx, qy;
            eval_point(g, i0 + ii, j, u, v, qx, qy);
            val += poke(g, qx, qy, mx, my);
          }
        acc += row[ii] * val;
      }
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) red[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) gw[b * g.n_act * g.n_act + a] = (red[0] + red[1] + red[2] + red[3]) * g.norm;
}

// Separable adjoint (no z tilt): the per-axis factors exp(-ddx^2 / (2 sigma^2)) and squared
// distances of the window's columns and rows, per dither, are tabulated in LDS first (`tab_len`
// doubles of dynamic LDS hold the 2 ovs (wi + hj) of the largest window, at most kSepLds), then
// every window pixel needs multiply-adds only.  The window's pixels, numbered row by row, are
// dealt to the 256 threads in turn (pixel t to thread t mod 256), so every lane has a pixel
// whatever the window's width (64-lane row pieces left 44 % of the lanes idle on 72-pixel
// windows); a thread issues the loads of kAdjBatch of its pixels before it uses any, since the
// loads hit L2 and one at a time the wave waited out their latency in turn (1024^2 pupil, 48 x 48
// actuators: 4 pixels 53.5 us at ovs 1 and 94.0 at ovs 2, 8 pixels 55.3 / 95.4, 16 61.8 / 105).
constexpr int kSepLds = 4096;
constexpr int kAdjBatch = 4;

__global__ __launch_bounds__(256) void dm_adjoint_sep_kernel(DmGeom g, const double* __restrict__ gs,
                                                             double* __restrict__ gw, int64_t b_off, int tab_len) {
  __shared__ double red[4];
  extern __shared__ double tab[];
  const int64_t a = adjoint_actuator(blockIdx.x, g.n_act * g.n_act);
  const int64_t b = blockIdx.y + b_off;
  const int64_t npix = g.nx * g.ny;
  const double* gsb = gs + b * npix;
  const double mx = g.ax0 + (double)(a % g.n_act) * g.pitch;
  const double my = g.ay0 + (double)(a / g.n_act) * g.pitch;
  // the same pixel window as the general adjoint (its centre without the rotation)
  const double cx = mx / g.sx;
  const double cy = my / g.sy;
  const int64_t i0 = max((int64_t)0, (int64_t)floor((cx - g.radius - g.x0) / g.dx) - 1);
  const int64_t i1 = min(g.nx - 1, (int64_t)ceil((cx + g.radius - g.x0) / g.dx) + 1);
  const int64_t j0 = max((int64_t)0, (int64_t)floor((cy - g.radius - g.y0) / g.dy) - 1);
  const int64_t j1 = min(g.ny - 1, (int64_t)ceil((cy + g.radius - g.y0) / g.dy) + 1);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc = 0.0;
  const int wi = (int)(i1 - i0 + 1), hj = (int)(j1 - j0 + 1);
  const int ovs = g.ovs;
  if (wi > 0 && hj > 0 && 2 * ovs * (wi + hj) <= tab_len) {
    // tab: ex[u][wi], dx2[u][wi], ey[v][hj], dy2[v][hj]
    double* ex = tab;
    double* dx2 = ex + ovs * wi;
    double* ey = dx2 + ovs * wi;
    double* dy2 = ey + ovs * hj;
    for (int t = threadIdx.x; t < ovs * wi; t += blockDim.x) {
      const int u = t / wi, ii = t - u * wi;
      const double fx = ((u + 0.5) / ovs - 0.5) * g.dx;
      const double d = (g.x0 + (double)(i0 + ii) * g.dx + fx) * g.sx - mx;
      dx2[t] = d * d;
      ex[t] = exp(-dx2[t] * g.inv_two_sigma2);
    }
    for (int t = threadIdx.x; t < ovs * hj; t += blockDim.x) {
      const int v = t / hj, jj = t - v * hj;
      const double fy = ((v + 0.5) / ovs - 0.5) * g.dy;
      const double d = (g.y0 + (double)(j0 + jj) * g.dy + fy) * g.sy - my;
      dy2[t] = d * d;
      ey[t] = exp(-dy2[t] * g.inv_two_sigma2);
    }
    __syncthreads();
    // pixel t = jj wi + ii of the window; t < wi hj exactly when jj < hj.  A step of 256 pixels
    // advances (jj, ii) by (256 / wi, 256 mod wi) with one carry, as 256 mod wi < wi.
    const int step_j = 256 / wi, step_i = 256 % wi;
    int jj = (int)threadIdx.x / wi, ii = (int)threadIdx.x % wi;
    const double* win = gsb + j0 * g.nx + i0;
    for (int t0 = 0; t0 < wi * hj; t0 += 256 * kAdjBatch) {
      double gv[kAdjBatch];
      int pj[kAdjBatch], pi[kAdjBatch];
#pragma unroll
      for (int q = 0; q < kAdjBatch; ++q) {
        pj[q] = jj;
        pi[q] = ii;
        gv[q] = (jj < hj) ? win[(int64_t)jj * g.nx + ii] : 0.0;
        ii += step_i;
        jj += step_j;
        if (ii >= wi) {
          ii -= wi;
          ++jj;
        }
      }
#pragma unroll
      for (int q = 0; q < kAdjBatch; ++q) {
        if (pj[q] >= hj) break;
        double val = 0.0;
        for (int v = 0; v < ovs; ++v) {
          const double eyv = ey[v * hj + pj[q]], dyv = dy2[v * hj + pj[q]];
          for (int u = 0; u < ovs; ++u) {
            const double r2 = dx2[u * wi + pi[q]] + dyv;
            if (r2 <= g.r2max) val += ex[u * wi + pi[q]] * eyv - g.offset;
          }
        }
        acc += gv[q] * val;
      }
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) red[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) gw[b * g.n_act * g.n_act + a] = (red[0] + red[1] + red[2] + red[3]) * g.norm;
}

// Doubles of the separable adjoint's tables for the largest window (a window spans
// ceil(A + 2R/d) - floor(A) + 3 <= 2R/d + 5 pixels per axis); every window fits when this is at
// most kSepLds.
double adjoint_sep_tab(const DmGeom& g) {
  const double wi = std::min(std::ceil(2.0 * g.radius / g.dx) + 6.0, (double)g.nx);
  const double hj = std::min(std::ceil(2.0 * g.radius / g.dy) + 6.0, (double)g.ny);
  return 2.0 * g.ovs * (wi + hj);
}

// ---------------------------------------------------------------------------------------------
// Influence functions sampled in a table (the Xinetics DM5V2 actuator shape): IF_a(p) is the
// bilinear interpolation of the table at q - m_a (zero outside the table), with q the same tilted
// pupil point as above (no dithers), times cos(y_tilt) / cos(x_tilt) as the reference writes it
// (hcipy/optics/deformable_mirror.py:95-138).  The table (at most 128 x 128 FP64) is staged in
// LDS by every workgroup; grid-stride loops keep the number of workgroups, and so of table
// loads, bounded.

constexpr int kTableMax = 128 * 128;

struct DmTableGeom {
  int64_t nx, ny, n_act;
  double x0, y0, dx, dy;
  double ax0, ay0, pitch;
  double sx, sy, cz, sz;
  double norm;
  int tnx, tny;
  double tx0, ty0, tdx, tdy;  // table coordinates tx0 + i tdx
  double tx1, ty1;            // last table coordinates
};

DmTableGeom make_table_geom(const hcipy_dm_table_desc& d) {
  DmTableGeom g;
  g.nx = d.nx;
  g.ny = d.ny;
  g.n_act = d.n_act;
  g.x0 = d.x0;
  g.y0 = d.y0;
  g.dx = d.dx;
  g.dy = d.dy;
  g.ax0 = d.act_x0;
  g.ay0 = d.act_y0;
  g.pitch = d.pitch;
  g.sx = 1.0 / std::cos(d.y_tilt);
  g.sy = 1.0 / std::cos(d.x_tilt);
  g.cz = std::cos(d.z_tilt);
  g.sz = std::sin(d.z_tilt);
  g.norm = std::cos(d.y_tilt) / std::cos(d.x_tilt);
  g.tnx = (int)d.table_nx;
  g.tny = (int)d.table_ny;
  g.tx0 = d.table_x0;
  g.ty0 = d.table_y0;
  g.tdx = d.table_dx;
  g.tdy = d.table_dy;
  g.tx1 = d.table_x0 + (double)(d.table_nx - 1) * d.table_dx;
  g.ty1 = d.table_y0 + (double)(d.table_ny - 1) * d.table_dy;
  return g;
}

__device__ inline void table_point(const DmTableGeom& g, int64_t i, int64_t j, double& qx, double& qy) {
  const double px = (g.x0 + (double)i * g.dx) * g.sx;
  const double py = (g.y0 + (double)j * g.dy) * g.sy;
  qx = g.cz * px + g.sz * py;
  qy = -g.sz * px + g.cz * py;
}

// Bilinear interpolation of the LDS table at (u, v); zero outside [tx0, tx1] x [ty0, ty1].
__device__ inline double table_interp(const DmTableGeom& g, const double* T, double u, double v) {
  if (!(u >= g.tx0 && u <= g.tx1 && v >= g.ty0 && v <= g.ty1)) return 0.0;
  const double fx = (u - g.tx0) / g.tdx, fy = (v - g.ty0) / g.tdy;
  const int i = min((int)fx, g.tnx - 2), j = min((int)fy, g.tny - 2);
  const double a = fx - (double)i, b = fy - (double)j;
  const double* r0 = T + j * g.tnx + i;
  const double* r1 = r0 + g.tnx;
  return (1.0 - b) * ((1.0 - a) * r0[0] + a * r0[1]) + b * ((1.0 - a) * r1[0] + a * r1[1]);
}

__device__ inline void stage_table(const double* __restrict__ table, double* T, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) T[t] = table[t];
  __syncthreads();
}

__global__ __launch_bounds__(256) void dm_table_surface_kernel(DmTableGeom g, const double* __restrict__ table,
                                                               const double* __restrict__ w,
                                                               double* __restrict__ surface, int64_t b_off) {
  extern __shared__ double dm_tab[];
  stage_table(table, dm_tab, g.tnx * g.tny);
  const int64_t npix = g.nx * g.ny;
  const int64_t b = blockIdx.y + b_off;
  const double* wb = w + b * g.n_act * g.n_act;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = p % g.nx, j = p / g.nx;
    double qx, qy;
    table_point(g, i, j, qx, qy);
    // actuators m with q - m inside the table
    const int64_t ia0 = max((int64_t)0, (int64_t)ceil((qx - g.tx1 - g.ax0) / g.pitch));
    const int64_t ia1 = min(g.n_act - 1, (int64_t)floor((qx - g.tx0 - g.ax0) / g.pitch));
    const int64_t ja0 = max((int64_t)0, (int64_t)ceil((qy - g.ty1 - g.ay0) / g.pitch));
    const int64_t ja1 = min(g.n_act - 1, (int64_t)floor((qy - g.ty0 - g.ay0) / g.pitch));
    double acc = 0.0;
    for (int64_t ja = ja0; ja <= ja1; ++ja) {
      const double my = g.ay0 + (double)ja * g.pitch;
      for (int64_t ia = ia0; ia <= ia1; ++ia) {
        const double wa = wb[ja * g.n_act + ia];
        if (wa != 0.0) acc += wa * table_interp(g, dm_tab, qx - (g.ax0 + (double)ia * g.pitch), qy - my);
      }
    }
    surface[b * npix + p] = acc * g.norm;
  }
}

// grad_w[b][a] = sum_pixel grad_s[b][pixel] IF_a(pixel); one workgroup per (actuator, batch
// item); the pixel window is the bounding box of the table's support mapped back to the pupil.
__global__ __launch_bounds__(256) void dm_table_adjoint_kernel(DmTableGeom g, const double* __restrict__ table,
                                                               const double* __restrict__ gs,
                                                               double* __restrict__ gw, int64_t b_off) {
  extern __shared__ double dm_tab[];
  __shared__ double red[4];
  stage_table(table, dm_tab, g.tnx * g.tny);
  const int64_t a = blockIdx.x;
  const int64_t b = blockIdx.y + b_off;
  const int64_t npix = g.nx * g.ny;
  const double* gsb = gs + b * npix;
  const double mx = g.ax0 + (double)(a % g.n_act) * g.pitch;
  const double my = g.ay0 + (double)(a / g.n_act) * g.pitch;
  // inverse of q = R(-z) diag(sx, sy) p: p = diag(1/sx, 1/sy) R(z) q, at the four support corners
  double pxmin = 1e300, pxmax = -1e300, pymin = 1e300, pymax = -1e300;
  for (int c = 0; c < 4; ++c) {
    const double qx = mx + ((c & 1) ? g.tx1 : g.tx0);
    const double qy = my + ((c & 2) ? g.ty1 : g.ty0);
    const double px = (g.cz * qx - g.sz * qy) / g.sx;
    const double py = (g.sz * qx + g.cz * qy) / g.sy;
    pxmin = fmin(pxmin, px);
    pxmax = fmax(pxmax, px);
    pymin = fmin(pymin, py);
    pymax = fmax(pymax, py);
  }
  const int64_t i0 = max((int64_t)0, (int64_t)floor((pxmin - g.x0) / g.dx) - 1);
  const int64_t i1 = min(g.nx - 1, (int64_t)ceil((pxmax - g.x0) / g.dx) + 1);
  const int64_t j0 = max((int64_t)0, (int64_t)floor((pymin - g.y0) / g.dy) - 1);
  const int64_t j1 = min(g.ny - 1, (int64_t)ceil((pymax - g.y0) / g.dy) + 1);
  double acc = 0.0;
  if (i0 <= i1 && j0 <= j1) {
    const int64_t wi = i1 - i0 + 1;
    const int64_t cnt = wi * (j1 - j0 + 1);
    for (int64_t t = threadIdx.x; t < cnt; t += blockDim.x) {
      const int64_t i = i0 + t % wi, j = j0 + t / wi;
      double qx, qy;
      table_point(g, i, j, qx, qy);
      acc += gsb[j * g.nx + i] * table_interp(g, dm_tab, qx - mx, qy - my);
    }
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) gw[b * g.n_act * g.n_act + a] = (red[0] + red[1] + red[2] + red[3]) * g.norm;
}

// ---------------------------------------------------------------------------------------------
// Poke windows: for every single-actuator poke of a poke set, the change of the pupil field
//   dE = E (exp(i s IF_a) - 1),   s = 2 k h (phase of the poke height h),
// evaluated only on the wx x wy pixel window that holds IF_a's support (dE is exactly 0 outside it),
// and written as rows of the stage-1 data operand in the plan's format: FP16 hi/lo or BF16
// b0/b1/b2 splits with a power-of-two row scale, or raw complex64 / complex128.  exp(i phi) - 1 is
// formed as (-2 sin^2(phi/2), sin(phi)) so that small phases keep their relative accuracy.
// One wave per window row; the row's values stay in registers (wx <= kPokeMaxW) for the row
// maximum of the split formats.

constexpr int kPokeMaxW = kPokeWindowMax;

struct PokeWinItemDev {
  DmGeom g;
  const void* field;     // base pupil field, complex64 (complex128 when prec is FP64)
  double phase_scale;    // s
  int prec;              // 0 FP64, 1 FP32, 2 FP16 hi/lo, 3 BF16 three-way (the Prec of mft_gemm.hpp)
  int wx, wy;            // window width and height in pixels (multiples of 64, wx <= kPokeMaxW)
  int64_t row0;          // first global window row of this item
  int64_t n_rows;        // n_act^2 * wy
  const int* org;        // window origin (x0, y0) per actuator
  void* out[3];          // [actuator][wy][2 wx] values of each output array
  float* row_scale;      // [actuator][wy] (split formats)
};

// Window rows per wave of the poke window kernel.  Window heights are multiples of 64, so an
// item's rows and an actuator's rows come in whole groups: one item search per group.
constexpr int kPokeRows = 8;

template <int NPL>  // window pixels per lane at most (wx <= 64 NPL for every item)
__global__ __launch_bounds__(256) void poke_window_kernel(const PokeWinItemDev* __restrict__ items, int n_items,
                                                          int64_t total_rows) {
  const int lane = threadIdx.x & 63;
  const int64_t R0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * kPokeRows;
  if (R0 >= total_rows) return;
  int lo = 0, hi = n_items - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].row0 <= R0) lo = mid;
    else hi = mid - 1;
  }
  const PokeWinItemDev& it = items[lo];
  const DmGeom& g = it.g;
  const int w = it.wx;
  const int64_t local0 = R0 - it.row0;
  const int64_t a = local0 / it.wy;
  const int r0 = (int)(local0 - a * it.wy);
  const int64_t x0 = it.org[2 * a], y0 = it.org[2 * a + 1];
  const double mx = g.ax0 + (double)(a % g.n_act) * g.pitch;
  const double my = g.ay0 + (double)(a / g.n_act) * g.pitch;
  for (int rr = 0; rr < kPokeRows; ++rr) {
    const int r = r0 + rr;
    const int64_t j = y0 + r;
    // the row's values stay in registers: pixel c = lane + 64 t of the window
    double2 val[NPL];
    double mxv = 0.0;
#pragma unroll
    for (int t = 0; t < NPL; ++t) {
      const int c = lane + 64 * t;
      const int64_t i = x0 + c;
      double2 d = make_double2(0.0, 0.0);
      if (c < w && i < g.nx && j < g.ny) {
        double ifv = 0.0;
        for (int v = 0; v < g.ovs; ++v)
          for (int u = 0; u < g.ovs; ++u) {
            double qx, qy;
            eval_point(g, i, j, u, v, qx, qy);
            ifv += poke(g, qx, qy, mx, my);
          }
        if (ifv != 0.0) {
          // exp(i phi) - 1 = (-2 sin^2(phi/2), 2 sin(phi/2) cos(phi/2)): one sincos, no cancellation
          double sh, ch;
          sincos(0.5 * it.phase_scale * ifv * g.norm, &sh, &ch);
          const double er = -2.0 * sh * sh, ei = 2.0 * sh * ch;
          double fr, fi;
          if (it.prec == 0) {
            const double2 e = static_cast<const double2*>(it.field)[j * g.nx + i];
            fr = e.x;
            fi = e.y;
          } else {
            const float2 e = static_cast<const float2*>(it.field)[j * g.nx + i];
            fr = e.x;
            fi = e.y;
          }
          d = make_double2(fr * er - fi * ei, fr * ei + fi * er);
          if (it.prec != 0) {
            // complex64 values of the poke change, as the split of a complex64 field would see them
            d = make_double2((double)(float)d.x, (double)(float)d.y);
            mxv = fmax(mxv, fmax(fabs(d.x), fabs(d.y)));
          }
        }
      }
      val[t] = d;
    }
    const int64_t row_off = (a * it.wy + r) * (int64_t)(2 * w);  // reals
    if (it.prec == 0 || it.prec == 1) {
#pragma unroll
      for (int t = 0; t < NPL; ++t) {
        const int c = lane + 64 * t;
        if (c >= w) continue;
        if (it.prec == 0) static_cast<double2*>(it.out[0])[row_off / 2 + c] = val[t];
        else static_cast<float2*>(it.out[0])[row_off / 2 + c] = make_float2((float)val[t].x, (float)val[t].y);
      }
      continue;
    }
    const float m = wave_max((float)mxv);
    float s = 1.0f;
    if (m > 0.0f && m <= 3.0e38f) {
      int e;
      frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)  =>  m 2^-e < 1
      s = ldexpf(1.0f, -e);
    }
    if (lane == 0) it.row_scale[a * it.wy + r] = s;
#pragma unroll
    for (int t = 0; t < NPL; ++t) {
      const int c = lane + 64 * t;
      if (c >= w) continue;
      const float vr = (float)val[t].x * s, vi = (float)val[t].y * s;
      if (it.prec == 2) {
        const _Float16 hr = (_Float16)vr, hi = (_Float16)vi;
        static_cast<half2_t*>(it.out[0])[row_off / 2 + c] = half2_t{hr, hi};
        static_cast<half2_t*>(it.out[1])[row_off / 2 + c] =
            half2_t{(_Float16)((vr - (float)hr) * 2048.0f), (_Float16)((vi - (float)hi) * 2048.0f)};
      } else {
        const float v2[2] = {vr, vi};
        __bf16 q0[2], q1[2], q2[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          q0[e] = (__bf16)v2[e];
          const float rr = v2[e] - (float)q0[e];
          q1[e] = (__bf16)rr;
          q2[e] = (__bf16)(rr - (float)q1[e]);
        }
        static_cast<bf16x2_t*>(it.out[0])[row_off / 2 + c] = bf16x2_t{q0[0], q0[1]};
        static_cast<bf16x2_t*>(it.out[1])[row_off / 2 + c] = bf16x2_t{q1[0], q1[1]};
        static_cast<bf16x2_t*>(it.out[2])[row_off / 2 + c] = bf16x2_t{q2[0], q2[1]};
      }
    }
  }
}

int check_table_desc(const hcipy_dm_table_desc* d) {
  HCIPY_REQUIRE(d != nullptr, "NULL descriptor");
  HCIPY_REQUIRE(d->nx > 0 && d->ny > 0 && d->n_act > 0, "bad grid or actuator count");
  HCIPY_REQUIRE(d->dx > 0 && d->dy > 0 && d->pitch > 0, "bad geometry");
  HCIPY_REQUIRE(d->table != nullptr && d->table_nx >= 2 && d->table_ny >= 2 &&
                    d->table_nx * d->table_ny <= kTableMax && d->table_dx > 0 && d->table_dy > 0,
                "bad influence table (2..%d samples per axis, at most %d in total)", kTableMax, kTableMax);
  HCIPY_REQUIRE(std::cos(d->x_tilt) > 0 && std::cos(d->y_tilt) > 0, "tilts must be below 90 degrees");
  return kOk;
}

int check_desc(const hcipy_dm_gaussian_desc* d) {
  HCIPY_REQUIRE(d != nullptr, "NULL descriptor");
  HCIPY_REQUIRE(d->nx > 0 && d->ny > 0 && d->n_act > 0, "bad grid or actuator count");
  HCIPY_REQUIRE(d->dx > 0 && d->dy > 0 && d->pitch > 0 && d->sigma > 0 && d->radius > 0, "bad geometry");
  HCIPY_REQUIRE(d->oversampling >= 1 && d->oversampling <= 16, "oversampling must be 1..16");
  HCIPY_REQUIRE(std::cos(d->x_tilt) > 0 && std::cos(d->y_tilt) > 0, "tilts must be below 90 degrees");
  return kOk;
}

}  // namespace

int check_gaussian_desc(const hcipy_dm_gaussian_desc* d) { return check_desc(d); }

// Pixel window [i0, i1] x [j0, j1] (inclusive, clipped to the grid) that holds the support of
// actuator a's influence function, widened by one pixel for the dithers.
void gaussian_support_window(const hcipy_dm_gaussian_desc& d, int64_t a, int64_t* i0, int64_t* i1, int64_t* j0,
                             int64_t* j1) {
  const DmGeom g = make_geom(d);
  const double mx = g.ax0 + (double)(a % g.n_act) * g.pitch;
  const double my = g.ay0 + (double)(a / g.n_act) * g.pitch;
  const double cx = (g.cz * mx - g.sz * my) / g.sx;
  const double cy = (g.sz * mx + g.cz * my) / g.sy;
  *i0 = std::max((int64_t)0, (int64_t)std::floor((cx - g.radius - g.x0) / g.dx) - 1);
  *i1 = std::min(g.nx - 1, (int64_t)std::ceil((cx + g.radius - g.x0) / g.dx) + 1);
  *j0 = std::max((int64_t)0, (int64_t)std::floor((cy - g.radius - g.y0) / g.dy) - 1);
  *j1 = std::min(g.ny - 1, (int64_t)std::ceil((cy + g.radius - g.y0) / g.dy) + 1);
}

int poke_window_bytes(int n_items) { return n_items * (int)sizeof(PokeWinItemDev); }

// Writes the device descriptors of `n` items to `dev_items` (device memory of poke_window_bytes)
// through the pinned host buffer `host` and returns; the copy is ordered on `s`.
int upload_poke_items(const PokeWinItem* items, int n, void* host, void* dev_items, hipStream_t s) {
  PokeWinItemDev* h = static_cast<PokeWinItemDev*>(host);
  for (int k = 0; k < n; ++k) {
    h[k].g = make_geom(items[k].dm);
    h[k].field = items[k].field;
    h[k].phase_scale = items[k].phase_scale;
    h[k].prec = items[k].prec;
    h[k].wx = items[k].wx;
    h[k].wy = items[k].wy;
    h[k].row0 = items[k].row0;
    h[k].n_rows = items[k].dm.n_act * items[k].dm.n_act * (int64_t)items[k].wy;
    h[k].org = items[k].org;
    for (int q = 0; q < 3; ++q) h[k].out[q] = items[k].out[q];
    h[k].row_scale = items[k].row_scale;
  }
  HCIPY_HIP(hipMemcpyAsync(dev_items, host, (size_t)n * sizeof(PokeWinItemDev), hipMemcpyHostToDevice, s));
  return kOk;
}

int launch_poke_windows(const void* dev_items, int n_items, int64_t total_rows, int max_wx, hipStream_t s) {
  if (total_rows <= 0) return kOk;
  HCIPY_REQUIRE(ceil_div(total_rows, 4 * kPokeRows) <= INT32_MAX, "too many poke window rows");
  HCIPY_REQUIRE(total_rows % kPokeRows == 0, "poke window rows must come in groups of %d", kPokeRows);
  HCIPY_REQUIRE(max_wx >= 1 && max_wx <= kPokeMaxW, "poke window width %d", max_wx);
  const dim3 g((unsigned)ceil_div(total_rows, 4 * kPokeRows));
  const auto* it = static_cast<const PokeWinItemDev*>(dev_items);
  if (max_wx <= 64) hipLaunchKernelGGL(poke_window_kernel<1>, g, dim3(256), 0, s, it, n_items, total_rows);
  else if (max_wx <= 128) hipLaunchKernelGGL(poke_window_kernel<2>, g, dim3(256), 0, s, it, n_items, total_rows);
  else if (max_wx <= 256) hipLaunchKernelGGL(poke_window_kernel<4>, g, dim3(256), 0, s, it, n_items, total_rows);
  else if (max_wx <= 512) hipLaunchKernelGGL(poke_window_kernel<8>, g, dim3(256), 0, s, it, n_items, total_rows);
  else hipLaunchKernelGGL(poke_window_kernel<16>, g, dim3(256), 0, s, it, n_items, total_rows);
  return check_hip(hipGetLastError(), "poke_window_kernel launch");
}

}  // namespace hcipy

using namespace hcipy;

extern "C" {

int hcipy_dm_gaussian_surface(const hcipy_dm_gaussian_desc* desc, const double* actuators, double* surface,
                              int64_t batch, void* stream) {
  if (int st = check_desc(desc)) return st;
  HCIPY_REQUIRE(batch >= 0, "bad batch");
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(actuators != nullptr && surface != nullptr, "NULL pointer");
  const DmGeom g = make_geom(*desc);
  HCIPY_REQUIRE(ceil_div(g.ny, kSurfRows) <= 65535, "pupil grid has too many rows");
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const dim3 grid((unsigned)ceil_div(g.nx, 256), (unsigned)ceil_div(g.ny, kSurfRows),
                    (unsigned)std::min<int64_t>(65535, batch - b0));
    if (separable(g))
      hipLaunchKernelGGL(dm_surface_sep_kernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream), g, actuators,
                         surface, b0);
    else
      hipLaunchKernelGGL(dm_surface_kernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream), g, actuators,
                         surface, b0);
    HCIPY_HIP(hipGetLastError());
  }
  return kOk;
}

int hcipy_dm_gaussian_surface_adjoint(const hcipy_dm_gaussian_desc* desc, const double* grad_surface,
                                      double* grad_actuators, int64_t batch, void* stream) {
  if (int st = check_desc(desc)) return st;
  HCIPY_REQUIRE(batch >= 0, "bad batch");
  HCIPY_REQUIRE(desc->n_act * desc->n_act <= INT32_MAX, "too many actuators");
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(grad_surface != nullptr && grad_actuators != nullptr, "NULL pointer");
  const DmGeom g = make_geom(*desc);
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const dim3 grid((unsigned)(g.n_act * g.n_act), (unsigned)std::min<int64_t>(65535, batch - b0));
    const double tab = adjoint_sep_tab(g);
    if (separable(g) && tab <= kSepLds)
      hipLaunchKernelGGL(dm_adjoint_sep_kernel, grid, dim3(256), (size_t)tab * sizeof(double),
                         static_cast<hipStream_t>(stream), g, grad_surface, grad_actuators, b0, (int)tab);
    else
      hipLaunchKernelGGL(dm_adjoint_kernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream), g, grad_surface,
                         grad_actuators, b0);
    HCIPY_HIP(hipGetLastError());
  }
  return kOk;
}

int hcipy_dm_table_surface(const hcipy_dm_table_desc* desc, const double* actuators, double* surface, int64_t batch,
                           void* stream) {
  if (int st = check_table_desc(desc)) return st;
  HCIPY_REQUIRE(batch >= 0, "bad batch");
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(actuators != nullptr && surface != nullptr, "NULL pointer");
  const DmTableGeom g = make_table_geom(*desc);
  const size_t lds = (size_t)g.tnx * g.tny * sizeof(double);
  HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(dm_table_surface_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int64_t npix = g.nx * g.ny;
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const dim3 grid((unsigned)std::min<int64_t>(ceil_div(npix, 256), 1024), (unsigned)std::min<int64_t>(65535, batch - b0));
    hipLaunchKernelGGL(dm_table_surface_kernel, grid, dim3(256), lds, static_cast<hipStream_t>(stream), g, desc->table,
                       actuators, surface, b0);
    HCIPY_HIP(hipGetLastError());
  }
  return kOk;
}

int hcipy_dm_table_surface_adjoint(const hcipy_dm_table_desc* desc, const double* grad_surface, double* grad_actuators,
                                   int64_t batch, void* stream) {
  if (int st = check_table_desc(desc)) return st;
  HCIPY_REQUIRE(batch >= 0, "bad batch");
  HCIPY_REQUIRE(desc->n_act * desc->n_act <= INT32_MAX, "too many actuators");
  if (batch == 0) return kOk;
  HCIPY_REQUIRE(grad_surface != nullptr && grad_actuators != nullptr, "NULL pointer");
  const DmTableGeom g = make_table_geom(*desc);
  const size_t lds = (size_t)g.tnx * g.tny * sizeof(double);
  HCIPY_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(dm_table_adjoint_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const dim3 grid((unsigned)(g.n_act * g.n_act), (unsigned)std::min<int64_t>(65535, batch - b0));
    hipLaunchKernelGGL(dm_table_adjoint_kernel, grid, dim3(256), lds, static_cast<hipStream_t>(stream), g, desc->table,
                       grad_surface, grad_actuators, b0);
    HCIPY_HIP(hipGetLastError());
  }
  return kOk;
}

}  // extern "C"
