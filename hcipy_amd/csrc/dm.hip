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

// surface[b][pixel] = sum_a w[b][a] IF_a(pixel)
__global__ __launch_bounds__(256) void dm_surface_kernel(DmGeom g, const double* __restrict__ w,
                                                         double* __restrict__ surface, int64_t b_off) {
  const int64_t npix = g.nx * g.ny;
  const int64_t b = blockIdx.y + b_off;
  const double* wb = w + b * g.n_act * g.n_act;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = p % g.nx, j = p / g.nx;
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
    surface[b * npix + p] = acc * g.norm;
  }
}

// Without a z tilt the transformed point is (px sx, py sy) and the Gaussian factorises,
//   exp(-(ddx^2 + ddy^2) / (2 sigma^2)) = exp(-ddx^2 / (2 sigma^2)) exp(-ddy^2 / (2 sigma^2)),
// so a point needs one exp per actuator row and one per actuator column in reach instead of one per
// actuator; the support test compares the same ddx^2 + ddy^2 as the general kernels.  At most
// kSepMax actuators per axis may lie within reach of a point (a support radius of up to 3 pitches,
// the reference's default cutoff; else the general kernels run).
constexpr int kSepMax = 8;

bool separable(const DmGeom& g) { return g.sz == 0.0 && g.cz == 1.0 && 2.0 * g.radius / g.pitch + 2.0 <= kSepMax; }

__global__ __launch_bounds__(256) void dm_surface_sep_kernel(DmGeom g, const double* __restrict__ w,
                                                             double* __restrict__ surface, int64_t b_off) {
  const int64_t npix = g.nx * g.ny;
  const int64_t b = blockIdx.y + b_off;
  const double* wb = w + b * g.n_act * g.n_act;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = p % g.nx, j = p / g.nx;
    double acc = 0.0;
    for (int v = 0; v < g.ovs; ++v) {
      const double fy = ((v + 0.5) / g.ovs - 0.5) * g.dy;
      const double qy = (g.y0 + (double)j * g.dy + fy) * g.sy;
      const int64_t ja0 = max((int64_t)0, (int64_t)ceil((qy - g.radius - g.ay0) / g.pitch));
      const int64_t ja1 = min(g.n_act - 1, (int64_t)floor((qy + g.radius - g.ay0) / g.pitch));
      const int nja = (int)(ja1 - ja0 + 1);
      double ey[kSepMax], dy2[kSepMax];
#pragma unroll
      for (int k = 0; k < kSepMax; ++k) {
        const double dya = qy - (g.ay0 + (double)(ja0 + k) * g.pitch);
        dy2[k] = dya * dya;
        ey[k] = 0.0;
        if (k < nja) ey[k] = exp(-dy2[k] * g.inv_two_sigma2);
      }
      for (int u = 0; u < g.ovs; ++u) {
        const double fx = ((u + 0.5) / g.ovs - 0.5) * g.dx;
        const double qx = (g.x0 + (double)i * g.dx + fx) * g.sx;
        const int64_t ia0 = max((int64_t)0, (int64_t)ceil((qx - g.radius - g.ax0) / g.pitch));
        const int64_t ia1 = min(g.n_act - 1, (int64_t)floor((qx + g.radius - g.ax0) / g.pitch));
        for (int64_t ia = ia0; ia <= ia1; ++ia) {
          const double dxa = qx - (g.ax0 + (double)ia * g.pitch);
          const double dx2 = dxa * dxa;
          if (dx2 > g.r2max) continue;
          const double ex = exp(-dx2 * g.inv_two_sigma2);
#pragma unroll
          for (int k = 0; k < kSepMax; ++k) {
            if (k < nja && dx2 + dy2[k] <= g.r2max) acc += wb[(ja0 + k) * g.n_act + ia] * (ex * ey[k] - g.offset);
          }
        }
      }
    }
    surface[b * npix + p] = acc * g.norm;
  }
}

// grad_w[b][a] = sum_pixel grad_s[b][pixel] IF_a(pixel): one workgroup per (actuator, batch item),
// a fixed thread -> pixel assignment and a fixed reduction tree (deterministic).
__global__ __launch_bounds__(256) void dm_adjoint_kernel(DmGeom g, const double* __restrict__ gs,
                                                         double* __restrict__ gw, int64_t b_off) {
  __shared__ double red[4];
  const int64_t a = blockIdx.x;
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
  double acc = 0.0;
  if (i0 <= i1 && j0 <= j1) {
    const int64_t wi = i1 - i0 + 1;
    const int64_t cnt = wi * (j1 - j0 + 1);
    for (int64_t t = threadIdx.x; t < cnt; t += blockDim.x) {
      const int64_t i = i0 + t % wi, j = j0 + t / wi;
      double val = 0.0;
      for (int v = 0; v < g.ovs; ++v)
        for (int u = 0; u < g.ovs; ++u) {
          double qx, qy;
          eval_point(g, i, j, u, v, qx, qy);
          val += poke(g, qx, qy, mx, my);
        }
      acc += gsb[j * g.nx + i] * val;
    }
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) gw[b * g.n_act * g.n_act + a] = (red[0] + red[1] + red[2] + red[3]) * g.norm;
}

// Separable adjoint (no z tilt): the per-axis factors exp(-ddx^2 / (2 sigma^2)) and squared
// distances of the window's columns and rows, per dither, are tabulated in LDS first
// (kSepLds doubles hold 2 ovs (wi + hj) of them), then every window pixel needs multiply-adds only.
constexpr int kSepLds = 4096;

__global__ __launch_bounds__(256) void dm_adjoint_sep_kernel(DmGeom g, const double* __restrict__ gs,
                                                             double* __restrict__ gw, int64_t b_off) {
  __shared__ double red[4];
  __shared__ double tab[kSepLds];
  const int64_t a = blockIdx.x;
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
  double acc = 0.0;
  const int wi = (int)(i1 - i0 + 1), hj = (int)(j1 - j0 + 1);
  const int ovs = g.ovs;
  if (wi > 0 && hj > 0 && 2 * ovs * (wi + hj) <= kSepLds) {
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
    const int64_t cnt = (int64_t)wi * hj;
    for (int64_t t = threadIdx.x; t < cnt; t += blockDim.x) {
      const int jj = (int)(t / wi), ii = (int)(t - (int64_t)jj * wi);
      double val = 0.0;
      for (int v = 0; v < ovs; ++v) {
        const double eyv = ey[v * hj + jj], dyv = dy2[v * hj + jj];
        for (int u = 0; u < ovs; ++u) {
          const double r2 = dx2[u * wi + ii] + dyv;
          if (r2 <= g.r2max) val += ex[u * wi + ii] * eyv - g.offset;
        }
      }
      acc += gsb[(j0 + jj) * g.nx + (i0 + ii)] * val;
    }
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) gw[b * g.n_act * g.n_act + a] = (red[0] + red[1] + red[2] + red[3]) * g.norm;
}

// Every actuator's window fits the separable adjoint's tables: a window spans
// ceil(A + 2R/d) - floor(A) + 3 <= 2R/d + 5 pixels per axis.
bool adjoint_sep_fits(const DmGeom& g) {
  const double wi = std::min(std::ceil(2.0 * g.radius / g.dx) + 6.0, (double)g.nx);
  const double hj = std::min(std::ceil(2.0 * g.radius / g.dy) + 6.0, (double)g.ny);
  return 2.0 * g.ovs * (wi + hj) <= kSepLds;
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

template <int NPL>  // window pixels per lane at most (wx <= 64 NPL for every item)
__global__ __launch_bounds__(256) void poke_window_kernel(const PokeWinItemDev* __restrict__ items, int n_items,
                                                          int64_t total_rows) {
  const int lane = threadIdx.x & 63;
  const int64_t R = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (R >= total_rows) return;
  int lo = 0, hi = n_items - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].row0 <= R) lo = mid;
    else hi = mid - 1;
  }
  const PokeWinItemDev& it = items[lo];
  const DmGeom& g = it.g;
  const int w = it.wx;
  const int64_t local = R - it.row0;
  const int64_t a = local / it.wy;
  const int r = (int)(local - a * it.wy);
  const int64_t x0 = it.org[2 * a], y0 = it.org[2 * a + 1];
  const int64_t j = y0 + r;
  const double mx = g.ax0 + (double)(a % g.n_act) * g.pitch;
  const double my = g.ay0 + (double)(a / g.n_act) * g.pitch;
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
    return;
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
  HCIPY_REQUIRE(ceil_div(total_rows, 4) <= INT32_MAX, "too many poke window rows");
  HCIPY_REQUIRE(max_wx >= 1 && max_wx <= kPokeMaxW, "poke window width %d", max_wx);
  const dim3 g((unsigned)ceil_div(total_rows, 4));
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
  const int64_t npix = g.nx * g.ny;
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    const dim3 grid((unsigned)std::min<int64_t>(ceil_div(npix, 256), 8192), (unsigned)std::min<int64_t>(65535, batch - b0));
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
    if (separable(g) && adjoint_sep_fits(g))
      hipLaunchKernelGGL(dm_adjoint_sep_kernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream), g, grad_surface,
                         grad_actuators, b0);
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
