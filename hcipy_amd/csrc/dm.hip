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
#include <cmath>

#include "common.hpp"
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
    hipLaunchKernelGGL(dm_surface_kernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream), g, actuators, surface,
                       b0);
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
