// sph_theta.hip -- deterministic gtheta reduction (no float atomics): the FP64
// form's gather and the node sums of both forms (the f32c form's keyed gather
// is in sph_f32c.hip, where P1 runs it in its own launch).
//
// The reference scatters each particle's four bilinear weights into gtheta
// with float atomics (adj_design_kernel!, src/adjoint.jl:166-185), whose
// summation order changes from run to run.  Here the scatter becomes two
// gathers in a fixed order:
//   A  eight lanes per theta cell (i0, j0): sweep the SPH cells that can hold
//      particles whose interp_cell is (i0, j0) (plus one cell of margin for
//      the displacement since the last rebuild), recompute interp_cell exactly
//      as the reference does, and sum the four weighted contributions
//      (ga*(1-fx))*(1-fy), (ga*fx)*(1-fy), (ga*(1-fx))*fy, (ga*fx)*fy;
//      the lanes take candidates round-robin and combine by a fixed butterfly;
//      The partials of the steps of one neighbour-list epoch add up in
//      the partial array (the epoch's first step writes it);
//   B  once per epoch, one thread per theta node: add the partials of the
//      (up to) four cells that touch the node, in a fixed order, to gtheta.
// Every contribution is the reference's product, bit for bit; only the order
// of the sum is fixed.  FP64 throughout; compiled with -ffp-contract=off.
#pragma clang fp contract(off)
#include "sph_f32c.h"

namespace sph {

namespace {

inline int nblocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// SPH cell (0-based, clamped) of coordinate v for the division rule.
__device__ __forceinline__ int cell_of(double v, double cs, int n) {
  return cell_coord_div(v / cs, n);
}

constexpr int kLanes = 8;   // lanes that share one theta cell

template <class GA>
__global__ void __launch_bounds__(kBlock)
k_theta_cells(const double2 *__restrict__ X, const GA *__restrict__ galpha,
              const int *__restrict__ cstart, Grid g, ThetaGrid T, int first,
              double4 *__restrict__ part, const unsigned long long *__restrict__ lev, int stamp) {
  const int ncx = T.ngx - 1, ncy = T.ngy - 1;
  const int lane = threadIdx.x & (kLanes - 1);
  // theta cells in x-fastest order, so that neighbouring groups read
  // neighbouring slot runs (and, through xcd_block, stay on one XCD)
  const int t = (xcd_block() * blockDim.x + threadIdx.x) / kLanes;
  if (t >= ncx * ncy) return;      // the kLanes lanes of a group leave together
  const int i0 = t % ncx;
  const int j0 = t / ncx;
  // SPH cell range holding particles with interp cell (i0, j0), widened by a
  // one-cell margin, or by the step's larger sweep level (raised by its R pass
  // for a slot that moved further since the rebuild)
  const int lv = level_of(lev, stamp);
  const int mg = lv > 1 ? lv : 1;
  int cx0 = (i0 == 0) ? 0 : cell_of((double)i0 * T.hx, T.cs, g.nx) - mg;
  int cx1 = (i0 == ncx - 1) ? g.nx - 1 : cell_of((double)(i0 + 1) * T.hx, T.cs, g.nx) + mg;
  int cy0 = (j0 == 0) ? 0 : cell_of((double)j0 * T.hy, T.cs, g.ny) - mg;
  int cy1 = (j0 == ncy - 1) ? g.ny - 1 : cell_of((double)(j0 + 1) * T.hy, T.cs, g.ny) + mg;
  cx0 = cx0 < 0 ? 0 : cx0;
  cy0 = cy0 < 0 ? 0 : cy0;
  cx1 = cx1 > g.nx - 1 ? g.nx - 1 : cx1;
  cy1 = cy1 > g.ny - 1 ? g.ny - 1 : cy1;
  // cheap pre-filter in grid units (generous margin; the exact test follows)
  const double lox = (i0 == 0) ? -1e300 : (double)i0 - 0.01;
  const double hix = (i0 == ncx - 1) ? 1e300 : (double)(i0 + 1) + 0.01;
  const double loy = (j0 == 0) ? -1e300 : (double)j0 - 0.01;
  const double hiy = (j0 == ncy - 1) ? 1e300 : (double)(j0 + 1) + 0.01;
  double s00 = 0.0, s10 = 0.0, s01 = 0.0, s11 = 0.0;
  for (int cy = cy0; cy <= cy1; ++cy) {
    const int b = cstart[cy * g.nx + cx0], e = cstart[cy * g.nx + cx1 + 1];
    for (int p = b + lane; p < e; p += kLanes) {
      const double2 x = X[p];
      const double ax = x.x * T.inv_hx, ay = x.y * T.inv_hy;
      if (ax < lox || ax >= hix || ay < loy || ay >= hiy) continue;
      // interp_cell (src/params.jl:53-68), exactly
      double gx = x.x / T.hx;
      double gy = x.y / T.hy;
      double gxc = (gx > T.gxclamp) ? T.gxclamp : ((gx < 0.0) ? 0.0 : gx);
      double gyc = (gy > T.gyclamp) ? T.gyclamp : ((gy < 0.0) ? 0.0 : gy);
      double fxf = floor(gxc), fyf = floor(gyc);
      int pi = (fxf > -1e9 && fxf < 1e9) ? (int)fxf : -1000000000;
      int pj = (fyf > -1e9 && fyf < 1e9) ? (int)fyf : -1000000000;
      pi = pi < 0 ? 0 : (pi > ncx - 1 ? ncx - 1 : pi);
      pj = pj < 0 ? 0 : (pj > ncy - 1 ? ncy - 1 : pj);
      if (pi != i0 || pj != j0) continue;
      const double fx = gxc - (double)pi;
      const double fy = gyc - (double)pj;
      const double ga = (double)galpha[p];
      s00 += (ga * (1.0 - fx)) * (1.0 - fy);
      s10 += (ga * fx) * (1.0 - fy);
      s01 += (ga * (1.0 - fx)) * fy;
      s11 += (ga * fx) * fy;
    }
  }
  // fixed butterfly over the group's lanes: the same order on every run
#pragma unroll
  for (int m = kLanes / 2; m > 0; m >>= 1) {
    s00 += __shfl_xor(s00, m, kLanes);
    s10 += __shfl_xor(s10, m, kLanes);
    s01 += __shfl_xor(s01, m, kLanes);
    s11 += __shfl_xor(s11, m, kLanes);
  }
  if (lane == 0) {
    double4 &pt = part[j0 + i0 * ncy];
    if (first) {
      pt = make_double4(s00, s10, s01, s11);
    } else {
      const double4 o = pt;
      pt = make_double4(o.x + s00, o.y + s10, o.z + s01, o.w + s11);
    }
  }
}

// node sums of the theta-cell partials, per scene (workgroups dealt in order
// from Scene32::nboff)
__global__ void k_theta_nodes_b(const double4 *__restrict__ part, Batch B,
                                double *__restrict__ gtheta) {
  const Scene32 &Sc = B.sc[B.nb2s[blockIdx.x]];
  const int l = (blockIdx.x - Sc.nboff) * blockDim.x + threadIdx.x;
  const int ngx = Sc.T.ngx, ngy = Sc.T.ngy;
  if (l >= ngx * ngy) return;
  const double4 *pt = part + Sc.tcoff;
  const int j = l % ngy, i = l / ngy;
  const int ncx = ngx - 1, ncy = ngy - 1;
  double acc = 0.0;
  if (i < ncx && j < ncy) acc += pt[j + i * ncy].x;                    // l00 corner
  if (i > 0 && j < ncy) acc += pt[j + (i - 1) * ncy].y;                // l00 + ny
  if (i < ncx && j > 0) acc += pt[(j - 1) + i * ncy].z;                // l00 + 1
  if (i > 0 && j > 0) acc += pt[(j - 1) + (i - 1) * ncy].w;            // l00 + ny + 1
  gtheta[Sc.toff + l] += acc;
}

__global__ void k_theta_nodes(const double4 *__restrict__ part, ThetaGrid T,
                              double *__restrict__ gtheta) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= T.ngx * T.ngy) return;
  const int j = l % T.ngy, i = l / T.ngy;
  const int ncx = T.ngx - 1, ncy = T.ngy - 1;
  double acc = 0.0;
  if (i < ncx && j < ncy) acc += part[j + i * ncy].x;                    // l00 corner
  if (i > 0 && j < ncy) acc += part[j + (i - 1) * ncy].y;                // l00 + ny
  if (i < ncx && j > 0) acc += part[(j - 1) + i * ncy].z;                // l00 + 1
  if (i > 0 && j > 0) acc += part[(j - 1) + (i - 1) * ncy].w;            // l00 + ny + 1
  gtheta[l] += acc;
}

}  // namespace

void launch_theta_cells_f64(hipStream_t s, const double2 *X, const double *galpha,
                            const int *cstart, Grid g, ThetaGrid T, bool first, double4 *part,
                            const unsigned long long *lev, int stamp) {
  const long long ncells = (long long)(T.ngx - 1) * (T.ngy - 1);
  k_theta_cells<double><<<nblocks(ncells * kLanes), kBlock, 0, s>>>(X, galpha, cstart, g, T,
                                                                    first ? 1 : 0, part, lev,
                                                                    stamp);
}

void launch_theta_nodes_f64(hipStream_t s, const double4 *part, ThetaGrid T, double *gtheta) {
  k_theta_nodes<<<nblocks((long long)T.ngx * T.ngy), kBlock, 0, s>>>(part, T, gtheta);
}

void launch_theta_nodes_keyed_f32(hipStream_t s, const double4 *part, const Batch &B,
                                  double *gtheta) {
  k_theta_nodes_b<<<B.nnblk, kBlock, 0, s>>>(part, B, gtheta);
}

}  // namespace sph
