// sph_theta.hip -- deterministic gtheta reduction (no float atomics).
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
              double4 *__restrict__ part) {
  const int ncx = T.ngx - 1, ncy = T.ngy - 1;
  const int lane = threadIdx.x & (kLanes - 1);
  // theta cells in x-fastest order, so that neighbouring groups read
  // neighbouring slot runs (and, through xcd_block, stay on one XCD)
  const int t = (xcd_block() * blockDim.x + threadIdx.x) / kLanes;
  if (t >= ncx * ncy) return;      // the kLanes lanes of a group leave together
  const int i0 = t % ncx;
  const int j0 = t / ncx;
  // SPH cell range holding particles with interp cell (i0, j0), +1 margin
  int cx0 = (i0 == 0) ? 0 : cell_of((double)i0 * T.hx, T.cs, g.nx) - 1;
  int cx1 = (i0 == ncx - 1) ? g.nx - 1 : cell_of((double)(i0 + 1) * T.hx, T.cs, g.nx) + 1;
  int cy0 = (j0 == 0) ? 0 : cell_of((double)j0 * T.hy, T.cs, g.ny) - 1;
  int cy1 = (j0 == ncy - 1) ? g.ny - 1 : cell_of((double)(j0 + 1) * T.hy, T.cs, g.ny) + 1;
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

// The same gather over per-slot keys written by the f32c adjoint's pointwise
// pass: key[p] = (the scene's first theta cell) + j0 + i0*(ngy-1) of the slot's
// interp cell, so that a candidate costs one key compare; a match's FP64
// fractions (fx, fy) are formed from its position exactly as the f32c
// interp_cell forms them (x * 1/hx, clamped, minus the cell index); the
// contributions and their order are those of k_theta_cells.
// A workgroup takes a strip of kThetaStrip theta cells of one theta row
// (kThetaLanes lanes per cell) and first copies, for each SPH cell row the
// strip sweeps, the keys of the slot run that covers the SPH cell ranges of
// all its cells into LDS (coalesced); each cell then compares keys there and
// reads positions and galpha only for its own slots, kMatchCap matches at a
// time; the cell starts of the strip's rows are staged too, so that a cell's
// sweep issues no dependent global loads before its matches.  A strip whose
// runs exceed kThetaStage keys, that sweeps more than kThetaRows rows or more
// than kThetaCols - 1 SPH cells per row, reads keys and cell starts in global
// memory (same order, same sums).
// Workgroups are dealt to the scenes of a batched state in order
// (Scene32::tboff).
constexpr int kThetaStage = 2048;
constexpr int kThetaRows = 12;
constexpr int kThetaCols = 64;
constexpr int kMatchCap = 8;
constexpr int kMatchBatch = 4;

template <class GA>
__global__ void __launch_bounds__(kBlock)
k_theta_strip_keyed(const int *__restrict__ key, const double2 *__restrict__ X,
                    const GA *__restrict__ galpha, const int *__restrict__ cstart,
                    const int *__restrict__ tfar, int tstamp, Batch B, int first,
                    double4 *__restrict__ part) {
  static_assert(kThetaStrip * kLanes == kBlock, "one strip per workgroup");
  __shared__ int skey[kThetaStage];
  __shared__ int slo[kThetaRows], soff[kThetaRows + 1];
  __shared__ int scs[kThetaRows][kThetaCols];   // cstart of cells X0 .. X1+1 of each row
  __shared__ int sok, sx0;
  __shared__ int smatch[kMatchCap][kBlock];      // each lane's matched slots, in order
  const int blk = blockIdx.x;
  const Scene32 &Sc = B.sc[B.tb2s[blk]];
  const Grid g = Sc.g;
  const ThetaGrid &T = Sc.T;
  const int coff = Sc.coff, tcoff = Sc.tcoff;
  const int ncx = T.ngx - 1, ncy = T.ngy - 1;
  const int nstrip = (ncx + kThetaStrip - 1) / kThetaStrip;
  const int lb = blk - Sc.tboff;
  const int j0 = lb / nstrip;
  const int ifirst = (lb % nstrip) * kThetaStrip;
  const int ilast = min(ifirst + kThetaStrip, ncx) - 1;
  const int lane = threadIdx.x & (kLanes - 1);
  const int i0 = ifirst + (int)threadIdx.x / kLanes;
  // SPH cell ranges holding particles with interp cell (i, j0), tabulated per
  // scene: a quarter cell of margin (theta_cell_range_tight) unless the R pass
  // found a slot further out this step, else one cell (theta_cell_range)
  const int2 *tr = (*tfar == tstamp ? B.trange : B.trange_tight) + Sc.troff;
  auto xrange = [&](int i, int &cx0, int &cx1) {
    const int2 r = tr[i];
    cx0 = r.x;
    cx1 = r.y;
  };
  const int2 ry = tr[ncx + j0];
  const int cy0 = ry.x, cy1 = ry.y;
  const int nrow = cy1 - cy0 + 1;
  // plan: row k of the strip stages slots [lo_k, lo_k + len_k) at soff[k]
  if (threadIdx.x < 64) {
    const int k = (int)threadIdx.x;
    int X0, X1, dummy;
    xrange(ifirst, X0, dummy);
    xrange(ilast, dummy, X1);
    int lo = 0, len = 0;
    if (k < nrow && k < kThetaRows) {
      const int base = coff + (cy0 + k) * g.nx;
      lo = cstart[base + X0];
      len = cstart[base + X1 + 1] - lo;
    }
    int inc = len;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
      const int t = __shfl_up(inc, d, 64);
      if (k >= d) inc += t;
    }
    if (k < kThetaRows) {
      slo[k] = lo;
      soff[k] = inc - len;
    }
    const int total = __shfl(inc, kThetaRows - 1, 64);
    if (k == 0) {
      soff[kThetaRows] = total;
      sx0 = X0;
      sok = (nrow <= kThetaRows && total <= kThetaStage && X1 - X0 + 2 <= kThetaCols) ? 1 : 0;
    }
  }
  __syncthreads();
  const bool staged = sok != 0;
  if (staged) {
    // cell starts of the rows, then the keys; each thread issues the loads of
    // four elements before storing them
    const int X0 = sx0;
    const int ncol = g.nx - X0 + 1 < kThetaCols ? g.nx - X0 + 1 : kThetaCols;
    const int ncs = nrow * kThetaCols, total = soff[kThetaRows];
    int off[kThetaRows];
#pragma unroll
    for (int q = 0; q < kThetaRows; ++q) off[q] = soff[q];
    for (int i0 = (int)threadIdx.x; i0 < ncs + total; i0 += 4 * kBlock) {
      int v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = i0 + u * kBlock;
        v[u] = 0;
        if (i < ncs) {
          const int k = i / kThetaCols, c = i % kThetaCols;
          if (c < ncol) v[u] = cstart[coff + (cy0 + k) * g.nx + X0 + c];
        } else if (i < ncs + total) {
          const int t = i - ncs;
          int k = 0;
#pragma unroll
          for (int q = 1; q < kThetaRows; ++q) k += (t >= off[q]) ? 1 : 0;
          v[u] = key[slo[k] + (t - off[k])];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = i0 + u * kBlock;
        if (i < ncs) (&scs[0][0])[i] = v[u];
        else if (i < ncs + total) skey[i - ncs] = v[u];
      }
    }
    __syncthreads();
  }
  if (i0 > ilast) return;          // the kLanes lanes of a group leave together
  const int mykey = tcoff + j0 + i0 * ncy;
  const PhysF32 &P = Sc.P;
  auto frac = [&](double2 x) {
    double gx = x.x * P.inv_hx64, gy = x.y * P.inv_hy64;
    gx = (gx > P.gxclamp) ? P.gxclamp : ((gx < 0.0) ? 0.0 : gx);
    gy = (gy > P.gyclamp) ? P.gyclamp : ((gy < 0.0) ? 0.0 : gy);
    return make_double2(gx - (double)i0, gy - (double)j0);
  };
  int cx0, cx1;
  xrange(i0, cx0, cx1);
  double s00 = 0.0, s10 = 0.0, s01 = 0.0, s11 = 0.0;
  auto add = [&](double2 f, double ga) {
    s00 += (ga * (1.0 - f.x)) * (1.0 - f.y);
    s10 += (ga * f.x) * (1.0 - f.y);
    s01 += (ga * (1.0 - f.x)) * f.y;
    s11 += (ga * f.x) * f.y;
  };
  // the lane's matches in sweep order, collected up to kMatchCap at a time
  // (their positions, in LDS), then their data read in batches of
  // kMatchBatch independent loads; the sweep resumes where it stopped
  const int tid = (int)threadIdx.x;
  int cy = cy0, p = 0, e = 0, ro = 0;
  auto row = [&](int y) {   // start the sweep of SPH cell row y
    const int k = y - cy0;
    const int b = staged ? scs[k][cx0 - sx0] : cstart[coff + y * g.nx + cx0];
    e = staged ? scs[k][cx1 + 1 - sx0] : cstart[coff + y * g.nx + cx1 + 1];
    ro = staged ? soff[k] - slo[k] : 0;
    p = b + lane;
  };
  row(cy);
  for (;;) {
    int n = 0;
    while (n < kMatchCap) {
      if (p >= e) {
        if (++cy > cy1) break;
        row(cy);
        continue;
      }
      if ((staged ? skey[p + ro] : key[p]) == mykey) smatch[n++][tid] = p;
      p += kLanes;
    }
    for (int m0 = 0; m0 < n; m0 += kMatchBatch) {
      double2 f[kMatchBatch];
      double ga[kMatchBatch];
#pragma unroll
      for (int q = 0; q < kMatchBatch; ++q) {
        if (m0 + q < n) {
          const int pm = smatch[m0 + q][tid];
          f[q] = X[pm];
          ga[q] = (double)galpha[pm];
        }
      }
#pragma unroll
      for (int q = 0; q < kMatchBatch; ++q)
        if (m0 + q < n) add(frac(f[q]), ga[q]);
    }
    if (n < kMatchCap) break;   // the sweep has ended
  }
#pragma unroll
  for (int m = kLanes / 2; m > 0; m >>= 1) {
    s00 += __shfl_xor(s00, m, kLanes);
    s10 += __shfl_xor(s10, m, kLanes);
    s01 += __shfl_xor(s01, m, kLanes);
    s11 += __shfl_xor(s11, m, kLanes);
  }
  if (lane == 0) {
    double4 &pt = part[tcoff + j0 + i0 * ncy];
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
                            const int *cstart, Grid g, ThetaGrid T, bool first, double4 *part) {
  const long long ncells = (long long)(T.ngx - 1) * (T.ngy - 1);
  k_theta_cells<double><<<nblocks(ncells * kLanes), kBlock, 0, s>>>(X, galpha, cstart, g, T,
                                                                    first ? 1 : 0, part);
}

void launch_theta_nodes_f64(hipStream_t s, const double4 *part, ThetaGrid T, double *gtheta) {
  k_theta_nodes<<<nblocks((long long)T.ngx * T.ngy), kBlock, 0, s>>>(part, T, gtheta);
}

void launch_theta_gather_keyed_f32(hipStream_t s, const int *key, const double2 *X,
                                   const float *galpha, const int *cstart, const int *tfar,
                                   int tstamp, const Batch &B, bool first, double4 *part) {
  static_assert(kLanes == kThetaLanes, "theta-cell launch geometry");
  k_theta_strip_keyed<float><<<B.ntblk, kBlock, 0, s>>>(key, X, galpha, cstart, tfar, tstamp, B,
                                                        first ? 1 : 0, part);
}

void launch_theta_nodes_keyed_f32(hipStream_t s, const double4 *part, const Batch &B,
                                  double *gtheta) {
  k_theta_nodes_b<<<B.nnblk, kBlock, 0, s>>>(part, B, gtheta);
}

}  // namespace sph
