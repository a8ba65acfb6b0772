// sph_theta.hip -- deterministic gtheta reduction (no float atomics): the FP64
// form's gather and the node sums of both forms (the f32c form's keyed gather
// is in sph_f32c.hip, where P1 runs it in its own launch).
//
// The reference scatters each particle's four bilinear weights into gtheta
// with float atomics (adj_design_kernel!, src/adjoint.jl:166-185), whose
// summation order changes from run to run.  Here the scatter becomes two
// gathers in a fixed order:
//   A  eight lanes per theta cell (i0, j0): sweep the SPH cells that can hold
//      particles whose interp_cell is (i0, j0) (plus a margin for the
//      displacement since the last rebuild), keep the candidates whose
//      interp_cell (stored by the R pass) is (i0, j0), recompute their
//      fractions exactly as the reference does, and sum the four weighted
//      contributions
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

constexpr int kLanes = 8;   // lanes that share one theta cell

// Gather A.  Each slot's key is its interp cell l00 = j0 + i0*ngy, stored by
// the step's R pass (k_adj_r / k_adj_rpw, interp_cell of src/params.jl:53-68),
// so that a candidate costs one key compare; a match's fractions are
// recomputed from its position by the same rule (x / hx, clamped, minus the
// cell index).  The gather is bound by dependent memory rounds, so a lane
// reads the cell starts of kKeyRows rows at once, then the keys of up to
// kKeyBatch of its candidates over those rows together, then the positions
// and galpha of the matches among them kHitBatch at a time, adding them in
// candidate order.  (Measured on the 1M block, FP64 form: 32.6 us a step for
// a version that loaded every candidate's position and recomputed its interp
// cell over the one-cell ranges; 30.1 us over the keys, with bit-identical
// sums; 19.8 us over the quarter-cell ranges.  4 lanes a cell: 21.5 us,
// 16 lanes: 23.5 us.)
constexpr int kKeyRows = 4;
constexpr int kKeyBatch = 16;
constexpr int kHitBatch = 4;
__global__ void __launch_bounds__(kBlock)
k_theta_cells_keyed(const double2 *__restrict__ X, const double *__restrict__ galpha,
                    const int *__restrict__ key, const int *__restrict__ cstart,
                    const int2 *__restrict__ trange, Grid g, ThetaGrid T, int first,
                    double4 *__restrict__ part, const unsigned long long *__restrict__ lev,
                    int stamp) {
  const int ncx = T.ngx - 1, ncy = T.ngy - 1;
  const int lane = threadIdx.x & (kLanes - 1);
  // theta cells in x-fastest order, so that neighbouring groups read
  // neighbouring slot runs (and, through xcd_block, stay on one XCD)
  const int t = (xcd_block() * blockDim.x + threadIdx.x) / kLanes;
  if (t >= ncx * ncy) return;      // the kLanes lanes of a group leave together
  const int i0 = t % ncx;
  const int j0 = t / ncx;
  // SPH cell range holding particles with interp cell (i0, j0): a quarter cell
  // of margin unless the step's R pass found a slot further out (level 1: one
  // cell; level m >= 2: the one-cell range widened by m - 1 cells a side)
  const int lv = level_of(lev, stamp);
  const int2 *tr = lv >= 1 ? trange : trange + (ncx + ncy);
  const int wid = lv > 1 ? lv - 1 : 0;
  const int2 rxw = tr[i0], ryw = tr[ncx + j0];
  const int cx0 = max(rxw.x - wid, 0), cx1 = min(rxw.y + wid, g.nx - 1);
  const int cy0 = max(ryw.x - wid, 0), cy1 = min(ryw.y + wid, g.ny - 1);
  const int mykey = j0 + i0 * T.ngy;
  double s00 = 0.0, s10 = 0.0, s01 = 0.0, s11 = 0.0;
  // the lane's candidates of kKeyRows rows, in sweep order: candidate u of
  // row q is slot rb[q] + kLanes u (u < its count), numbered k = pre[q] + u
  for (int ya = cy0; ya <= cy1; ya += kKeyRows) {
    int rb[kKeyRows], pre[kKeyRows];
    int total = 0;
#pragma unroll
    for (int q = 0; q < kKeyRows; ++q) {
      const int y = ya + q;
      int b = 0, e = 0;
      if (y <= cy1) {
        b = cstart[y * g.nx + cx0];
        e = cstart[y * g.nx + cx1 + 1];
      }
      rb[q] = b + lane;
      pre[q] = total;
      total += e > rb[q] ? (e - rb[q] + kLanes - 1) / kLanes : 0;
    }
    auto slot_of = [&](int k) {   // slot of candidate k < total
      int q = 0;
#pragma unroll
      for (int r = 1; r < kKeyRows; ++r) q += (k >= pre[r]) ? 1 : 0;
      return rb[q] + kLanes * (k - pre[q]);
    };
    for (int k0 = 0; k0 < total; k0 += kKeyBatch) {
      int kv[kKeyBatch];
#pragma unroll
      for (int u = 0; u < kKeyBatch; ++u) kv[u] = k0 + u < total ? key[slot_of(k0 + u)] : -1;
      uint32_t hits = 0;
#pragma unroll
      for (int u = 0; u < kKeyBatch; ++u) hits |= (kv[u] == mykey ? 1u : 0u) << u;
      while (hits != 0) {
        bool ok[kHitBatch];
        double2 x[kHitBatch];
        double ga[kHitBatch];
#pragma unroll
        for (int u = 0; u < kHitBatch; ++u) {
          ok[u] = hits != 0;
          if (ok[u]) {
            const int p = slot_of(k0 + __builtin_ctz(hits));
            hits &= hits - 1;
            x[u] = X[p];
            ga[u] = galpha[p];
          }
        }
#pragma unroll
        for (int u = 0; u < kHitBatch; ++u) {
          if (!ok[u]) continue;
          // the fractions of interp_cell (src/params.jl:53-68), exactly
          const double gx = x[u].x / T.hx, gy = x[u].y / T.hy;
          const double gxc = (gx > T.gxclamp) ? T.gxclamp : ((gx < 0.0) ? 0.0 : gx);
          const double gyc = (gy > T.gyclamp) ? T.gyclamp : ((gy < 0.0) ? 0.0 : gy);
          const double fx = gxc - (double)i0;
          const double fy = gyc - (double)j0;
          s00 += (ga[u] * (1.0 - fx)) * (1.0 - fy);
          s10 += (ga[u] * fx) * (1.0 - fy);
          s01 += (ga[u] * (1.0 - fx)) * fy;
          s11 += (ga[u] * fx) * fy;
        }
      }
    }
  }
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

void launch_theta_cells_keyed_f64(hipStream_t s, const double2 *X, const double *galpha,
                                  const int *key, const int *cstart, const int2 *trange, Grid g,
                                  ThetaGrid T, bool first, double4 *part,
                                  const unsigned long long *lev, int stamp) {
  const long long ncells = (long long)(T.ngx - 1) * (T.ngy - 1);
  k_theta_cells_keyed<<<nblocks(ncells * kLanes), kBlock, 0, s>>>(
      X, galpha, key, cstart, trange, g, T, first ? 1 : 0, part, lev, stamp);
}

void launch_theta_nodes_f64(hipStream_t s, const double4 *part, ThetaGrid T, double *gtheta) {
  k_theta_nodes<<<nblocks((long long)T.ngx * T.ngy), kBlock, 0, s>>>(part, T, gtheta);
}

void launch_theta_nodes_keyed_f32(hipStream_t s, const double4 *part, const Batch &B,
                                  double *gtheta) {
  k_theta_nodes_b<<<B.nnblk, kBlock, 0, s>>>(part, B, gtheta);
}

}  // namespace sph
