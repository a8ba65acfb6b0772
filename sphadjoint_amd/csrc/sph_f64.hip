// sph_f64.hip -- plain FP64 physics kernels for gfx950.
//
// Forward (src/forward.jl) and gather-form adjoint (src/adjoint.jl) over the
// slot layout.  As in the reference, each rebuild materialises a transposed
// fixed-row neighbour list (src/neighbors.jl:116-245): entry m of slot s is at
// nl[m*N + s], so at a given m all lanes of a wavefront read consecutive
// words.  The list holds every slot within 2h(1+skin) at build time, in the
// reference's order (ddy outer, ddx inner, ascending particle id inside a
// cell); the pair kernels keep the pairs with r^2 < (2h)^2 at the current
// positions.  Every sum is therefore the reference's, term for term and in the
// same order.  This file is compiled with -ffp-contract=off: a*b+c stays two
// rounded operations as in Julia, so the FP64 results equal the reference's
// CPU Float64 arithmetic.
//
// Kernels: list build (per rebuild); forward density (pterm, invrho) and
// force + implicit-drag integrate; adjoint R (density recompute, integrate
// and drag adjoints), P1 (pair adjoint and design term), P2 (density adjoint,
// gX += gXs).  gtheta is the deterministic gather of sph_theta.hip.  The pair
// passes copy the neighbour rows of their window of slots into LDS first
// (sph_stage.h) and read the list's neighbours there.
#pragma clang fp contract(off)
#include "sph_internal.h"
#include "sph_stage.h"

namespace sph {

namespace {

inline int nblocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

__device__ __forceinline__ double wall_accel(double x, double L, double kw) {
  double lo = (-x > 0.0) ? -x : 0.0;
  double hi = (x - L > 0.0) ? x - L : 0.0;
  return kw * (lo * lo - hi * hi);
}
__device__ __forceinline__ double wall_daccel(double x, double L, double kw) {
  double lo = (-x > 0.0) ? -x : 0.0;
  double hi = (x - L > 0.0) ? x - L : 0.0;
  return kw * (-2.0 * lo - 2.0 * hi);
}

struct Interp {
  int l00;            // 0-based linear index of theta[j, i] (column-major, ny rows)
  double fx, fy, ax, ay;
};

// interp_cell (src/params.jl:53-68)
__device__ __forceinline__ Interp interp_cell(double x, double y, const PhysF64 &P) {
  Interp r;
  double gx = x / P.hx;
  double gy = y / P.hy;
  r.ax = (gx > 0.0 && gx < P.gxmax) ? 1.0 : 0.0;
  r.ay = (gy > 0.0 && gy < P.gymax) ? 1.0 : 0.0;
  double gxc = (gx > P.gxclamp) ? P.gxclamp : ((gx < 0.0) ? 0.0 : gx);
  double gyc = (gy > P.gyclamp) ? P.gyclamp : ((gy < 0.0) ? 0.0 : gy);
  double fxf = floor(gxc), fyf = floor(gyc);
  long long i0 = (fxf > -1e15 && fxf < 1e15) ? (long long)fxf : -(1LL << 60);
  long long j0 = (fyf > -1e15 && fyf < 1e15) ? (long long)fyf : -(1LL << 60);
  i0 = i0 < 0 ? 0 : (i0 > P.ngx - 2 ? P.ngx - 2 : i0);
  j0 = j0 < 0 ? 0 : (j0 > P.ngy - 2 ? P.ngy - 2 : j0);
  r.fx = gxc - (double)i0;
  r.fy = gyc - (double)j0;
  r.l00 = (int)(j0 + i0 * P.ngy);
  return r;
}

// interp_alpha (src/params.jl:70-82)
__device__ __forceinline__ double interp_alpha(const double *__restrict__ th, const Interp &c,
                                               const PhysF64 &P, double *dadx, double *dady) {
  double t00 = th[c.l00];
  double t01 = th[c.l00 + P.ngy];
  double t10 = th[c.l00 + 1];
  double t11 = th[c.l00 + P.ngy + 1];
  double omx = 1.0 - c.fx, omy = 1.0 - c.fy;
  double a = (((omx * omy) * t00 + (c.fx * omy) * t01) + (omx * c.fy) * t10) + (c.fx * c.fy) * t11;
  if (dadx) {
    *dadx = ((omy * (t01 - t00) + c.fy * (t11 - t10)) / P.hx) * c.ax;
    *dady = ((omx * (t10 - t00) + c.fx * (t11 - t01)) / P.hy) * c.ay;
  }
  return a;
}

// ---- neighbour reads of the pair passes ------------------------------------
// Each pass walks the slot's list in order and reads the neighbours' data
// either from the state arrays or from the LDS rows of its window
// (sph_stage.h); both give the same values, so the sums are the same.
//
// kPairMode, bit 0: stage the window's rows in LDS; bit 1: fetch the list
// entries kAhead at a time, one group ahead of their use.  On a 1M-particle
// block staging made every FP64 pass slower (force 95 -> 146 us, density
// 49 -> 56 us: the L1-cached gathers are as fast as the LDS reads, and the
// copy, the index map and the lower occupancy are extra), while fetching
// ahead sped P1 up by 12 %; the default is 2 (build with
// -DSPH_F64_PAIR_MODE=3 for the staged passes).
#ifndef SPH_F64_PAIR_MODE
#define SPH_F64_PAIR_MODE 2
#endif
#ifndef SPH_F64_AHEAD
#define SPH_F64_AHEAD 4
#endif
constexpr int kPairMode = SPH_F64_PAIR_MODE;
constexpr int kAhead = SPH_F64_AHEAD;
// staging capacity in slots (a 256-slot window of a fluid block with ~8 slots
// per cell stages ~830 slots; windows over two cell rows ~1100)
constexpr int kStageF64 = 1024;

// LDS position of list entry j of a slot whose own cell row is segment row k
// of the window: the sweep's three rows ascend in slot order, so the row of
// j is (j >= lo[k+1]) + (j >= lo[k+2]) (lo = INT_MAX for a row outside the
// grid), and its position is j minus that segment's delta.
struct StagedIdx {
  int t1, t2, d0, d1, d2;
  __device__ __forceinline__ int operator()(int j) const {
    return j - (j >= t2 ? d2 : (j >= t1 ? d1 : d0));
  }
};
__device__ __forceinline__ StagedIdx staged_idx(const StageTab &T, int bc, const Grid &g) {
  const int k = (cell_xy(bc, g).y - T.ya) * 3;
  return StagedIdx{T.lo[k + 1], T.lo[k + 2], T.delta[k], T.delta[k + 1], T.delta[k + 2]};
}
struct SlotIdx {
  __device__ __forceinline__ int operator()(int j) const { return j; }
};

// Stages the rows of the window starting at slot w0 (st(i, ld(j)) stores
// slot j's data at LDS position i); every thread of the workgroup calls it.
// True when the calling wave reads from LDS: the window fits, and every slot
// of the wave sweeps around its binning cell (a slot's sweep cell, from
// x*(1/cs), and its binning cell, from x/cs, differ only where the two
// roundings fall on either side of a cell boundary; such a wave reads from
// global memory).
template <class Ld, class St>
__device__ __forceinline__ bool stage_rows(StageTab &T, const EpochCells &ec, int w0,
                                           const Grid &g, Ld &&ld, St &&st) {
  stage_plan_at(T, ec.bcell, [&](int i) { return ec.cstart[i]; }, w0, min(w0 + kBlock, g.N), g,
                0, kStageF64);
  if (!T.ok) return false;
  stage_copy(T, ld, st);
  const int s = w0 + (int)threadIdx.x;
  return __all(s >= g.N || ec.scell[s] == ec.bcell[s]);
}

// fn(j) for the entries of row in order; with kAhead the entries of the next
// group are loaded before the current group is evaluated
template <bool kFetchAhead, class Fn>
__device__ __forceinline__ void nl_walk(const NbRow &row, Fn &&fn) {
  if constexpr (!kFetchAhead) {
    for (int m = 0; m < row.n; ++m) fn(row.at(m));
  } else {
    int cur[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) cur[u] = u < row.n ? row.at(u) : 0;
    for (int m = 0; m < row.n; m += kAhead) {
      int nxt[kAhead];
#pragma unroll
      for (int u = 0; u < kAhead; ++u)
        nxt[u] = m + kAhead + u < row.n ? row.at(m + kAhead + u) : 0;
#pragma unroll
      for (int u = 0; u < kAhead; ++u)
        if (m + u < row.n) fn(cur[u]);
#pragma unroll
      for (int u = 0; u < kAhead; ++u) cur[u] = nxt[u];
    }
  }
}

// density sum of slot s (density_kernel!, src/forward.jl:25-65); xs[idx(j)]
// is the position of list entry j
template <class Idx>
__device__ __forceinline__ double density_at(const double2 *__restrict__ xs, Idx idx, double2 xi,
                                             const NbRow &row, const PhysF64 &P) {
  double acc = 0.0;
  nl_walk<(kPairMode & 2) != 0>(row, [&](int j) {
    double2 xj = xs[idx(j)];
    double dx = xi.x - xj.x;
    double dy = xi.y - xj.y;
    double r2 = dx * dx + dy * dy;
    if (r2 < P.r2max) {
      double q = sqrt(r2) * P.invh;
      double u = 1.0 - 0.5 * q;
      double u2 = u * u;
      acc += P.m * (((P.A * u2) * u2) * (2.0 * q + 1.0));
    }
  });
  return acc + P.m * P.A;
}

// the density of slot s from the window's LDS rows sx when staged, else from X
__device__ __forceinline__ double density_of(const double2 *__restrict__ X, const double2 *sx,
                                             bool lds, const StageTab &T, const EpochCells &ec,
                                             int s, double2 xi, const NbRow &row, const Grid &g,
                                             const PhysF64 &P) {
  if ((kPairMode & 1) && lds) return density_at(sx, staged_idx(T, ec.bcell[s], g), xi, row, P);
  return density_at(X, SlotIdx{}, xi, row, P);
}

// _nl_build! (src/neighbors.jl:192-245) into slot space: the candidates of
// the 3x3 block around the sweep cell, kept when r2min < r2 < rc2.
// Each lane keeps its entries in LDS as 16-bit offsets into the row run they
// come from (the sweep visits the three runs in order, so per-row counts
// tell the rows apart), and the wavefront writes them out entry by entry:
// at each m the lanes store consecutive words of nl[m*N + .], instead of
// every lane storing into its own line at its own m during the sweep.
// Lists longer than kNlLds entries (maxnb > kNlLds) and offsets past 16 bits
// (a row run of more than 65,535 slots) are stored directly.
constexpr int kNlLds = 64;
__global__ void __launch_bounds__(kBlock)
k_nl_build(const double2 *__restrict__ X, const int *__restrict__ scell,
           const int *__restrict__ cstart, Grid g, CellConstF64 k, int maxnb,
           int *__restrict__ nl, int *__restrict__ nbcnt, int *__restrict__ flags) {
  __shared__ unsigned short ent[kNlLds][kBlock];
  const int tid = (int)threadIdx.x;
  int s = xcd_block() * blockDim.x + tid;
  if (s >= g.N) return;   // (no barrier below: each lane uses its own LDS column)
  const double2 xi = X[s];
  Runs r = sweep_runs(cstart, scell[s], g.nx, g.ny);
  int m = 0, mlds = 0;        // entries, and how many of them sit in LDS (a prefix)
  int cnt[3] = {0, 0, 0};     // LDS entries from each row run
  bool lds = maxnb <= kNlLds, over = false;
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    for (int j = r.b[w]; j < r.e[w]; ++j) {
      double2 xj = X[j];
      double dx = xi.x - xj.x;
      double dy = xi.y - xj.y;
      double r2 = dx * dx + dy * dy;
      if (k.r2min < r2 && r2 < k.rc2) {
        if (m < maxnb) {
          const int b = j - r.b[w];
          if (lds && b < 65536) {
            ent[m][tid] = (unsigned short)b;
            cnt[w] += 1;
            mlds = m + 1;
          } else {
            lds = false;
            nl[(size_t)m * g.N + s] = j;
          }
          ++m;
        } else {
          over = true;
        }
      }
    }
  }
  for (int q = 0; q < kNlLds; ++q) {
    if (!__any(q < mlds)) break;
    if (q < mlds) {
      const int base = q < cnt[0] ? r.b[0] : (q < cnt[0] + cnt[1] ? r.b[1] : r.b[2]);
      nl[(size_t)q * g.N + s] = base + (int)ent[q][tid];
    }
  }
  nbcnt[s] = m;
  if (over) flags[0] = 1;
}

// (the staged passes keep one array per neighbour quantity, of kStageF64
// slots when staging is on)
constexpr int kStageArr = (kPairMode & 1) ? kStageF64 : 1;

__global__ void __launch_bounds__(kBlock)
k_density(const double2 *__restrict__ X, const int *__restrict__ nl,
          const int *__restrict__ nbcnt, Grid g, PhysF64 P, EpochCells ec,
          double2 *__restrict__ PI, double *__restrict__ rho_out, double2 *__restrict__ tapeX) {
  __shared__ StageTab T;
  __shared__ double2 sx[kStageArr];
  const int w0 = xcd_block() * kBlock, s = w0 + (int)threadIdx.x;
  bool lds = false;
  if constexpr ((kPairMode & 1) != 0)
    lds = stage_rows(T, ec, w0, g, [&](int j) { return X[j]; },
                     [&](int i, double2 d) { sx[i] = d; });
  if (s >= g.N) return;
  double2 xi = X[s];
  double rho = density_of(X, sx, lds, T, ec, s, xi, nb_row(nl, nbcnt, s, g.N), g, P);
  // eos_kernel! (src/forward.jl:68-75)
  double inv = 1.0 / rho;
  PI[s] = make_double2(((P.c2 * (rho - P.rho0)) * inv) * inv, inv);
  if (rho_out) rho_out[s] = rho;
  if (tapeX) tapeX[s] = xi;
}

// accel_kernel! + integrate_kernel! (src/forward.jl:77-147), fused: the new
// state goes to separate buffers because neighbours still read the old one.
// pair sum of accel_kernel! (src/forward.jl:77-130) for slot s: xs, pis, vs
// at idx(j) hold the position, (pterm, invrho) and velocity of entry j
template <class Idx>
__device__ __forceinline__ double2 force_sum(const double2 *__restrict__ xs,
                                             const double2 *__restrict__ pis,
                                             const double2 *__restrict__ vs, Idx idx, double2 xi,
                                             double2 vi, double pti, double iri,
                                             const NbRow &row, const PhysF64 &P) {
  double ax = 0.0, ay = 0.0;
  nl_walk<(kPairMode & 2) != 0>(row, [&](int j) {
    const int i = idx(j);
    double2 xj = xs[i];
    double dx = xi.x - xj.x;
    double dy = xi.y - xj.y;
    double r2 = dx * dx + dy * dy;
    if (r2 < P.r2max) {
      double q = sqrt(r2) * P.invh;
      double u = 1.0 - 0.5 * q;
      double F = ((P.Fc * u) * u) * u;
      double2 pij = pis[i];
      double Pij = P.m * (pti + pij.x);
      ax -= (Pij * F) * dx;
      ay -= (Pij * F) * dy;
      double C = ((P.twomumass * iri) * pij.y) * F;
      double2 vj = vs[i];
      ax += C * (vi.x - vj.x);
      ay += C * (vi.y - vj.y);
    }
  });
  return make_double2(ax, ay);
}

struct NbXPV64 {
  double2 x, pi, v;
};

__global__ void __launch_bounds__(kBlock)
k_force(const double2 *__restrict__ X, const double2 *__restrict__ V,
        const double2 *__restrict__ PI, const double *__restrict__ theta,
        const int *__restrict__ nl, const int *__restrict__ nbcnt, Grid g, PhysF64 P,
        EpochCells ec, double2 *__restrict__ Xn, double2 *__restrict__ Vn,
        double2 *__restrict__ tapeV, double2 *__restrict__ a_out,
        double *__restrict__ alpha_out) {
  __shared__ StageTab T;
  __shared__ double2 sx[kStageArr], spi[kStageArr], sv[kStageArr];
  const int w0 = xcd_block() * kBlock, s = w0 + (int)threadIdx.x;
  bool lds = false;
  if constexpr ((kPairMode & 1) != 0)
    lds = stage_rows(T, ec, w0, g, [&](int j) { return NbXPV64{X[j], PI[j], V[j]}; },
                     [&](int i, const NbXPV64 &d) {
                       sx[i] = d.x;
                       spi[i] = d.pi;
                       sv[i] = d.v;
                     });
  if (s >= g.N) return;
  double2 xi = X[s], vi = V[s], pii = PI[s];
  double pti = pii.x, iri = pii.y;
  const NbRow row = nb_row(nl, nbcnt, s, g.N);
  double2 a2 = ((kPairMode & 1) && lds)
                   ? force_sum(sx, spi, sv, staged_idx(T, ec.bcell[s], g), xi, vi, pti, iri, row, P)
                   : force_sum(X, PI, V, SlotIdx{}, xi, vi, pti, iri, row, P);
  double ax = a2.x, ay = a2.y;
  Interp ic = interp_cell(xi.x, xi.y, P);
  double al = interp_alpha(theta, ic, P, nullptr, nullptr);
  ax += P.gx + wall_accel(xi.x, P.Lx, P.kw);
  ay += P.gy + wall_accel(xi.y, P.Ly, P.kw);
  if (a_out) {  // auxiliary recompute of st.a / st.alpha, no integration
    a_out[s] = make_double2(ax, ay);
    alpha_out[s] = al;
    return;
  }
  double D = 1.0 / (1.0 + P.dt * al);
  double v1 = (vi.x + P.dt * ax) * D;
  double v2 = (vi.y + P.dt * ay) * D;
  Vn[s] = make_double2(v1, v2);
  Xn[s] = make_double2(xi.x + P.dt * v1, xi.y + P.dt * v2);
  if (tapeV) tapeV[s] = vi;
}

// The step's gtheta gather (k_theta_cells_keyed) sweeps the quarter-cell SPH
// ranges of each theta cell at level 0, the one-cell ranges at level 1 and
// those widened by m - 1 cells at level m >= 2.  Its level must reach slot s,
// whose binning cell at the epoch's rebuild is L.bcell[s] and whose interp
// cell is l00: raise it when s lies beyond the quarter-cell range (a slot that
// moved more than a quarter cell past its binning cell since the rebuild).
// The interp cell is also stored as the slot's key.
__device__ __forceinline__ void sweep_level(const LevelArgs &L, const Grid &g, int s, int l00) {
  L.key[s] = l00;
  const int i0 = l00 / L.ngy, j0 = l00 - i0 * L.ngy;
  const int c = L.bcell[s];
  const int cy = c / g.nx, cx = c - cy * g.nx;
  const int ncy = L.ngy - 1;
  const int2 *tt = L.trange + (L.ncx + ncy);
  const int2 rx = tt[i0], ry = tt[L.ncx + j0];
  if (cx < rx.x || cx > rx.y || cy < ry.x || cy > ry.y) {
    const int m = max(margin_beyond(cx, L.trange[i0]), margin_beyond(cy, L.trange[L.ncx + j0]));
    raise_level(L.lev, L.stamp, m > 1 ? m : 1);
  }
}

// Adjoint pass R: density recompute, _adj_integrate! and _adj_drag!
// (src/adjoint.jl:193-228).  v' is the taped V_{n+1}, which is bit-identical
// to the (V + dt a) D that the reference recomputes.
__global__ void __launch_bounds__(kBlock)
k_adj_r(const double2 *__restrict__ X, const double2 *__restrict__ Vnext,
        const double *__restrict__ theta, const int *__restrict__ nl,
        const int *__restrict__ nbcnt, Grid g, PhysF64 P, EpochCells ec,
        double2 *__restrict__ PI, double2 *__restrict__ gX, double2 *__restrict__ gV,
        double2 *__restrict__ abar, double *__restrict__ galpha, LevelArgs L) {
  __shared__ StageTab T;
  __shared__ double2 sx[kStageArr];
  const int w0 = xcd_block() * kBlock, s = w0 + (int)threadIdx.x;
  bool lds = false;
  if constexpr ((kPairMode & 1) != 0)
    lds = stage_rows(T, ec, w0, g, [&](int j) { return X[j]; },
                     [&](int i, double2 d) { sx[i] = d; });
  if (s >= g.N) return;
  double2 xi = X[s];
  double rho = density_of(X, sx, lds, T, ec, s, xi, nb_row(nl, nbcnt, s, g.N), g, P);
  double inv = 1.0 / rho;
  PI[s] = make_double2(((P.c2 * (rho - P.rho0)) * inv) * inv, inv);
  Interp ic = interp_cell(xi.x, xi.y, P);
  sweep_level(L, g, s, ic.l00);
  double al = interp_alpha(theta, ic, P, nullptr, nullptr);
  double2 gx = gX[s], gv = gV[s], vp = Vnext[s];
  double gvx = gv.x + P.dt * gx.x;
  double gvy = gv.y + P.dt * gx.y;
  double D = 1.0 / (1.0 + P.dt * al);
  abar[s] = make_double2((P.dt * D) * gvx, (P.dt * D) * gvy);
  galpha[s] = ((-P.dt) * D) * (vp.x * gvx + vp.y * gvy);
  gV[s] = make_double2(D * gvx, D * gvy);
}

// Adjoint pass R without the density recompute, for tapes that kept each
// step's (pterm, invrho): the same _adj_integrate! and _adj_drag! arithmetic.
__global__ void __launch_bounds__(kBlock)
k_adj_rpw(const double2 *__restrict__ X, const double2 *__restrict__ Vnext,
          const double *__restrict__ theta, Grid g, PhysF64 P, double2 *__restrict__ gX,
          double2 *__restrict__ gV, double2 *__restrict__ abar, double *__restrict__ galpha,
          LevelArgs L) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.N) return;
  double2 xi = X[s];
  Interp ic = interp_cell(xi.x, xi.y, P);
  sweep_level(L, g, s, ic.l00);
  double al = interp_alpha(theta, ic, P, nullptr, nullptr);
  double2 gx = gX[s], gv = gV[s], vp = Vnext[s];
  double gvx = gv.x + P.dt * gx.x;
  double gvy = gv.y + P.dt * gx.y;
  double D = 1.0 / (1.0 + P.dt * al);
  abar[s] = make_double2((P.dt * D) * gvx, (P.dt * D) * gvy);
  galpha[s] = ((-P.dt) * D) * (vp.x * gvx + vp.y * gvy);
  gV[s] = make_double2(D * gvx, D * gvy);
}

// Adjoint pass P1: adj_pass1_kernel! (src/adjoint.jl:30-122) followed by the
// design term of adj_design_kernel! (src/adjoint.jl:166-185) and gV += gVs.
// pair sums of adj_pass1_kernel! (src/adjoint.jl:30-122) for slot s: xs,
// pis, abs, vs at idx(j) hold the position, (pterm, invrho), abar and
// velocity of entry j
struct P1Sums {
  double gx, gy, gvx, gvy, grh, sumS;
};
template <class Idx>
__device__ __forceinline__ P1Sums p1_sums(const double2 *__restrict__ xs,
                                          const double2 *__restrict__ pis,
                                          const double2 *__restrict__ abs,
                                          const double2 *__restrict__ vs, Idx idx, double2 xi,
                                          double2 vi, double2 ab, double pti, double iri,
                                          const NbRow &row, const PhysF64 &P) {
  double gx = 0.0, gy = 0.0, gvx = 0.0, gvy = 0.0, grh = 0.0, sumS = 0.0;
  nl_walk<(kPairMode & 2) != 0>(row, [&](int j) {
    const int i = idx(j);
    double2 xj = xs[i];
    double dx = xi.x - xj.x;
    double dy = xi.y - xj.y;
    double r2 = dx * dx + dy * dy;
    if (r2 < P.r2max) {
      double rr = sqrt(r2);
      double q = rr * P.invh;
      double u = 1.0 - 0.5 * q;
      double F = ((P.Fc * u) * u) * u;
      double G = (P.Gc * u) * u;
      double2 pij = pis[i];
      double Pij = P.m * (pti + pij.x);
      double Cbase = (P.twomumass * iri) * pij.y;
      double C = Cbase * F;
      double2 abj = abs[i];
      double2 vj = vs[i];
      double dax = ab.x - abj.x;
      double day = ab.y - abj.y;
      double dvx = vi.x - vj.x;
      double dvy = vi.y - vj.y;
      double ad_d = -(dax * dx + day * dy);
      double ad_v = dax * dvx + day * dvy;
      gx -= (Pij * F) * dax;
      gy -= (Pij * F) * day;
      double Fbar = Pij * ad_d + Cbase * ad_v;
      double wgt = ((Fbar * G) * P.invh) / rr;
      gx += wgt * dx;
      gy += wgt * dy;
      gvx += C * dax;
      gvy += C * day;
      grh -= (C * ad_v) * iri;
      sumS += F * ad_d;
    }
  });
  return P1Sums{gx, gy, gvx, gvy, grh, sumS};
}

struct NbXPAV64 {
  double2 x, pi, ab, v;
};

__global__ void __launch_bounds__(kBlock)
k_adj_p1(const double2 *__restrict__ X, const double2 *__restrict__ V,
         const double2 *__restrict__ PI, const double2 *__restrict__ abar,
         const double *__restrict__ galpha, const double *__restrict__ theta,
         const int *__restrict__ nl, const int *__restrict__ nbcnt, Grid g, PhysF64 P,
         EpochCells ec, double2 *__restrict__ gXs, double2 *__restrict__ gV,
         double *__restrict__ grho) {
  __shared__ StageTab T;
  __shared__ double2 sx[kStageArr], spi[kStageArr], sab[kStageArr], sv[kStageArr];
  const int w0 = xcd_block() * kBlock, s = w0 + (int)threadIdx.x;
  bool lds = false;
  if constexpr ((kPairMode & 1) != 0)
    lds = stage_rows(T, ec, w0, g, [&](int j) { return NbXPAV64{X[j], PI[j], abar[j], V[j]}; },
                     [&](int i, const NbXPAV64 &d) {
                       sx[i] = d.x;
                       spi[i] = d.pi;
                       sab[i] = d.ab;
                       sv[i] = d.v;
                     });
  if (s >= g.N) return;
  double2 xi = X[s], vi = V[s], ab = abar[s], pii = PI[s];
  double pti = pii.x, iri = pii.y;
  const NbRow row = nb_row(nl, nbcnt, s, g.N);
  const P1Sums q = ((kPairMode & 1) && lds)
                       ? p1_sums(sx, spi, sab, sv, staged_idx(T, ec.bcell[s], g), xi, vi, ab, pti,
                                 iri, row, P)
                       : p1_sums(X, PI, abar, V, SlotIdx{}, xi, vi, ab, pti, iri, row, P);
  double gx = q.gx, gy = q.gy, gvx = q.gvx, gvy = q.gvy, grh = q.grh, sumS = q.sumS;
  grh += (((-2.0 * P.m) * pti) * iri) * sumS;
  grh += (((P.c2 * P.m) * iri) * iri) * sumS;
  gx += ab.x * wall_daccel(xi.x, P.Lx, P.kw);
  gy += ab.y * wall_daccel(xi.y, P.Ly, P.kw);
  // adj_design_kernel!: gXs += galpha * grad(alpha); its gtheta scatter is the
  // deterministic gather of sph_theta.hip
  double ga = galpha[s];
  Interp ic = interp_cell(xi.x, xi.y, P);
  double dadx, dady;
  interp_alpha(theta, ic, P, &dadx, &dady);
  gx += ga * dadx;
  gy += ga * dady;
  gXs[s] = make_double2(gx, gy);
  grho[s] = grh;
  double2 gvo = gV[s];
  gV[s] = make_double2(gvo.x + gvx, gvo.y + gvy);
}

// Adjoint pass P2: adj_pass2_kernel! (src/adjoint.jl:126-156) and gX += gXs.
// pair sum of adj_pass2_kernel! (src/adjoint.jl:126-156) for slot s: xs and
// grs at idx(j) hold the position and rho-bar of entry j
template <class Idx>
__device__ __forceinline__ double2 p2_sum(const double2 *__restrict__ xs,
                                          const double *__restrict__ grs, Idx idx, double2 xi,
                                          double gri, const NbRow &row, const PhysF64 &P) {
  double gx = 0.0, gy = 0.0;
  nl_walk<(kPairMode & 2) != 0>(row, [&](int j) {
    const int i = idx(j);
    double2 xj = xs[i];
    double dx = xi.x - xj.x;
    double dy = xi.y - xj.y;
    double r2 = dx * dx + dy * dy;
    if (r2 < P.r2max) {
      double q = sqrt(r2) * P.invh;
      double u = 1.0 - 0.5 * q;
      double wgt = (P.m * (((P.Fc * u) * u) * u)) * (gri + grs[i]);
      gx += wgt * dx;
      gy += wgt * dy;
    }
  });
  return make_double2(gx, gy);
}

struct NbXG64 {
  double2 x;
  double gr;
};

__global__ void __launch_bounds__(kBlock)
k_adj_p2(const double2 *__restrict__ X, const double *__restrict__ grho,
         const double2 *__restrict__ gXs, const int *__restrict__ nl,
         const int *__restrict__ nbcnt, Grid g, PhysF64 P, EpochCells ec,
         double2 *__restrict__ gX) {
  __shared__ StageTab T;
  __shared__ double2 sx[kStageArr];
  __shared__ double sgr[kStageArr];
  const int w0 = xcd_block() * kBlock, s = w0 + (int)threadIdx.x;
  bool lds = false;
  if constexpr ((kPairMode & 1) != 0)
    lds = stage_rows(T, ec, w0, g, [&](int j) { return NbXG64{X[j], grho[j]}; },
                     [&](int i, const NbXG64 &d) {
                       sx[i] = d.x;
                       sgr[i] = d.gr;
                     });
  if (s >= g.N) return;
  double2 xi = X[s];
  double gri = grho[s];
  const NbRow row = nb_row(nl, nbcnt, s, g.N);
  const double2 q = ((kPairMode & 1) && lds)
                        ? p2_sum(sx, sgr, staged_idx(T, ec.bcell[s], g), xi, gri, row, P)
                        : p2_sum(X, grho, SlotIdx{}, xi, gri, row, P);
  double2 gs = gXs[s];
  double tx = gs.x + q.x;
  double ty = gs.y + q.y;
  double2 go = gX[s];
  gX[s] = make_double2(go.x + tx, go.y + ty);
}

// target_objective (src/driver.jl:219-234) in FP64: per-block partial sums
// in a fixed order, then one block sums the partials, so J is bit-identical
// from run to run.
// (seed[s] in slot order, or seed[ids[s]] in particle-id order when ids is given)
__global__ void k_objective_partial(const double2 *__restrict__ X, int N, double xc, double yc,
                                    double s2, double *__restrict__ partials,
                                    double2 *__restrict__ seed, const int *__restrict__ ids) {
  __shared__ double red[kBlock];
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (s < N) {
    double2 x = X[s];
    double dx = x.x - xc;
    double dy = x.y - yc;
    double e = exp(-(dx * dx + dy * dy) / (2.0 * s2));
    v = -e;
    if (seed) seed[ids ? ids[s] : s] = make_double2((e * dx) / s2, (e * dy) / s2);
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void k_objective_final(const double *__restrict__ partials, int nb, double *J) {
  __shared__ double red[kBlock];
  double v = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) v += partials[i];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *J = red[0];
}

}  // namespace

void launch_nl_build_f64(hipStream_t s, const double2 *X, const int *scell, const int *cstart,
                         Grid g, CellConstF64 k, int maxnb, int *nl, int *nbcnt, int *flags) {
  if (g.N > 0)
    k_nl_build<<<nblocks(g.N), kBlock, 0, s>>>(X, scell, cstart, g, k, maxnb, nl, nbcnt, flags);
}
void launch_density_f64(hipStream_t s, const double2 *X, const int *nl, const int *nbcnt, Grid g,
                        PhysF64 P, const EpochCells &ec, double2 *PI, double *rho,
                        double2 *tapeX) {
  if (g.N > 0)
    k_density<<<nblocks(g.N), kBlock, 0, s>>>(X, nl, nbcnt, g, P, ec, PI, rho, tapeX);
}
void launch_force_f64(hipStream_t s, const double2 *X, const double2 *V, const double2 *PI,
                      const double *theta, const int *nl, const int *nbcnt, Grid g, PhysF64 P,
                      const EpochCells &ec, double2 *Xn, double2 *Vn, double2 *tapeV,
                      double2 *a_out, double *alpha_out) {
  if (g.N > 0)
    k_force<<<nblocks(g.N), kBlock, 0, s>>>(X, V, PI, theta, nl, nbcnt, g, P, ec, Xn, Vn, tapeV,
                                            a_out, alpha_out);
}
void launch_adj_r_f64(hipStream_t s, const double2 *X, const double2 *Vnext, const double *theta,
                      const int *nl, const int *nbcnt, Grid g, PhysF64 P, const EpochCells &ec,
                      double2 *PI, double2 *gX, double2 *gV, double2 *abar, double *galpha,
                      const LevelArgs &L) {
  if (g.N > 0)
    k_adj_r<<<nblocks(g.N), kBlock, 0, s>>>(X, Vnext, theta, nl, nbcnt, g, P, ec, PI, gX, gV,
                                            abar, galpha, L);
}
void launch_adj_rpw_f64(hipStream_t s, const double2 *X, const double2 *Vnext,
                        const double *theta, Grid g, PhysF64 P, double2 *gX, double2 *gV,
                        double2 *abar, double *galpha, const LevelArgs &L) {
  if (g.N > 0)
    k_adj_rpw<<<nblocks(g.N), kBlock, 0, s>>>(X, Vnext, theta, g, P, gX, gV, abar, galpha, L);
}
void launch_adj_p1_f64(hipStream_t s, const double2 *X, const double2 *V, const double2 *PI,
                       const double2 *abar, const double *galpha, const double *theta,
                       const int *nl, const int *nbcnt, Grid g, PhysF64 P, const EpochCells &ec,
                       double2 *gXs, double2 *gV, double *grho) {
  if (g.N > 0)
    k_adj_p1<<<nblocks(g.N), kBlock, 0, s>>>(X, V, PI, abar, galpha, theta, nl, nbcnt, g, P, ec,
                                             gXs, gV, grho);
}
void launch_adj_p2_f64(hipStream_t s, const double2 *X, const double *grho, const double2 *gXs,
                       const int *nl, const int *nbcnt, Grid g, PhysF64 P, const EpochCells &ec,
                       double2 *gX) {
  if (g.N > 0) k_adj_p2<<<nblocks(g.N), kBlock, 0, s>>>(X, grho, gXs, nl, nbcnt, g, P, ec, gX);
}
int objective_blocks(int N) { return N > 0 ? nblocks(N) : 1; }
void launch_objective_f64(hipStream_t s, const double2 *X, int N, double xc, double yc,
                          double sigma, double *partials, int nb, double *J, double2 *seed,
                          const int *ids) {
  double s2 = sigma * sigma;
  if (N > 0) k_objective_partial<<<nb, kBlock, 0, s>>>(X, N, xc, yc, s2, partials, seed, ids);
  else launch_fill_i32(s, (int *)partials, 2, 0);
  k_objective_final<<<1, kBlock, 0, s>>>(partials, N > 0 ? nb : 1, J);
}

}  // namespace sph
