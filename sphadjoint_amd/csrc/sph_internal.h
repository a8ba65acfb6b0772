// sph_internal.h -- shared definitions of libsphadj's kernels and host code.
//
// Device layout (HBM), per state:
//   particles live in "slots": sorted by (cell, particle id) at the last
//   neighbour rebuild, so each cell is a contiguous slot range and the three
//   rows of a 3x3 cell sweep are three contiguous slot runs;
//   X, V are double2 (interleaved x,y) per slot; ids[slot] = particle id;
//   scell[slot] = the cell whose 3x3 block the slot sweeps;
//   cstart[0..ncell] = first slot of each cell (exclusive prefix, cstart[ncell] = N).
// Cells and ids are 0-based on the device; the ABI converts to the
// reference's 1-based convention.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

constexpr int kBlock = 256;

// Cell grid of a state (0-based cells c = cx + cy*nx); inv_nx = 1/nx in FP64.
struct Grid {
  int nx, ny, ncell, N;
  double inv_nx;
};

// Constants of the FP64 kernels.  Every value is formed on the host with the
// operation order the reference uses in T (see sph_api.cpp: make_phys).
struct PhysF64 {
  double m, rho0, c2, mu, gx, gy, dt, Lx, Ly, kw;
  double A, invh, r2max, r2min, Fc, Gc, twomumass;
  double hx, hy, gxmax, gymax, gxclamp, gyclamp;  // theta grid: spacing, ngx-1, ngx-1-1e-6
  int ngx, ngy;
};

// Cell-index constants (src/neighbors.jl:38-42, 192-218): cs in T, 1/cs in T,
// squared cutoffs in T.
struct CellConstF64 {
  double cs, invcs, rc2, r2min, dmax2;
};

// Design-field grid for the gtheta gather (src/params.jl:30-32,53-68).
struct ThetaGrid {
  int ngx, ngy;
  double hx, hy, inv_hx, inv_hy, gxclamp, gyclamp, cs;
};

// Slot block of this workgroup.  Workgroups are dealt round-robin to the 8
// XCDs (each with its own L2); renumbering them so that every XCD walks one
// contiguous range of slots keeps the three cell rows a sweep reads in that
// XCD's L2 instead of fetching them once per XCD.
__device__ __forceinline__ int xcd_block() {
  const int w = blockIdx.x, nb = gridDim.x;
  const int x = w & 7, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + (w >> 3);
}

// ---- integer helpers -------------------------------------------------------

// Run heads of equal keys within a wavefront.  Lanes whose key is -1 are
// inactive.  Returns the ballot of heads; `head_lane` is the head of this
// lane's run and `run_len` is set on heads.
__device__ __forceinline__ unsigned long long wave_runs(int key, int &head_lane, int &run_len) {
  const int lane = threadIdx.x & 63;
  int prev = __shfl_up(key, 1, 64);
  bool head = (lane == 0) || (prev != key);
  unsigned long long heads = __ballot(head);
  // next head strictly above this lane (or 64)
  unsigned long long above = (lane == 63) ? 0ull : (heads >> (lane + 1)) << (lane + 1);
  int next = above ? __ffsll((long long)above) - 1 : 64;
  unsigned long long below = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
  head_lane = 63 - __clzll((long long)below);
  run_len = next - lane;
  return heads;
}


// unsafe_trunc(Int, floor(v)) + 2 clamped to [1, n], then made 0-based.
// NaN / out-of-range inputs give the lowest cell, as x86-64's cvttsd2si
// (typemin) followed by the clamp does in the reference.
__host__ __device__ inline int cell_coord_div(double v, int n) {
  double f = floor(v);
  long long t = (f > -1e15 && f < 1e15) ? (long long)f : (long long)(-(1LL << 60));
  long long c = t + 2;
  c = c < 1 ? 1 : (c > n ? n : c);
  return (int)(c - 1);
}
// unsafe_trunc(Int32, floor(v)) + Int32(2) clamped to [1, n], 0-based.
__device__ __forceinline__ int cell_coord_i32(double v, int n) {
  double f = floor(v);
  long long t = (f > -2147483648.0 && f < 2147483647.0) ? (long long)(int)f
                                                          : (long long)INT32_MIN;
  long long c = t + 2;
  c = c < 1 ? 1 : (c > n ? n : c);
  return (int)(c - 1);
}

// Quadrant of position x in its 0-based binning cell (cx, cy), which spans
// [(cx-1) cs, cx cs) x [(cy-1) cs, cy cs): 2*qx + qy, qx = 1 in the right
// half, qy = 1 in the upper half (v = x/cs as the binning divides, so a
// position left of a clamped edge cell counts as left, right of one as
// right).  The 5x5 half-cell sweep of a list build orders each cell's slots by
// quadrant.  A slot in the left half (qx = 0) of cell column cx has every
// position within cs of it in columns cx-1 and cx or in the left half of
// column cx+1; one in the right half, in the right half of cx-1 or in columns
// cx and cx+1; rows likewise.  So the sweep scans 5x5 half cells, about 69 %
// of the 3x3 block's candidates, and misses no pair closer than cs (up to the
// rounding of v at a half-cell edge, which only concerns pairs at a distance
// of cs within that rounding).
__host__ __device__ inline int cell_quadrant(double2 x, double cs, int cx, int cy) {
  const int qx = (x.x / cs - (double)(cx - 1)) >= 0.5 ? 1 : 0;
  const int qy = (x.y / cs - (double)(cy - 1)) >= 0.5 ? 1 : 0;
  return 2 * qx + qy;
}

// Sweep level of one reverse step's gtheta gather: 0 tight ranges (a quarter
// cell, f32c form), 1 a one-cell margin, m >= 2 an m-cell margin (the SPH
// cells holding a theta cell's positions, widened by m cells on each side:
// a slot that moved m cells since the rebuild is still swept).  The R pass
// of the step raises it with an atomic max of (stamp << 32 | level) for every
// slot its level would miss (the maximum does not depend on the order of the
// updates); the gather reads the level of its own stamp, 0 for any other.
__device__ __forceinline__ void raise_level(unsigned long long *lev, int stamp, int level) {
  atomicMax(lev, ((unsigned long long)(unsigned)stamp << 32) | (unsigned)level);
}
__device__ __forceinline__ int level_of(const unsigned long long *lev, int stamp) {
  const unsigned long long v = *lev;
  return (unsigned)(v >> 32) == (unsigned)stamp ? (int)(unsigned)v : 0;
}

// The three contiguous slot runs of a 3x3 sweep around cell `sc`, in the
// reference's order (ddy outer, ddx inner; src/neighbors.jl:220-231).
struct Runs {
  int b[3], e[3];
};
__device__ __forceinline__ Runs sweep_runs(const int *__restrict__ cstart, int sc, int nx,
                                           int ny) {
  Runs r;
  int cy = sc / nx;
  int cx = sc - cy * nx;
  int x0 = cx > 0 ? cx - 1 : 0;
  int x1 = cx < nx - 1 ? cx + 1 : nx - 1;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int jy = cy + k - 1;
    if (jy >= 0 && jy < ny) {
      r.b[k] = cstart[jy * nx + x0];
      r.e[k] = cstart[jy * nx + x1 + 1];
    } else {
      r.b[k] = 0;
      r.e[k] = 0;
    }
  }
  return r;
}

// Row of the transposed neighbour list (the reference's slot-major layout,
// src/neighbors.jl:124-131): entry m of slot s at nl[m*N + s], so the lanes
// of a wavefront read consecutive words for each m.
struct NbRow {
  const int *__restrict__ p;
  int n, N;
  __device__ __forceinline__ int at(int m) const { return p[(size_t)m * N]; }
};
__device__ __forceinline__ NbRow nb_row(const int *__restrict__ nl,
                                        const int *__restrict__ nbcnt, int s, int N) {
  NbRow r;
  r.p = nl + s;
  r.n = nbcnt[s];
  r.N = N;
  return r;
}

// ---- launchers ----------------------------------------------------------------

// cell list (sph_cells.hip)
void launch_fill_i32(hipStream_t s, int *p, int n, int v);
// (flags[2] = 1 when a position is not finite)
void launch_assign(hipStream_t s, const double2 *X, Grid g, double cs, int *cellof, int *counts,
                   int *flags);
// exclusive prefix of in[0..n) into out[0..n] (and cursor); clears in; lists
// the cells with more than 128 entries in big (big[0] = count, n + 1 entries)
void launch_scan_excl(hipStream_t s, int *in, int n, int *out, int *cursor, int *bsum,
                      int *big);
// entries at their cells' cursors with their rank keys: the ids, or with qX
// (the positions; g, cs the single grid of the state) quadrant * N + id, which
// orders each cell by quadrant, then id (N < 2^29)
void launch_fill(hipStream_t s, const int *cellof, const int *ids, int N, int *cursor,
                 int *tmp_id, int *tmp_slot, const double2 *qX = nullptr, Grid g = Grid{},
                 double cs = 0.0);
// rank by key inside each cell; with ids (keys that are not ids), new_ids
// takes ids[slot] instead of the key
void launch_rank(hipStream_t s, const int *cellof, const int *cstart, const int *tmp_id,
                 const int *tmp_slot, int N, const int *big, int *perm, int *new_ids,
                 const int *ids = nullptr);
// first slot of each quadrant of each cell of a quadrant-ordered layout:
// qs[4c + k], k = 0..3, and qs[4 ncell] = N (4 ncell + 1 entries)
void launch_quad_starts(hipStream_t s, const double2 *X, const int *cstart, Grid g, double cs,
                        int *qs);
// also writes bcell[slot] = the slot's binning cell (cellof, the division rule)
void launch_permute(hipStream_t s, const int *perm, const int *cellof, const double2 *X,
                    const double2 *V, const double2 *Xref, int check, Grid g, CellConstF64 k,
                    double2 *Xn, double2 *Vn, double2 *Xrefn, int *scell, int *bcell,
                    int *flags);
void launch_materialise(hipStream_t s, const double2 *Xref, const int *ids, const int *scell,
                        const int *cstart, Grid g, CellConstF64 k, int maxnb, long long sm,
                        long long si, int *indices, int *nbcount, int *flags);
void launch_scatter_i32(hipStream_t s, const int *src, const int *ids, int N, int *dst);
// an epoch's layout: ids, sweep cells and cell starts; binning cells (FP64
// form) and the pair passes' thread order (f32c form) where present
struct EpochLayout {
  const int *ids, *scell, *bcell, *lane, *cstart;
};
struct EpochLayoutOut {
  int *ids, *scell, *bcell, *lane, *cstart;
};
void launch_epoch_record(hipStream_t s, const EpochLayout &src, int N, int ncs,
                         const EpochLayoutOut &dst);
void launch_inverse(hipStream_t s, const int *ids, int N, int *inv);
void launch_gather_remap2(hipStream_t s, const double2 *src, const int *inv_src,
                          const int *ids_dst, int N, double2 *dst);
void launch_scatter2(hipStream_t s, const double2 *src, const int *ids, int N, double2 *dst);
void launch_gather2(hipStream_t s, const double2 *src, const int *ids, int N, double2 *dst);
void launch_scatter1(hipStream_t s, const double *src, const int *ids, int N, double *dst);
void launch_copy2(hipStream_t s, const double2 *src, int N, double2 *dst);

// FP64 physics (sph_f64.hip)
void launch_nl_build_f64(hipStream_t s, const double2 *X, const int *scell, const int *cstart,
                         Grid g, CellConstF64 k, int maxnb, int *nl, int *nbcnt, int *flags);
// The layout of the list's epoch, for the LDS staging of the FP64 pair
// passes: each slot's binning cell (bcell) and sweep cell (scell), and the
// cell starts.
struct EpochCells {
  const int *bcell, *scell, *cstart;
};
void launch_density_f64(hipStream_t s, const double2 *X, const int *nl, const int *nbcnt, Grid g,
                        PhysF64 P, const EpochCells &ec, double2 *PI, double *rho,
                        double2 *tapeX);
void launch_force_f64(hipStream_t s, const double2 *X, const double2 *V, const double2 *PI,
                      const double *theta, const int *nl, const int *nbcnt, Grid g, PhysF64 P,
                      const EpochCells &ec, double2 *Xn, double2 *Vn, double2 *tapeV,
                      double2 *a_out, double *alpha_out);
// what an R pass needs to raise its step's gtheta sweep level (raise_level):
// each slot's binning cell at the epoch's rebuild, the level word and the
// step's stamp, and the theta grid the gather sweeps for
struct LevelArgs {
  const int *bcell;
  const int2 *trange;   // one-cell-margin ranges: ncx theta columns, then the theta
                        // rows; then the quarter-cell ranges in the same order
  unsigned long long *lev;
  int stamp, ncx, ngy;
  int *key;             // out: each slot's interp cell l00, for the keyed gather
};
// the margin (>= 2) a sweep needs to reach binning cell c when c lies beyond
// its one-cell-margin range w (theta_cell_range, which clamps to the axis), else 0
__host__ __device__ inline int margin_beyond(int c, int2 w) {
  return c < w.x ? w.x + 1 - c : (c > w.y ? c - w.y + 1 : 0);
}
// the FP64 gtheta gather over the keys (LevelArgs::key) of the step's R pass
// (trange: the table of LevelArgs::trange)
void launch_theta_cells_keyed_f64(hipStream_t s, const double2 *X, const double *galpha,
                                  const int *key, const int *cstart, const int2 *trange, Grid g,
                                  ThetaGrid T, bool first, double4 *part,
                                  const unsigned long long *lev, int stamp);
void launch_adj_r_f64(hipStream_t s, const double2 *X, const double2 *Vnext, const double *theta,
                      const int *nl, const int *nbcnt, Grid g, PhysF64 P, const EpochCells &ec,
                      double2 *PI, double2 *gX, double2 *gV, double2 *abar, double *galpha,
                      const LevelArgs &L);
void launch_adj_rpw_f64(hipStream_t s, const double2 *X, const double2 *Vnext,
                        const double *theta, Grid g, PhysF64 P, double2 *gX, double2 *gV,
                        double2 *abar, double *galpha, const LevelArgs &L);
void launch_adj_p1_f64(hipStream_t s, const double2 *X, const double2 *V, const double2 *PI,
                       const double2 *abar, const double *galpha, const double *theta,
                       const int *nl, const int *nbcnt, Grid g, PhysF64 P, const EpochCells &ec,
                       double2 *gXs, double2 *gV, double *grho);
void launch_adj_p2_f64(hipStream_t s, const double2 *X, const double *grho, const double2 *gXs,
                       const int *nl, const int *nbcnt, Grid g, PhysF64 P, const EpochCells &ec,
                       double2 *gX);
// J = the fixed-order sum of nblocks partials (partials: nblocks scratch
// doubles, at least 2); seed (may be NULL) in slot order, or in particle-id
// order through ids when ids is given
void launch_objective_f64(hipStream_t s, const double2 *X, int N, double xc, double yc,
                          double sigma, double *partials, int nblocks, double *J,
                          double2 *seed, const int *ids = nullptr);
int objective_blocks(int N);

// node sums of the theta-cell partials (sph_theta.hip), added to gtheta once
// per neighbour-list epoch
void launch_theta_nodes_f64(hipStream_t s, const double4 *part, ThetaGrid T, double *gtheta);

// design-field chain of the topology optimiser (sph_design.hip): an
// (ngy, ngx) grid, the density filter radius r (grid points) and the
// Borrvall-Petersson constants
struct DesignGrid {
  int ngy, ngx, r;
  double alpha_max, qp;
};
void launch_design_filter(hipStream_t s, const double *x, const double *wtab, const double *snorm,
                          const DesignGrid &D, double *y);
void launch_design_filter_adjoint(hipStream_t s, const double *g, const double *wtab,
                                  const double *snorm, const DesignGrid &D, double *out);
void launch_design_alpha(hipStream_t s, const double *rdf, const DesignGrid &D, double *theta);
void launch_design_chain(hipStream_t s, const double *gtheta, const double *rdf,
                         const DesignGrid &D, double *grdf);
// out = project_volume(in - step * g / max|g|) (g may be NULL: project_volume(in));
// gmax (device, may be NULL) receives max|g|
void launch_design_project(hipStream_t s, const double *in, const double *g, double step, int n,
                           double volfrac, double *out, double *gmax);


}  // namespace sph
