// sph_f32c.h -- constants and launchers of the FP32-compensated form.
#pragma once
#include "sph_internal.h"

namespace sph {

// Constants of the f32c kernels.  Theta-grid quantities stay in FP64 because
// interpolation reads the FP64 position; pair constants are FP32.
struct PhysF32 {
  float m, rho0, c2, mu, gx, gy, dt, Lx, Ly, kw;
  float mA, invh, r2max, r2min, Fc, Gc, twomumass, cs;
  float inv_hx, inv_hy;
  // products folded once on the host: u = 1 - q/2 = fma(r, mhh, 1) with
  // mhh = -1/(2h); 2q + 1 = fma(r, twoinvh, 1); mFc = m*Fc; Gch = Gc/h
  float mhh, twoinvh, mFc, Gch;
  double dt64, hx, hy, gxmax, gymax, gxclamp, gyclamp;
  double inv_hx64, inv_hy64;   // 1/hx, 1/hy: interp_cell's grid coordinates
  int ngx, ngy;
};

// One scene of an f32c state.  A state holds one or more independent scenes
// (batched line-search candidates, say) side by side: scene k owns the slots
// [off, off + g.N), the cells [coff, coff + g.ncell) of the global cell
// numbering (cstart holds global slots), the design-field nodes
// [toff, toff + ngx*ngy) and the design cells [tcoff, tcoff + (ngx-1)*(ngy-1)).
// Every per-slot launch covers all scenes at once: its workgroups are dealt to
// the scenes in order (Batch::boff[k] = scene k's first workgroup), so a
// workgroup never straddles two scenes and reads its scene's constants once,
// through scalar loads.
struct Scene32 {
  Grid g;                  // the scene's own cell grid; g.N = its particle count
  PhysF32 P;
  ThetaGrid T;
  double cs64;             // cell width in FP64 (binning by the division rule)
  double dmax2;            // squared displacement allowed between rebuilds (skin*h)^2
  int off, coff;           // first slot, first cell
  int toff, tcoff, tboff;  // first theta node, first theta cell, first workgroup of
                           // the keyed gtheta gather (theta_strip_blocks per scene)
  int nboff;               // first workgroup of theta-node launches
  int troff;               // first entry of the scene's SPH ranges in Batch::trange
};

// Launch geometry of an f32c state: device scene table, workgroup -> scene
// maps of the three launch shapes, and totals.  Per-slot launches use bs
// threads per workgroup (one of kSlotBlocks; the autotuner picks it per
// kernel), with b2s and boff built for that size.
struct Batch {
  const Scene32 *sc;
  const int *b2s, *boff, *tb2s, *nb2s;
  // per scene (from Scene32::troff): for each theta column i the SPH cell
  // columns (cx0, cx1) that can hold slots whose interp cell has column i,
  // then for each theta row j the SPH cell rows (cy0, cy1); see
  // theta_cell_range.  trange allows a slot one cell beyond its binning cell,
  // trange_tight a quarter cell (used for a step in which no slot is further out)
  const int2 *trange, *trange_tight;
  int bs;
  int nostage;   // 1: the pair passes read neighbours from global memory (no LDS staging)
  int nscenes, nblk, ntblk, nnblk;
  int N, ncell, ntnode, ntcell;
};
constexpr int kSlotBlocks[3] = {64, 128, 256};
// kernels whose workgroup size is tuned (sph_autotune)
enum TunedKernel { TK_NLBUILD, TK_DENSITY, TK_FORCE, TK_ADJ_R, TK_ADJ_P1, TK_ADJ_P2, TK_COUNT };
constexpr int kThetaLanes = 4;   // lanes that share one theta cell in the gtheta gather
constexpr int kThetaStrip = 64;  // theta cells of one theta row per gtheta-gather workgroup
// SPH cell range [c0, c1] holding slots whose interp cell has index i along
// an axis of nt theta cells of width ht: the cells of the theta cell's
// interval (the division rule of cell_index) widened by one cell on each side
// for the displacement since the last rebuild, the whole axis for the edge
// cells (the interpolation clamps there)
inline int2 theta_cell_range(int i, int nt, double ht, double cs, int n) {
  int c0 = (i == 0) ? 0 : cell_coord_div((double)i * ht / cs, n) - 1;
  int c1 = (i == nt - 1) ? n - 1 : cell_coord_div((double)(i + 1) * ht / cs, n) + 1;
  c0 = c0 < 0 ? 0 : c0;
  c1 = c1 > n - 1 ? n - 1 : c1;
  return make_int2(c0, c1);
}
// The same range when a slot may lie at most ex beyond its binning cell: a slot
// in binning cell c at x has (c-1) cs - ex <= x < c cs + ex, so c lies in
// [cell_of(x - ex), cell_of(x + ex)], and x in [i ht, (i+1) ht) bounds both
// (slack covers the rounding of the interpolation coordinate)
inline int2 theta_cell_range_tight(int i, int nt, double ht, double cs, int n, double ex) {
  const double slack = 1e-9 * cs;
  int c0 = (i == 0) ? 0 : cell_coord_div(((double)i * ht - ex - slack) / cs, n);
  int c1 = (i == nt - 1) ? n - 1 : cell_coord_div(((double)(i + 1) * ht + ex + slack) / cs, n);
  return make_int2(c0, c1);
}
// workgroups of the keyed gtheta gather of one scene: one per strip
inline long long theta_strip_blocks(int ngx, int ngy) {
  return (long long)(ngy - 1) * ((ngx - 1 + kThetaStrip - 1) / kThetaStrip);
}

// Arguments of the adjoint R pass without the density recompute (tapes that
// kept each step's pterm/invrho), which a P2 launch can also run for the
// previous step (same slot order) as each slot's epilogue
struct RpwArgs {
  const double2 *X;      // positions of the step
  const float2 *Vnext;   // velocities after the step
  const double *theta;
  float2 *xr;            // out: relative coordinates of the step
  float2 *gV, *abar;
  float *galpha;
  int *tkey;
  unsigned long long *tlev;   // the step's gtheta sweep level (raise_level)
  int tstamp;
};

// bitmask neighbour list: 32-bit words per row (runs of up to 32*kMaskWords slots)
constexpr int kMaskWords = 4;
// the kernels address a mask list (3*kMaskWords words per slot) with 32-bit
// byte offsets, which bounds the slots of an f32c state
constexpr long long kMaxSlotsF32C = 0xFFFFFFFFLL / (3 * kMaskWords * 4);

// (flags[2] = 1 when a position is not finite)
void launch_assign_f32c(hipStream_t s, const double2 *X, const Batch &B, int *cellof,
                        int *counts, int *flags);
// mask-list build; with qs (the quadrant starts of a quadrant-ordered
// single-scene layout, launch_quad_starts) the 5x5 half-cell sweep
void launch_nm_build_f32c(hipStream_t s, const float2 *xr, const int *bcell, const int *cstart,
                          const Batch &B, uint32_t *nm, int *flags, const int *qs = nullptr);
// first density pass after a rebuild, fused with the mask-list build: writes
// the mask list nm, the exact list em and PI (flags is not written: no pair is
// ever left out, see kMaskRun); qs as for launch_nm_build_f32c
void launch_density_build_f32c(hipStream_t s, const float2 *xr, const int *bcell,
                               const int *cstart, const int *lane, const Batch &B, float2 *PI,
                               uint32_t *nm, uint32_t *em, int *flags, const int *qs = nullptr);
// density + EOS over the mask list nm; when em is given, the exact list of
// these positions (pairs with r2 < (2h)^2) is written there for the force pass
void launch_density_f32c(hipStream_t s, const float2 *xr, const int *bcell, const uint32_t *nm,
                         const int *cstart, const int *lane, const Batch &B, float2 *PI,
                         double *rho, uint32_t *em);
// force + integrator over the exact list nm of the density pass
void launch_force_f32c(hipStream_t s, const double2 *X, const float2 *xr, const float2 *V,
                       const float2 *PI, const double *theta, const int *bcell, const uint32_t *nm,
                       const int *cstart, const int *lane, const Batch &B, double2 *Xn,
                       float2 *xrn, float2 *Vn, double2 *tapeX, float2 *tapeV, double2 *a_out,
                       double *alpha_out);
// relative coordinates xr of the slots; with `lane` (B: the 256-slot
// geometry, Bv[2]) also the thread -> slot order of the pair passes for these
// positions, which every pair launcher takes as `lane`
void launch_xrel_f32c(hipStream_t s, const double2 *X, const int *bcell, const Batch &B,
                      float2 *xr, int *lane = nullptr);
// the slot reorder of a rebuild, and with `lane` the new thread order as above
void launch_permute_f32c(hipStream_t s, const int *perm, const int *cellof, const double2 *X,
                         const float2 *V, const double2 *Xref, int check, const Batch &B,
                         double2 *Xn, float2 *Vn, float2 *xr, double2 *Xrefn, int *bcell,
                         int *flags, int *lane);
// a checkpoint snapshot of a rebuilt single-scene layout: to_snap copies
// (X, V, ids, scell, lane, cstart) into (qX, ...); otherwise copies them back,
// X also into Xref, and forms xr
void launch_snapshot_f32c(hipStream_t s, bool to_snap, const Batch &B, double2 *X, float2 *V,
                          int *ids, int *scell, int *lane, int *cstart, double2 *Xref,
                          float2 *xr, double2 *qX, float2 *qV, int *qids, int *qscell,
                          int *qlane, int *qcstart);
// R passes also write each slot's interp-cell key for the keyed gtheta gather
// and raise the step's sweep level in *tlev (raise_level, with stamp tstamp,
// which differs from step to step) for a slot beyond its interp cell's tight
// SPH range (Batch::trange_tight; level 1) or beyond the one-cell range
// (Batch::trange; the margin that reaches it)
void launch_adj_r_f32c(hipStream_t s, const double2 *X, const float2 *xr, const float2 *Vnext,
                       const double *theta, const int *bcell, const uint32_t *nm,
                       const int *cstart, const int *lane, const Batch &B, float2 *PI,
                       float2 *gX, float2 *gV, float2 *abar, float *galpha, uint32_t *em,
                       int *tkey, unsigned long long *tlev, int tstamp);
void launch_adj_rpw_f32c(hipStream_t s, const double2 *X, const int *bcell, const float2 *Vnext,
                         const double *theta, const Batch &B, float2 *xr, float2 *gX,
                         float2 *gV, float2 *abar, float *galpha, int *tkey,
                         unsigned long long *tlev, int tstamp);
// the same R pass on the adjoint state (gXo, gVo) of another slot order,
// gathered through inv[ids[s]] (V_{k+1}, Vno, too when remap_v, else it is in
// this order already): gX into gXn, gV into gVn
void launch_remap_rpw_f32c(hipStream_t s, const double2 *X, const int *bcell,
                           const float2 *gXo, const float2 *gVo, const float2 *Vno,
                           const int *inv, const int *ids, const double *theta, const Batch &B,
                           float2 *gXn, float2 *gVn, float2 *xr, float2 *abar, float *galpha,
                           int *tkey, unsigned long long *tlev, int tstamp, bool remap_v);
// P1 over nm: the exact list when `exact`, else the mask list, whose exact
// subset is then written to em for P2
void launch_adj_p1_f32c(hipStream_t s, const double2 *X, const float2 *xr, const float2 *V,
                        const float2 *PI, const float2 *abar, const float *galpha,
                        const double *theta, const int *bcell, const uint32_t *nm, bool exact,
                        const int *cstart, const int *lane, const Batch &B, float2 *gXs,
                        float2 *gV, float *grho, uint32_t *em);
// P2 over the exact list nm; with `prev`, each slot then runs the R pass of
// the previous step (rpw) on its final gX (prev->xr must not alias xr)
void launch_adj_p2_f32c(hipStream_t s, const float2 *xr, const float *grho, const float2 *gXs,
                        const int *bcell, const uint32_t *nm, const int *cstart, const int *lane,
                        const Batch &B, float2 *gX, const RpwArgs *prev = nullptr);
// deterministic gtheta gather over the keys of the R passes: the theta-cell
// partials of one step, written (first != 0) or added to part; the gather
// sweeps the SPH ranges of the step's level (*tlev at stamp tstamp: 0 the
// tight ranges, 1 the one-cell ranges, m >= 2 an m-cell margin).
// launch_theta_nodes_keyed_f32 (sph_theta.hip) adds the node sums of part to
// gtheta (once per neighbour-list epoch).
struct ThetaKeyed {
  const int *key;
  const double2 *X;
  const float *galpha;
  const int *cstart;
  const unsigned long long *tlev;
  int tstamp, first;
  double4 *part;
};
void launch_theta_gather_keyed_f32(hipStream_t s, const ThetaKeyed &A, const Batch &B);
void launch_theta_nodes_keyed_f32(hipStream_t s, const double4 *part, const Batch &B,
                                  double *gtheta);
// launch_adj_p1_f32c and launch_theta_gather_keyed_f32 as one launch (same
// results); B must have 256-thread workgroups (B.bs == kBlock)
void launch_adj_p1_theta_f32c(hipStream_t s, const double2 *X, const float2 *xr,
                              const float2 *V, const float2 *PI, const float2 *abar,
                              const float *galpha, const double *theta, const int *bcell,
                              const uint32_t *nm, bool exact, const int *cstart, const int *lane,
                              const Batch &B, float2 *gXs, float2 *gV, float *grho, uint32_t *em,
                              const ThetaKeyed &A);

// float2 <-> double2 helpers and float2 layout remaps (sph_cells.hip)
void launch_f2_to_d2(hipStream_t s, const float2 *src, int N, double2 *dst);
void launch_d2_to_f2(hipStream_t s, const double2 *src, int N, float2 *dst);
// (da, db, dc)[s] = (a, b, c)[inv_src[ids_dst[s]]]
void launch_gather_remap3f(hipStream_t s, const float2 *a, const float2 *b, const float2 *c,
                           const int *inv_src, const int *ids_dst, int N, float2 *da, float2 *db,
                           float2 *dc);
void launch_gather_remap2f(hipStream_t s, const float2 *src, const int *inv_src,
                           const int *ids_dst, int N, float2 *dst);
void launch_gather2f(hipStream_t s, const double2 *src, const int *ids, int N, float2 *dst);
void launch_scatter2f(hipStream_t s, const float2 *src, const int *ids, int N, double2 *dst);
void launch_scatter_pi(hipStream_t s, const float2 *src, const int *ids, int N, double2 *dst);

}  // namespace sph
