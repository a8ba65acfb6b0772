// sph_f32c.hip -- FP32-compensated physics kernels for gfx950 (the "f32c" form).
//
// Arithmetic form:
//   * positions are kept in FP64 (X64) and advanced in FP64 (x += dt*v);
//   * every gather reads FP32 positions relative to the particle's cell origin
//     (xr = x - origin(cell), |xr| < cell width), so a pair difference keeps
//     ~1e-10 absolute accuracy where FP32 absolute coordinates would give ~1e-7;
//   * velocities, pterm/invrho and the adjoint state are FP32;
//   * density, acceleration and adjoint pair sums use TwoSum (Knuth) error-free
//     accumulation; rho - rho0 is formed from the compensated pair;
//   * gtheta is reduced in FP64 by the deterministic gather (sph_theta.hip),
//     J in FP64;
//   * square roots use v_sqrt_f32 / v_rsq_f32 (1 ulp) instead of the correctly
//     rounded expansions, and 1/r is an rsq instead of a division.
// Each rebuild builds a bitmask neighbour list: for each slot and each of the
// three cell rows of its 3x3 block, the row's cells form one contiguous slot
// run [rb, rb+len), and bit b of the row's mask flags slot rb+b as a list
// neighbour (r2min < r2 < rc2).  Walking the set bits in ascending order
// visits the reference's order (ddy outer, ddx inner, ascending id inside a
// cell).  The cell of bit b inside the row, and with it the exact FP32 shift
// between the two cell frames, follows from the run's cell boundaries.  A
// mask costs 8 bytes per row instead of 4 bytes per entry (~24 entries).
// Every sum has a fixed order: repeated runs are bit-identical.  Physics
// follows src/forward.jl:25-147 and src/adjoint.jl:30-228.
#include "sph_f32c.h"
#include "sph_stage.h"

#include <stdexcept>

// No implicit contraction: every FMA below is written out, so the rounding of
// each expression is fixed by the source and identical in every kernel that
// evaluates it (the checkpointed and full-tape reverse sweeps must agree bit
// for bit), and the compensated sums see the roundings they assume.
#pragma clang fp contract(off)

namespace sph {

namespace {

inline int nblocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// TwoSum accumulator: s + c carries the sum with an error of ~1 ulp of s.
struct CSum {
  float s, c;
  __device__ __forceinline__ CSum() : s(0.f), c(0.f) {}
  __device__ __forceinline__ void add(float t) {
    float n = s + t;
    float bp = n - s;
    float err = (s - (n - bp)) + (t - bp);
    s = n;
    c += err;
  }
  __device__ __forceinline__ float value() const { return s + c; }
};

// Fast2Sum accumulator: the error term is exact when |s| >= |t| for every
// added t (density terms are positive and none exceeds the self term, which
// is added first).
struct FSum {
  float s, c;
  __device__ __forceinline__ explicit FSum(float s0) : s(s0), c(0.f) {}
  __device__ __forceinline__ void add(float t) {
    float n = s + t;
    c += t - (n - s);
    s = n;
  }
  __device__ __forceinline__ float value() const { return s + c; }
};

// plain FP32 accumulator
struct PSum {
  float s;
  __device__ __forceinline__ PSum() : s(0.f) {}
  __device__ __forceinline__ void add(float t) { s += t; }
  __device__ __forceinline__ float value() const { return s; }
};

// Adjoint pair sums (P1, P2) accumulate in plain FP32: their rounding error
// is far below the error the FP32 state itself carries into the gradients
// (measured against FP64 with and without TwoSum: the same to two digits).
typedef PSum ASum;

__device__ __forceinline__ float wall_accel(float x, float L, float kw) {
  float lo = fmaxf(0.f, -x);
  float hi = fmaxf(0.f, x - L);
  return kw * (lo * lo - hi * hi);
}
__device__ __forceinline__ float wall_daccel(float x, float L, float kw) {
  float lo = fmaxf(0.f, -x);
  float hi = fmaxf(0.f, x - L);
  return kw * (-2.f * lo - 2.f * hi);
}

struct InterpF {
  int l00;
  float fx, fy, ax, ay;
  double fx64, fy64;   // the fractions before rounding to FP32
};

// interp_cell (src/params.jl:53-68) from the FP64 position
__device__ __forceinline__ InterpF interp_cell(double x, double y, const PhysF32 &P) {
  InterpF r;
  double gx = x * P.inv_hx64;
  double gy = y * P.inv_hy64;
  r.ax = (gx > 0.0 && gx < P.gxmax) ? 1.f : 0.f;
  r.ay = (gy > 0.0 && gy < P.gymax) ? 1.f : 0.f;
  double gxc = (gx > P.gxclamp) ? P.gxclamp : ((gx < 0.0) ? 0.0 : gx);
  double gyc = (gy > P.gyclamp) ? P.gyclamp : ((gy < 0.0) ? 0.0 : gy);
  double fxf = floor(gxc), fyf = floor(gyc);
  int i0 = (fxf > -1e9 && fxf < 1e9) ? (int)fxf : -1000000000;
  int j0 = (fyf > -1e9 && fyf < 1e9) ? (int)fyf : -1000000000;
  i0 = i0 < 0 ? 0 : (i0 > P.ngx - 2 ? P.ngx - 2 : i0);
  j0 = j0 < 0 ? 0 : (j0 > P.ngy - 2 ? P.ngy - 2 : j0);
  r.fx64 = gxc - (double)i0;
  r.fy64 = gyc - (double)j0;
  r.fx = (float)r.fx64;
  r.fy = (float)r.fy64;
  r.l00 = j0 + i0 * P.ngy;
  return r;
}

__device__ __forceinline__ float interp_alpha(const double *__restrict__ th, const InterpF &c,
                                              const PhysF32 &P, float *dadx, float *dady) {
  float t00 = (float)th[c.l00];
  float t01 = (float)th[c.l00 + P.ngy];
  float t10 = (float)th[c.l00 + 1];
  float t11 = (float)th[c.l00 + P.ngy + 1];
  float omx = 1.f - c.fx, omy = 1.f - c.fy;
  float a = omx * omy * t00 + c.fx * omy * t01 + omx * c.fy * t10 + c.fx * c.fy * t11;
  if (dadx) {
    *dadx = ((omy * (t01 - t00) + c.fy * (t11 - t10)) * P.inv_hx) * c.ax;
    *dady = ((omx * (t10 - t00) + c.fx * (t11 - t01)) * P.inv_hy) * c.ay;
  }
  return a;
}

// origin of 0-based cell (cx, cy): ((cx-1) cs32, (cy-1) cs32) in FP64
__device__ __forceinline__ float2 rel_coord(double2 x, int cell, const Grid &g, const PhysF32 &P) {
  const int2 c = cell_xy(cell, g);
  const int cx = c.x, cy = c.y;
  double ox = (double)(cx - 1) * (double)P.cs;
  double oy = (double)(cy - 1) * (double)P.cs;
  return make_float2((float)(x.x - ox), (float)(x.y - oy));
}

// ---- raw buffer access -------------------------------------------------------
// Neighbour-side reads go through buffer loads with a 32-bit byte offset: one
// offset serves every per-slot array of the pair, the 128-bit descriptor lives
// in scalar registers, and a read past the array returns 0 instead of
// faulting.  Descriptor word 3 for gfx9 raw buffers: 0x00020000.
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                           0x00020000);
}
__device__ __forceinline__ float2 ld2(Rsrc r, uint32_t off) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ float ld1(Rsrc r, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ uint32_t ldu(Rsrc r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}

// The scene of this workgroup and the slot s of this thread (threads past the
// scene's last slot leave); g and P are the scene's grid and constants.
#define SCENE_SLOT(B)                                                   \
  const int blk_ = xcd_block();                                         \
  const int scn_ = (B).b2s[blk_];                                       \
  const Scene32 S = (B).sc[scn_];                                       \
  const int sl_ = (blk_ - (B).boff[scn_]) * (B).bs + (int)threadIdx.x;  \
  if (sl_ >= S.g.N) return;                                             \
  const int s = S.off + sl_;                                            \
  const Grid &g = S.g;                                                  \
  const PhysF32 &P = S.P

// Pair passes: thread sl_ of the scene takes the slot lane[off + sl_] (see
// lane_order_window), a permutation of each aligned window of kLaneWin slots of the
// scene, so that the lanes of a wavefront sweep rows of similar length; each
// slot's arithmetic is unchanged by this assignment.  w0_..w1_ is the window
// of this workgroup's slots.  Threads past the scene's end (valid_ false) take
// part in the workgroup's barriers and then leave.
constexpr int kLaneWin = 256;
#define SCENE_WIN(B, lane) SCENE_WIN_AT(B, lane, xcd_block())
#define SCENE_WIN_AT(B, lane, blk)                                                     \
  const int blk_ = (blk);                                                              \
  const int scn_ = (B).b2s[blk_];                                                      \
  const Scene32 S = (B).sc[scn_];                                                      \
  const int bsl_ = (blk_ - (B).boff[scn_]) * (B).bs;                                   \
  const int sl_ = bsl_ + (int)threadIdx.x;                                             \
  const bool valid_ = sl_ < S.g.N;                                                     \
  const int w0_ = S.off + (bsl_ / kLaneWin) * kLaneWin;                                \
  const int w1_ = min(w0_ + kLaneWin, S.off + S.g.N);                                  \
  const int s = valid_ ? (lane)[S.off + sl_] : w0_;                                    \
  const Grid &g = S.g;                                                                 \
  const PhysF32 &P = S.P

// descriptors of the list and the per-slot arrays of one state
#define NB_RSRC                                                                \
  const Rsrc nmr = rsrc(nm, (uint32_t)B.N * (uint32_t)(3 * kMaskWords) * 4u);  \
  const Rsrc cst = rsrc(cstart, (uint32_t)(B.ncell + 1) * 4u);                  \
  const Rsrc xrr = rsrc(xr, (uint32_t)B.N * 8u)

// One row of the 3x3 sweep around cell (cx, row jy): the slot run of cells
// max(cx-1,0) .. min(cx+1,nx-1), and the offsets in it where cells cx and
// cx+1 begin (bits below s0 lie in cell cx-1, bits from s1 on in cell cx+1).
struct RowRun {
  int rb, len, s0, s1;
};
__device__ __forceinline__ RowRun row_run(Rsrc cst, int cx, int jy, const Grid &g, int coff) {
  const int base = coff + jy * g.nx;
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx < g.nx - 1 ? cx + 1 : g.nx - 1;
  RowRun r;
  r.rb = (int)ldu(cst, (uint32_t)(base + x0) * 4u);
  r.len = (int)ldu(cst, (uint32_t)(base + x1 + 1) * 4u) - r.rb;
  r.s0 = (int)ldu(cst, (uint32_t)(base + cx) * 4u) - r.rb;
  r.s1 = (int)ldu(cst, (uint32_t)(base + cx + 1) * 4u) - r.rb;
  return r;
}

// The candidates of row r (cell row jy, 0 below the slot's row, 2 above) in
// the 5x5 half-cell sweep of a list build over a quadrant-ordered layout (qs:
// the quadrant starts, sph_cells.hip): cand(b, xs) for each offset b < len of
// the row run in the slot's half-cell block (cell_quadrant), ascending, with
// xs the slot's x in the frame of b's cell.  (qx, qy): the slot's quadrant.
// A row whose both halves are in the block is one range of the run (less the
// half column away from the slot); the other row is the halves next to the
// slot of its two or three cells.
template <class Cand>
__device__ __forceinline__ void half_cell_row(Rsrc qsr, const RowRun &run, int len, int cx,
                                              int jy, int r, int qx, int qy, const Grid &g,
                                              float xm, float xi, float xp, Cand &&cand) {
  const int nx = g.nx;
  if (r == 1 || (r == 0) == (qy == 0)) {
    int a = 0, e = len;
    if (qx && cx > 0) a = (int)ldu(qsr, (uint32_t)(4 * (jy * nx + cx - 1) + 2) * 4u) - run.rb;
    if (!qx && cx < nx - 1) {
      const int t = (int)ldu(qsr, (uint32_t)(4 * (jy * nx + cx + 1) + 2) * 4u) - run.rb;
      e = t < e ? t : e;
    }
    for (int b = a; b < e; ++b) cand(b, b < run.s0 ? xm : (b < run.s1 ? xi : xp));
  } else {
    const int qyr = r == 0 ? 1 : 0;   // the half of row r next to the slot
    const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx < nx - 1 ? cx + 1 : nx - 1;
    for (int cc = x0; cc <= x1; ++cc) {
      const int d = cc - cx;
      const int cmask = d < 0 ? (qx ? 2 : 3) : (d > 0 ? (qx ? 3 : 1) : 3);   // halves in x
      const float xs = d < 0 ? xm : (d > 0 ? xp : xi);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (!((cmask >> h) & 1)) continue;
        const uint32_t qi = (uint32_t)(4 * (jy * nx + cc) + 2 * h + qyr);
        const int a = (int)ldu(qsr, qi * 4u) - run.rb;
        int e = (int)ldu(qsr, (qi + 1u) * 4u) - run.rb;
        e = e < len ? e : len;
        for (int b = a; b < e; ++b) cand(b, xs);
      }
    }
  }
}

// quadrant of slot s in its cell c from the quadrant starts
__device__ __forceinline__ int slot_quadrant(Rsrc qsr, int s, int c) {
  const uint32_t q = (uint32_t)(4 * c);
  return (s >= (int)ldu(qsr, (q + 1u) * 4u) ? 1 : 0) + (s >= (int)ldu(qsr, (q + 2u) * 4u) ? 1 : 0) +
         (s >= (int)ldu(qsr, (q + 3u) * 4u) ? 1 : 0);
}

// Bitmask list: word w (32 bits) of row r of slot s sits at
// nm[(w*3 + r)*N + s]; words beyond ceil(len/32) are neither written nor read.
// The masks of a run longer than kMaskRun slots cover its first kMaskRun
// slots; every pass tests the slots past them directly (for_each_nb), so a
// pile-up in one cell (particles beyond a wall gather in the clamped edge
// cells) leaves no pair out.
constexpr int kMaskRun = 32 * kMaskWords;

// What the walk of slot s reads from memory before its first pair: the three
// rows' runs and (kMasks) each row's first mask word, all independent of one
// another.  A pass loads them in one batch -- before its staging, so that
// their latency overlaps the staging's, or where registers are short right
// after it -- instead of paying two round trips per row in every wavefront.
// A row outside the grid has len = 0; its mask word is loaded (it lies in the
// list) but never used.
struct NbHead {
  RowRun run[3];
  uint32_t m0[3];
  int2 c;
};
template <bool kMasks = true>
__device__ __forceinline__ NbHead nb_head(Rsrc nmr, Rsrc cst, int s, int bc, const Grid &g,
                                          int coff, uint32_t Ntot) {
  NbHead h;
  h.c = cell_xy(bc, g);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int jy = h.c.y + r - 1;
    h.run[r] = (jy >= 0 && jy < g.ny) ? row_run(cst, h.c.x, jy, g, coff) : RowRun{0, 0, 0, 0};
    h.m0[r] = kMasks ? ldu(nmr, ((uint32_t)r * Ntot + (uint32_t)s) * 4u) : 0u;
  }
  return h;
}

// Loader index of a neighbour: slot j of row r is passed to the loader as
// j + ro.o[r].  Global loads use ro = 0 (the index is the slot); staged
// passes use the LDS position of the row's staged segment minus its first
// slot (stage_rowoff).
struct RowOff {
  int o[3];
};

// Calls fn(d, dx, dy, r2) for every neighbour flagged in the mask list m_in
// of slot s (own-frame position xi, rows h from nb_head), in the reference's sweep
// order, where d = ld(i) holds the neighbour's per-slot data (d.x must be its
// xr), i being its loader index (RowOff).  fn returns whether the pair is a
// true neighbour (r2 < (2h)^2); when em is given, those pairs form a mask list
// of the same layout, written to em, that later passes over the same
// positions iterate without testing r2.  The loads of the next neighbour are
// issued before the current pair is evaluated (two registers sets in turn),
// so a word exposes one load latency, not one per pair.  Slots of a row run
// past the kMaskRun slots its masks cover are tested directly (true
// neighbours, r2min < r2 < (2h)^2, at these positions) after the masked ones.
// one pair of for_each_nb: bit b of a word whose cells cx-1 / cx / cx+1 start
// at bits <0 / s0 / s1; xm, x0, xp are the own x in those cells' frames
template <class D, class Fn>
__device__ __forceinline__ void nb_pair(const D &d, int b, int s0, int s1, float xm, float x0,
                                        float xp, float yi, uint32_t &hit, Fn &fn) {
  const float xs = b < s0 ? xm : (b < s1 ? x0 : xp);
  const float dx = xs - d.x.x;
  const float dy = yi - d.x.y;
  if (fn(d, dx, dy, fmaf(dx, dx, dy * dy))) hit |= 1u << b;
}

// kPipe: issue the next neighbour's loads before evaluating the current pair
// (for loaders from global memory); loaders from LDS walk one pair at a time,
// the wavefronts sharing a SIMD hiding the LDS latency.
template <bool kPipe = true, class Ld, class Fn>
__device__ __forceinline__ void nb_row(int r, Rsrc nmr, uint32_t *__restrict__ em, const NbHead &h,
                                       int s, float2 xi, uint32_t Ntot, const PhysF32 &P,
                                       const RowOff &ro, Ld &&ld, Fn &&fn) {
  const RowRun run = h.run[r];
  if (run.len == 0) return;
  const float cs = P.cs;
  const float xm = xi.x + cs, xp = xi.x - cs;   // own x in the frames of cells cx-1, cx+1
  const float yi = xi.y - (float)(r - 1) * cs;
  int nw = (run.len + 31) >> 5;
  nw = nw < kMaskWords ? nw : kMaskWords;
  for (int w = 0; w < nw; ++w) {
    const uint32_t widx = (uint32_t)(w * 3 + r) * Ntot + (uint32_t)s;
    uint32_t m = w == 0 ? h.m0[r] : ldu(nmr, widx * 4u);
    uint32_t hit = 0;
    if (m != 0) {
      const int s0 = run.s0 - 32 * w, s1 = run.s1 - 32 * w;   // word-local cell bounds
      const int wb = run.rb + 32 * w + ro.o[r];
      if constexpr (!kPipe) {
        do {
          const int b = __builtin_ctz(m);
          m &= m - 1;
          nb_pair(ld(wb + b), b, s0, s1, xm, xi.x, xp, yi, hit, fn);
        } while (m != 0);
      } else {
      int b0 = __builtin_ctz(m);
      m &= m - 1;
      auto d0 = ld(wb + b0);
      for (;;) {
        const bool more1 = m != 0;
        const int b1 = more1 ? __builtin_ctz(m) : b0;
        m &= m - 1;
        auto d1 = ld(wb + b1);
        nb_pair(d0, b0, s0, s1, xm, xi.x, xp, yi, hit, fn);
        if (!more1) break;
        const bool more0 = m != 0;
        b0 = more0 ? __builtin_ctz(m) : b1;
        m &= m - 1;
        d0 = ld(wb + b0);
        nb_pair(d1, b1, s0, s1, xm, xi.x, xp, yi, hit, fn);
        if (!more0) break;
      }
      }
    }
    if (em) em[widx] = hit;
  }
  // slots of a row run beyond the kMaskRun slots its masks cover (a pile-up,
  // e.g. particles beyond a wall in a clamped edge cell): tested directly,
  // in slot order after the masked ones, so no pair is left out
  if (run.len > kMaskRun) {
    const int rb = run.rb + ro.o[r];
    for (int b = kMaskRun; b < run.len; ++b) {
      const auto d = ld(rb + b);
      const float xs = b < run.s0 ? xm : (b < run.s1 ? xi.x : xp);
      const float dx = xs - d.x.x;
      const float dy = yi - d.x.y;
      const float r2 = fmaf(dx, dx, dy * dy);
      if (r2 > P.r2min && r2 < P.r2max) fn(d, dx, dy, r2);
    }
  }
}

// all three rows in order (row r = 0 below the slot's cell row, 2 above)
template <bool kPipe = true, class Ld, class Fn>
__device__ __forceinline__ void for_each_nb(Rsrc nmr, uint32_t *__restrict__ em, const NbHead &h,
                                            int s, float2 xi, uint32_t Ntot, const PhysF32 &P,
                                            const RowOff &ro, Ld &&ld, Fn &&fn) {
#pragma unroll
  for (int r = 0; r < 3; ++r) nb_row<kPipe>(r, nmr, em, h, s, xi, Ntot, P, ro, ld, fn);
}

// ---- LDS staging of the neighbour rows of a window (sph_stage.h) -------------
// Capacities in slots.  A 256-slot window in a cell row two lattice rows high
// between rows three high stages ~1100 slots on the 1M block (every eighth
// cell row there, 5 % of the windows); the others stage 800-900.
// kStageCap, for the passes of 8 or 12 bytes a slot, holds them all: those
// passes' occupancy is bound by registers, not LDS.  Force (24 bytes) keeps
// the capacity of six workgroups per CU (kStageCapF) and reads the few larger
// windows from global memory: measured, a capacity of 1120 that holds them
// costs a workgroup per CU and more time than it saves (41 -> 43 us).  Force
// staging one row at a time as P1 does (below) is slower too: 40.8 -> 44.8
// us, its registers then spill at seven waves per SIMD.  Adjoint P1 (32
// bytes) stages one sweep row at a time (kStageCapP1, at adj_p1).
constexpr int kStageCap = 1280;
constexpr int kStageCapF = 1024;

// window plan over the cell starts of a buffer descriptor
template <bool kPhased = false>
__device__ __forceinline__ void stage_plan(StageTab &T, const int *__restrict__ bcell, Rsrc cst,
                                           int w0, int w1, const Grid &g, int coff, int cap) {
  stage_plan_at<kPhased>(T, bcell, [&](int i) { return (int)ldu(cst, (uint32_t)i * 4u); }, w0,
                         w1, g, coff, cap);
}

// loader offsets of slot s (binning cell bc) when its window is staged
__device__ __forceinline__ RowOff stage_rowoff(const StageTab &T, int bc, const Grid &g) {
  const int k = (cell_xy(bc, g).y - T.ya) * 3;
  RowOff r;
#pragma unroll
  for (int i = 0; i < 3; ++i) r.o[i] = T.off[k + i] - T.lo[k + i];
  return r;
}

// neighbour data of the pair passes
struct NbX {
  float2 x;
};
struct NbXPV {
  float2 x, pi, v;
};
struct NbXPAV {
  float2 x, pi, ab, v;
};
struct NbXG {
  float2 x;
  float gr;
};

// Mask-list build: the 3x3 sweep over every slot of the three row runs, or
// with qs (kSub) the 5x5 half-cell sweep over a quadrant-ordered layout (the
// same list: every list neighbour lies in the half-cell block).
template <bool kSub>
__global__ void __launch_bounds__(kBlock)
k_nm_build32(const float2 *__restrict__ xr, const int *__restrict__ bcell,
             const int *__restrict__ cstart, const int *__restrict__ qs, Batch B,
             uint32_t *__restrict__ nm, int *__restrict__ flags) {
  __shared__ StageTab T;
  __shared__ float2 sx[kStageCap];
  const int blk = xcd_block();
  const int scn = B.b2s[blk];
  const Scene32 S = B.sc[scn];
  const Grid &g = S.g;
  const PhysF32 &P = S.P;
  const int bsl = (blk - B.boff[scn]) * B.bs;
  const int sl = bsl + (int)threadIdx.x;
  const int w0 = S.off + bsl, w1 = min(w0 + B.bs, S.off + S.g.N);
  const int s = sl < S.g.N ? S.off + sl : w0;
  const Rsrc xrr = rsrc(xr, (uint32_t)B.N * 8u);
  const Rsrc cst = rsrc(cstart, (uint32_t)(B.ncell + 1) * 4u);
  // the slot's own loads before the staging plan, its row runs after it: both
  // arrive while the window is copied
  const float2 xi = xr[s];
  const int bc = bcell[s];
  stage_plan(T, bcell, cst, w0, w1, g, S.coff, B.nostage ? 0 : kStageCap);
  const NbHead h = nb_head<false>(cst, cst, s, bc, g, S.coff, B.N);
  if (T.ok)
    stage_copy(T, [&](int j) { return ld2(xrr, (uint32_t)j * 8u); },
               [&](int i, float2 v) { sx[i] = v; });
  if (sl >= S.g.N) return;
  const float rc2 = (float)(S.cs64 * S.cs64);
  const int2 c = h.c;
  const float xm = xi.x + P.cs, xp = xi.x - P.cs;
  const Rsrc qsr = rsrc(qs, kSub ? (uint32_t)(4 * B.ncell + 1) * 4u : 0u);
  const int q = kSub ? slot_quadrant(qsr, s, bc) : 0;
  auto rows = [&](const RowOff &ro, auto &&ld) __attribute__((always_inline)) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int jy = c.y + r - 1;
      if (jy < 0 || jy >= g.ny) continue;
      const RowRun run = h.run[r];
      const float yi = xi.y - (float)(r - 1) * P.cs;
      // masks cover the first kMaskRun slots of the run; the passes test the rest
      const int len = run.len < kMaskRun ? run.len : kMaskRun;
      const int rb = run.rb + ro.o[r];
      if constexpr (kSub) {
        uint32_t m = 0;
        int wc = 0;   // the word being formed; candidates come in ascending order
        half_cell_row(qsr, run, len, c.x, jy, r, q >> 1, q & 1, g, xm, xi.x, xp,
                      [&](int b, float xs) __attribute__((always_inline)) {
                        for (; wc < (b >> 5); ++wc) {
                          nm[(size_t)(wc * 3 + r) * B.N + s] = m;
                          m = 0;
                        }
                        const float2 xj = ld(rb + b);
                        const float dx = xs - xj.x;
                        const float dy = yi - xj.y;
                        const float r2 = fmaf(dx, dx, dy * dy);
                        if (r2 < rc2 && r2 > P.r2min) m |= 1u << (b & 31);
                      });
        for (; wc * 32 < len; ++wc) {
          nm[(size_t)(wc * 3 + r) * B.N + s] = m;
          m = 0;
        }
      } else {
        for (int w = 0; w * 32 < len; ++w) {
          uint32_t m = 0;
          const int e = len - 32 * w < 32 ? len - 32 * w : 32;
          for (int k = 0; k < e; ++k) {
            const int b = 32 * w + k;
            const float2 xj = ld(rb + b);
            const float xs = b < run.s0 ? xm : (b < run.s1 ? xi.x : xp);
            const float dx = xs - xj.x;
            const float dy = yi - xj.y;
            const float r2 = fmaf(dx, dx, dy * dy);
            if (r2 < rc2 && r2 > P.r2min) m |= 1u << k;
          }
          nm[(size_t)(w * 3 + r) * B.N + s] = m;
        }
      }
    }
  };
  if (T.ok)
    rows(stage_rowoff(T, bc, g), [&](int i) __attribute__((always_inline)) { return sx[i]; });
  else
    rows(RowOff{}, [&](int i) __attribute__((always_inline)) {
      return ld2(xrr, (uint32_t)i * 8u);
    });
}

// Thread -> slot assignment of the pair passes.  A pair pass walks the three
// cell rows of a slot's 3x3 block one after the other, so a wavefront spends
// on each row the longest row count of its 64 lanes.  How a slot's neighbours
// split over the rows depends on its height in its cell (near the bottom most
// lie in the row below, near the top in the row above), and consecutive slots
// of a cell climb through it (ids ascend with y on a lattice).  Each aligned
// window of kLaneWin slots of a scene is therefore dealt to its threads in
// order of that height (16 classes, then slot order), so that each wavefront
// holds slots of similar height and similar row counts.  The order is fixed
// (ballot ranks, no atomics) and changes which thread evaluates a slot, never
// what is evaluated.  The kernels that form a window's relative coordinates
// (k_permute32 at a rebuild, k_xrel at an epoch of the reverse sweep) deal it
// on the way: their workgroups of kLaneWin threads are the windows.
constexpr int kLaneClasses = 16;

// The window [base, base + n) (0 < n <= kLaneWin) of a workgroup of kLaneWin
// threads: thread t < n gives the relative height yr of slot base + t.  Every
// thread of the workgroup calls it (it holds a barrier).
__device__ __forceinline__ void lane_order_window(float yr, float cs, int base, int n,
                                                  int *__restrict__ lane) {
  __shared__ int wcnt[kLaneWin / 64][kLaneClasses];
  const int t = (int)threadIdx.x, w = t >> 6, l = t & 63;
  int key = kLaneClasses;   // threads past the window's end take no class
  if (t < n) {
    const float y = yr * ((float)kLaneClasses / cs);
    key = !(y > 0.f) ? 0 : (y >= (float)(kLaneClasses - 1) ? kLaneClasses - 1 : (int)y);
  }
  int below = 0;   // lanes of this wavefront below this one with the same class
  for (int c = 0; c < kLaneClasses; ++c) {
    const unsigned long long m = __ballot(key == c);
    if (l == 0) wcnt[w][c] = __popcll(m);
    if (key == c) below = __popcll(m & ((1ull << l) - 1ull));
  }
  __syncthreads();
  if (t < n) {
    int pos = below;
    for (int c = 0; c < key; ++c)
      for (int v = 0; v < kLaneWin / 64; ++v) pos += wcnt[v][c];
    for (int v = 0; v < w; ++v) pos += wcnt[v][key];
    lane[base + pos] = base + t;
  }
}

// W/(m A) = u^4 (2q + 1) of a pair at squared distance r2 < (2h)^2
__device__ __forceinline__ float density_term(float r2, const PhysF32 &P) {
  const float r = __builtin_amdgcn_sqrtf(r2);
  const float u = fmaf(r, P.mhh, 1.f);
  const float u2 = u * u;
  return u2 * u2 * fmaf(r, P.twoinvh, 1.f);
}
// rho = mA (s + c): the product's rounding error by an FMA, then rho - rho0
// with the exact subtraction first and the tails after; returns (rho, rho - rho0)
__device__ __forceinline__ float2 density_finish(const FSum &acc, const PhysF32 &P) {
  const float p = P.mA * acc.s;
  const float pe = fmaf(P.mA, acc.s, -p);
  const float drho = (p - P.rho0) + fmaf(P.mA, acc.c, pe);
  return make_float2(P.rho0 + drho, drho);
}

template <bool kPipe, class Ld>
__device__ __forceinline__ float2 density_pair(Rsrc nmr, uint32_t *em, const NbHead &h, int s,
                                               float2 xi, uint32_t Ntot, const PhysF32 &P,
                                               const RowOff &ro, Ld &&ld) {
  // sum of W/(m A) = u^4 (2q + 1) by Fast2Sum from the self term 1, the
  // largest term of the sum: every later term is positive and no larger, so
  // each step's error term is exact
  FSum acc(1.f);
  for_each_nb<kPipe>(nmr, em, h, s, xi, Ntot, P, ro, ld,
              [&, P](const NbX &, float, float, float r2) __attribute__((always_inline)) {
    if (r2 < P.r2max) {
      acc.add(density_term(r2, P));
      return true;
    }
    return false;
  });
  return density_finish(acc, P);
}

// density over the staged rows when the window fits, else over global loads
__device__ __forceinline__ float2 density_slot(Rsrc nmr, uint32_t *em, const NbHead &h, Rsrc xrr,
                                               const StageTab &T, const float2 *sx, int s, int bc,
                                               float2 xi, const Grid &g, uint32_t Ntot,
                                               const PhysF32 &P) {
  if (T.ok)
    return density_pair<false>(nmr, em, h, s, xi, Ntot, P, stage_rowoff(T, bc, g),
                        [&](int i) __attribute__((always_inline)) { return NbX{sx[i]}; });
  return density_pair<true>(nmr, em, h, s, xi, Ntot, P, RowOff{},
                      [&](int i) __attribute__((always_inline)) {
                        return NbX{ld2(xrr, (uint32_t)i * 8u)};
                      });
}


__device__ __forceinline__ float2 eos(float2 rd, const PhysF32 &P) {
  float inv = __builtin_amdgcn_rcpf(rd.x);
  return make_float2(P.c2 * rd.y * inv * inv, inv);    // (pterm, invrho)
}

__global__ void __launch_bounds__(kBlock)
k_density32(const float2 *__restrict__ xr, const int *__restrict__ bcell,
            const uint32_t *__restrict__ nm, const int *__restrict__ cstart,
            const int *__restrict__ lane, Batch B, float2 *__restrict__ PI,
            double *__restrict__ rho_out, uint32_t *__restrict__ em) {
  __shared__ StageTab T;
  __shared__ float2 sx[kStageCap];
  SCENE_WIN(B, lane);
  NB_RSRC;
  // the slot's own loads before the staging plan, its walk's row heads after
  // it: both arrive while the window is copied
  const int bc = bcell[s];
  const float2 xi = xr[s];
  stage_plan(T, bcell, cst, w0_, w1_, g, S.coff, B.nostage ? 0 : kStageCap);
  const NbHead h = nb_head(nmr, cst, s, bc, g, S.coff, B.N);
  if (T.ok)
    stage_copy(T, [&](int j) { return ld2(xrr, (uint32_t)j * 8u); },
               [&](int i, float2 v) { sx[i] = v; });
  if (!valid_) return;
  float2 rd = density_slot(nmr, em, h, xrr, T, sx, s, bc, xi, g, B.N, P);
  PI[s] = eos(rd, P);
  if (rho_out) rho_out[s] = (double)rd.x;
}

// The first density pass after a rebuild, fused with the mask-list build:
// one sweep over every slot of the three row runs sets the mask bit of each
// list neighbour (r2min < r2 < rc2, as k_nm_build32), and for those within
// (2h)^2 adds the density term and sets the exact-list bit, in the order in
// which k_density32 would visit them -- the same pairs, the same arithmetic,
// the same sums.  Writes nm, em, PI and the overflow flag.  With qs (kSub)
// the sweep is the 5x5 half-cell one of k_nm_build32<true>, which visits the
// list neighbours in the same ascending order.
template <bool kSub>
__global__ void __launch_bounds__(kBlock)
k_density_build32(const float2 *__restrict__ xr, const int *__restrict__ bcell,
                  const int *__restrict__ cstart, const int *__restrict__ qs,
                  const int *__restrict__ lane, Batch B, float2 *__restrict__ PI,
                  uint32_t *__restrict__ nm, uint32_t *__restrict__ em,
                  int *__restrict__ flags) {
  __shared__ StageTab T;
  __shared__ float2 sx[kStageCap];
  SCENE_WIN(B, lane);
  const Rsrc cst = rsrc(cstart, (uint32_t)(B.ncell + 1) * 4u);
  const Rsrc xrr = rsrc(xr, (uint32_t)B.N * 8u);
  const Rsrc qsr = rsrc(qs, kSub ? (uint32_t)(4 * B.ncell + 1) * 4u : 0u);
  // the slot's own loads before the staging plan, its row runs after it: both
  // arrive while the window is copied
  const float2 xi = xr[s];
  const int bc = bcell[s];
  stage_plan(T, bcell, cst, w0_, w1_, g, S.coff, B.nostage ? 0 : kStageCap);
  const NbHead h = nb_head<false>(cst, cst, s, bc, g, S.coff, B.N);
  if (T.ok)
    stage_copy(T, [&](int j) { return ld2(xrr, (uint32_t)j * 8u); },
               [&](int i, float2 v) { sx[i] = v; });
  if (!valid_) return;
  const float rc2 = (float)(S.cs64 * S.cs64);
  const int2 c = h.c;
  const float xm = xi.x + P.cs, xp = xi.x - P.cs;
  const int q = kSub ? slot_quadrant(qsr, s, bc) : 0;
  FSum acc(1.f);
  auto rows = [&](const RowOff &ro, auto &&ld) __attribute__((always_inline)) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int jy = c.y + r - 1;
      if (jy < 0 || jy >= g.ny) continue;
      const RowRun run = h.run[r];
      const float yi = xi.y - (float)(r - 1) * P.cs;
      const int len = run.len < kMaskRun ? run.len : kMaskRun;
      const int rb = run.rb + ro.o[r];
      if constexpr (kSub) {
        uint32_t m = 0, e = 0;
        int wc = 0;   // the word being formed; candidates come in ascending order
        auto flush = [&](int upto) __attribute__((always_inline)) {
          for (; wc < upto; ++wc) {
            const size_t widx = (size_t)(wc * 3 + r) * B.N + s;
            nm[widx] = m;
            em[widx] = e;
            m = 0;
            e = 0;
          }
        };
        half_cell_row(qsr, run, len, c.x, jy, r, q >> 1, q & 1, g, xm, xi.x, xp,
                      [&](int b, float xs) __attribute__((always_inline)) {
                        flush(b >> 5);
                        const float2 xj = ld(rb + b);
                        const float dx = xs - xj.x;
                        const float dy = yi - xj.y;
                        const float r2 = fmaf(dx, dx, dy * dy);
                        if (r2 < rc2 && r2 > P.r2min) {
                          m |= 1u << (b & 31);
                          if (r2 < P.r2max) {
                            acc.add(density_term(r2, P));
                            e |= 1u << (b & 31);
                          }
                        }
                      });
        flush((len + 31) >> 5);
      } else {
        for (int w = 0; w * 32 < len; ++w) {
          uint32_t m = 0, e = 0;
          const int n = len - 32 * w < 32 ? len - 32 * w : 32;
          for (int k = 0; k < n; ++k) {
            const int b = 32 * w + k;
            const float2 xj = ld(rb + b);
            const float xs = b < run.s0 ? xm : (b < run.s1 ? xi.x : xp);
            const float dx = xs - xj.x;
            const float dy = yi - xj.y;
            const float r2 = fmaf(dx, dx, dy * dy);
            if (r2 < rc2 && r2 > P.r2min) {
              m |= 1u << k;
              if (r2 < P.r2max) {
                acc.add(density_term(r2, P));
                e |= 1u << k;
              }
            }
          }
          const size_t widx = (size_t)(w * 3 + r) * B.N + s;
          nm[widx] = m;
          em[widx] = e;
        }
      }
      // the run's slots past the masks, as the passes over these lists take them
      for (int b = kMaskRun; b < run.len; ++b) {
        const float2 xj = ld(rb + b);
        const float xs = b < run.s0 ? xm : (b < run.s1 ? xi.x : xp);
        const float dx = xs - xj.x;
        const float dy = yi - xj.y;
        const float r2 = fmaf(dx, dx, dy * dy);
        if (r2 > P.r2min && r2 < P.r2max) acc.add(density_term(r2, P));
      }
    }
  };
  if (T.ok)
    rows(stage_rowoff(T, bc, g), [&](int i) __attribute__((always_inline)) { return sx[i]; });
  else
    rows(RowOff{}, [&](int i) __attribute__((always_inline)) {
      return ld2(xrr, (uint32_t)i * 8u);
    });
  PI[s] = eos(density_finish(acc, P), P);
}

__global__ void __launch_bounds__(kBlock)
k_force32(const double2 *__restrict__ X, const float2 *__restrict__ xr,
          const float2 *__restrict__ V, const float2 *__restrict__ PI,
          const double *__restrict__ theta, const int *__restrict__ bcell,
          const uint32_t *__restrict__ nm, const int *__restrict__ cstart,
          const int *__restrict__ lane, Batch B, double2 *__restrict__ Xn,
          float2 *__restrict__ xrn, float2 *__restrict__ Vn, double2 *__restrict__ tapeX,
          float2 *__restrict__ tapeV, double2 *__restrict__ a_out,
          double *__restrict__ alpha_out) {
  __shared__ StageTab T;
  // a pair reads (x, pi) as one 16-byte record and v beside it
  __shared__ float4 sxp[kStageCapF];
  __shared__ float2 sv[kStageCapF];
  SCENE_WIN(B, lane);
  NB_RSRC;
  // the slot's own loads and its drag coefficient first: they arrive while
  // the window is staged (the row heads, loaded after it, in one round trip)
  const double2 x64 = X[s];
  const float2 xi = xr[s], vi = V[s], pii = PI[s];
  const int bc = bcell[s];
  const float al = interp_alpha(theta + S.toff, interp_cell(x64.x, x64.y, P), P, nullptr, nullptr);
  stage_plan(T, bcell, cst, w0_, w1_, g, S.coff, B.nostage ? 0 : kStageCapF);
  const Rsrc pir = rsrc(PI, (uint32_t)B.N * 8u);
  const Rsrc vr = rsrc(V, (uint32_t)B.N * 8u);
  if (T.ok)
    stage_copy(T,
               [&](int j) {
                 const uint32_t o8 = (uint32_t)j * 8u;   // one offset for the three arrays
                 return NbXPV{ld2(xrr, o8), ld2(pir, o8), ld2(vr, o8)};
               },
               [&](int i, const NbXPV &d) {
                 sxp[i] = make_float4(d.x.x, d.x.y, d.pi.x, d.pi.y);
                 sv[i] = d.v;
               });
  if (!valid_) return;
  const NbHead h = nb_head(nmr, cst, s, bc, g, S.coff, B.N);
  // pair term F (C (v_i - v_j) - m (pt_i + pt_j) d), F = Fc u^3, with Fc and
  // m folded into the own-particle factors
  const float piF = P.mFc * pii.x;
  const float tmF = (P.Fc * P.twomumass) * pii.y;
  CSum ax, ay;
  auto pair = [&, P](const NbXPV &d, float dx, float dy, float r2) __attribute__((always_inline)) {
    const float u = fmaf(__builtin_amdgcn_sqrtf(r2), P.mhh, 1.f);
    const float u3 = u * u * u;
    const float2 pij = d.pi;
    const float PF = fmaf(P.mFc, pij.x, piF) * u3;
    const float C = (tmF * pij.y) * u3;
    const float2 vj = d.v;
    ax.add(fmaf(C, vi.x - vj.x, -(PF * dx)));
    ay.add(fmaf(C, vi.y - vj.y, -(PF * dy)));
    return true;
  };
  // nm is the exact list of the density pass over these positions
  if (T.ok) {
    for_each_nb<false>(nmr, nullptr, h, s, xi, B.N, P, stage_rowoff(T, bc, g),
                [&](int i) __attribute__((always_inline)) {
                  const float4 q = sxp[i];
                  return NbXPV{make_float2(q.x, q.y), make_float2(q.z, q.w), sv[i]};
                },
                pair);
  } else {
    for_each_nb(nmr, nullptr, h, s, xi, B.N, P, RowOff{},
                [&](int i) __attribute__((always_inline)) {
                  const uint32_t o8 = (uint32_t)i * 8u;
                  return NbXPV{ld2(xrr, o8), ld2(pir, o8), ld2(vr, o8)};
                },
                pair);
  }
  float axv = ax.value() + (P.gx + wall_accel((float)x64.x, P.Lx, P.kw));
  float ayv = ay.value() + (P.gy + wall_accel((float)x64.y, P.Ly, P.kw));
  if (a_out) {
    a_out[s] = make_double2(axv, ayv);
    alpha_out[s] = al;
    return;
  }
  float D = __builtin_amdgcn_rcpf(fmaf(P.dt, al, 1.f));
  float v1 = (vi.x + P.dt * axv) * D;
  float v2 = (vi.y + P.dt * ayv) * D;
  Vn[s] = make_float2(v1, v2);
  double2 xn = make_double2(x64.x + P.dt64 * (double)v1, x64.y + P.dt64 * (double)v2);
  Xn[s] = xn;
  xrn[s] = rel_coord(xn, bc, g, P);
  if (tapeX) {
    tapeX[s] = x64;
    tapeV[s] = vi;
  }
}

// relative coordinates of the slots; with `lane`, also their thread order
// (the workgroups are then the kLaneWin-slot windows)
__global__ void __launch_bounds__(kBlock)
k_xrel(const double2 *__restrict__ X, const int *__restrict__ bcell, Batch B,
       float2 *__restrict__ xr, int *__restrict__ lane) {
  const int blk = xcd_block();
  const int scn = B.b2s[blk];
  const Scene32 S = B.sc[scn];
  const int bsl = (blk - B.boff[scn]) * B.bs;
  const int sl = bsl + (int)threadIdx.x;
  float yr = 0.f;
  if (sl < S.g.N) {
    const int s = S.off + sl;
    const float2 r = rel_coord(X[s], bcell[s], S.g, S.P);
    xr[s] = r;
    yr = r.y;
  }
  if (lane) lane_order_window(yr, S.P.cs, S.off + bsl, min(B.bs, S.g.N - bsl), lane);
}

// the slot reorder of a rebuild; with `lane`, also the new thread order
__global__ void __launch_bounds__(kBlock)
k_permute32(const int *__restrict__ perm, const int *__restrict__ cellof,
            const double2 *__restrict__ X, const float2 *__restrict__ V,
            const double2 *__restrict__ Xref, int check, Batch B, double2 *__restrict__ Xn,
            float2 *__restrict__ Vn, float2 *__restrict__ xr, double2 *__restrict__ Xrefn,
            int *__restrict__ bcell, int *__restrict__ flags, int *__restrict__ lane) {
  // slots keep their scene through the reorder: scenes own contiguous cell
  // ranges, and cells own contiguous slot ranges
  const int blk = xcd_block();
  const int scn = B.b2s[blk];
  const Scene32 S = B.sc[scn];
  const int bsl = (blk - B.boff[scn]) * B.bs;
  const int sl = bsl + (int)threadIdx.x;
  float yr = 0.f;
  if (sl < S.g.N) {
    const int t = S.off + sl;
    int o = perm[t];
    double2 x = X[o];
    if (check) {
      double2 r = Xref[o];
      double ddx = x.x - r.x, ddy = x.y - r.y;
      if (ddx * ddx + ddy * ddy > S.dmax2) flags[1] = 1;
    }
    int c = cellof[o] - S.coff;     // the scene's own cell numbering
    Xn[t] = x;
    Vn[t] = V[o];
    Xrefn[t] = x;
    bcell[t] = c;
    const float2 r = rel_coord(x, c, S.g, S.P);
    xr[t] = r;
    yr = r.y;
  }
  if (lane) lane_order_window(yr, S.P.cs, S.off + bsl, min(B.bs, S.g.N - bsl), lane);
}

// A checkpoint snapshot of a rebuilt single-scene layout, in one pass:
// to_snap copies (X, V, ids, binning cells, thread order, cell starts) into
// the snapshot; otherwise they are copied back, X also into the reference
// positions, and the relative coordinates are formed.  Thread i covers slot i
// and cell-start entry i.
__global__ void __launch_bounds__(kBlock)
k_snapshot32(int to_snap, Batch B, double2 *__restrict__ X, float2 *__restrict__ V,
             int *__restrict__ ids, int *__restrict__ scell, int *__restrict__ lane,
             int *__restrict__ cstart, double2 *__restrict__ Xref, float2 *__restrict__ xr,
             double2 *__restrict__ qX, float2 *__restrict__ qV, int *__restrict__ qids,
             int *__restrict__ qscell, int *__restrict__ qlane, int *__restrict__ qcstart) {
  const int i = blockIdx.x * kBlock + (int)threadIdx.x;
  const Scene32 &S = B.sc[0];
  if (i < B.N) {
    if (to_snap) {
      qX[i] = X[i];
      qV[i] = V[i];
      qids[i] = ids[i];
      qscell[i] = scell[i];
      qlane[i] = lane[i];
    } else {
      const double2 x = qX[i];
      const int c = qscell[i];
      X[i] = x;
      Xref[i] = x;
      V[i] = qV[i];
      ids[i] = qids[i];
      scell[i] = c;
      lane[i] = qlane[i];
      xr[i] = rel_coord(x, c, S.g, S.P);
    }
  }
  if (i <= B.ncell) {
    if (to_snap) qcstart[i] = cstart[i];
    else cstart[i] = qcstart[i];
  }
}

// cell of every slot (the scene's grid, the division rule of cell_index,
// src/neighbors.jl:38-42) in the global cell numbering, counts by
// wave-aggregated atomics
__global__ void __launch_bounds__(kBlock)
k_assign32(const double2 *__restrict__ X, Batch B, int *__restrict__ cellof,
           int *__restrict__ counts, int *__restrict__ flags) {
  const int blk = blockIdx.x;     // all lanes of the workgroup take part in the run search
  const int scn = B.b2s[blk];
  const Scene32 S = B.sc[scn];
  const int sl = (blk - B.boff[scn]) * B.bs + (int)threadIdx.x;
  int c = -1;
  if (sl < S.g.N) {
    const int s = S.off + sl;
    const double2 x = X[s];
    if (!isfinite(x.x) || !isfinite(x.y)) flags[2] = 1;   // binned into a clamped cell
    const int cx = cell_coord_div(x.x / S.cs64, S.g.nx);
    const int cy = cell_coord_div(x.y / S.cs64, S.g.ny);
    c = S.coff + cx + cy * S.g.nx;
    cellof[s] = c;
  }
  int head_lane, run_len;
  wave_runs(c, head_lane, run_len);
  const int lane = threadIdx.x & 63;
  if (c >= 0 && head_lane == lane) atomicAdd(&counts[c], run_len);
}

// Raises this step's gtheta sweep level when slot s (binning cell bc) lies
// beyond the tight SPH range of its interp cell ic (level 1: the one-cell
// ranges), or beyond the one-cell range (the margin that reaches it), so that
// the step's gather visits every slot (a level of an earlier step's stamp
// selects nothing).
__device__ __forceinline__ void check_level(const InterpF &ic, int bc, const Grid &g,
                                            const PhysF32 &P, const Scene32 &S, const Batch &B,
                                            unsigned long long *tlev, int tstamp) {
  const int i0 = ic.l00 / P.ngy, j0 = ic.l00 - i0 * P.ngy;
  const int2 c = cell_xy(bc, g);
  const int rowoff = S.troff + (P.ngx - 1);
  const int2 rx = B.trange_tight[S.troff + i0];
  const int2 ry = B.trange_tight[rowoff + j0];
  if (c.x < rx.x || c.x > rx.y || c.y < ry.x || c.y > ry.y) {
    const int m = max(margin_beyond(c.x, B.trange[S.troff + i0]),
                      margin_beyond(c.y, B.trange[rowoff + j0]));
    raise_level(tlev, tstamp, m > 1 ? m : 1);
  }
}

// adjoint R: density recompute, _adj_integrate!, _adj_drag! (src/adjoint.jl:193-228)
__global__ void __launch_bounds__(kBlock)
k_adj_r32(const double2 *__restrict__ X, const float2 *__restrict__ xr,
          const float2 *__restrict__ Vnext, const double *__restrict__ theta,
          const int *__restrict__ bcell, const uint32_t *__restrict__ nm,
          const int *__restrict__ cstart, const int *__restrict__ lane, Batch B,
          float2 *__restrict__ PI, float2 *__restrict__ gX, float2 *__restrict__ gV,
          float2 *__restrict__ abar, float *__restrict__ galpha, uint32_t *__restrict__ em,
          int *__restrict__ tkey, unsigned long long *__restrict__ tlev, int tstamp) {
  __shared__ StageTab T;
  __shared__ float2 sx[kStageCap];
  SCENE_WIN(B, lane);
  NB_RSRC;
  // own loads before the staging plan, row heads after it (as k_density32)
  const double2 x64 = X[s];
  const int bc = bcell[s];
  const float2 xi = xr[s];
  stage_plan(T, bcell, cst, w0_, w1_, g, S.coff, B.nostage ? 0 : kStageCap);
  const NbHead h = nb_head(nmr, cst, s, bc, g, S.coff, B.N);
  if (T.ok)
    stage_copy(T, [&](int j) { return ld2(xrr, (uint32_t)j * 8u); },
               [&](int i, float2 v) { sx[i] = v; });
  if (!valid_) return;
  float2 rd = density_slot(nmr, em, h, xrr, T, sx, s, bc, xi, g, B.N, P);
  PI[s] = eos(rd, P);
  InterpF ic = interp_cell(x64.x, x64.y, P);
  tkey[s] = S.tcoff + ic.l00 - ic.l00 / P.ngy;   // scene's theta cell j0 + i0*(ngy-1)
  check_level(ic, bcell[s], g, P, S, B, tlev, tstamp);
  float al = interp_alpha(theta + S.toff, ic, P, nullptr, nullptr);
  float2 gx = gX[s], gv = gV[s], vp = Vnext[s];
  float gvx = gv.x + P.dt * gx.x;
  float gvy = gv.y + P.dt * gx.y;
  float D = __builtin_amdgcn_rcpf(fmaf(P.dt, al, 1.f));
  float dtD = P.dt * D;
  abar[s] = make_float2(dtD * gvx, dtD * gvy);
  galpha[s] = -dtD * (vp.x * gvx + vp.y * gvy);
  gV[s] = make_float2(D * gvx, D * gvy);
}

// adjoint R without the density recompute (tapes that kept each step's
// pterm/invrho) for slot s with binning cell bc, given the slot's gX: also
// forms the step's relative coordinates xr
// the R pass's own-slot inputs; a P2 launch that runs the pass loads them
// before its pair loop, so that they arrive while the loop computes
struct RpwIn {
  double2 x64;
  float2 gv, vp;
};
__device__ __forceinline__ RpwIn rpw_load(const RpwArgs &R, int s) {
  return RpwIn{R.X[s], R.gV[s], R.Vnext[s]};
}
__device__ __forceinline__ void rpw_finish(const RpwArgs &R, const RpwIn &in, int s, int bc,
                                           float2 gx, const Scene32 &S, const Batch &B) {
  const PhysF32 &P = S.P;
  const Grid &g = S.g;
  const double2 x64 = in.x64;
  R.xr[s] = rel_coord(x64, bc, g, P);
  InterpF ic = interp_cell(x64.x, x64.y, P);
  R.tkey[s] = S.tcoff + ic.l00 - ic.l00 / P.ngy;   // scene's theta cell j0 + i0*(ngy-1)
  check_level(ic, bc, g, P, S, B, R.tlev, R.tstamp);
  float al = interp_alpha(R.theta + S.toff, ic, P, nullptr, nullptr);
  const float2 gv = in.gv, vp = in.vp;
  float gvx = gv.x + P.dt * gx.x;
  float gvy = gv.y + P.dt * gx.y;
  float D = __builtin_amdgcn_rcpf(fmaf(P.dt, al, 1.f));
  float dtD = P.dt * D;
  R.abar[s] = make_float2(dtD * gvx, dtD * gvy);
  R.galpha[s] = -dtD * (vp.x * gvx + vp.y * gvy);
  R.gV[s] = make_float2(D * gvx, D * gvy);
}
__device__ __forceinline__ void rpw_slot(const RpwArgs &R, int s, int bc, float2 gx,
                                         const Scene32 &S, const Batch &B) {
  rpw_finish(R, rpw_load(R, s), s, bc, gx, S, B);
}

__global__ void __launch_bounds__(kBlock)
k_adj_rpw32(RpwArgs R, const int *__restrict__ bcell, const float2 *__restrict__ gX, Batch B) {
  const int blk = xcd_block();
  const int scn = B.b2s[blk];
  const Scene32 S = B.sc[scn];
  const int sl = (blk - B.boff[scn]) * B.bs + (int)threadIdx.x;
  if (sl >= S.g.N) return;
  const int s = S.off + sl;
  rpw_slot(R, s, bcell[s], gX[s], S, B);
}

// The pointwise R pass of a reverse step that opens an epoch, fused with the
// gather that brings the adjoint state (and with kRemapV V_{k+1}) from
// another slot order into the epoch's: slot s takes gX, gV (and V_{k+1}) from
// slot inv[ids[s]] of that order (inv: the inverse of its ids), keeps gX in
// gXn and runs rpw_finish on them, R.gV receiving this order's gV (a buffer
// other than gVo).  The same values as the remap gathers followed by
// k_adj_rpw32.
template <bool kRemapV>
__global__ void __launch_bounds__(kBlock)
k_remap_rpw32(RpwArgs R, const int *__restrict__ bcell, const float2 *__restrict__ gXo,
              const float2 *__restrict__ gVo, const float2 *__restrict__ Vno,
              const int *__restrict__ inv, const int *__restrict__ ids, Batch B,
              float2 *__restrict__ gXn) {
  const int blk = xcd_block();
  const int scn = B.b2s[blk];
  const Scene32 S = B.sc[scn];
  const int sl = (blk - B.boff[scn]) * B.bs + (int)threadIdx.x;
  if (sl >= S.g.N) return;
  const int s = S.off + sl;
  const int o = inv[ids[s]];
  const float2 gx = gXo[o];
  gXn[s] = gx;
  rpw_finish(R, RpwIn{R.X[s], gVo[o], Vno[kRemapV ? o : s]}, s, bcell[s], gx, S, B);
}

// ---- gtheta gather over the keys of the pointwise pass -----------------------

// The gather of sph_theta.hip over per-slot keys written by the f32c adjoint's
// pointwise pass: key[p] = (the scene's first theta cell) + j0 + i0*(ngy-1)
// of the slot's interp cell, so that a candidate costs one key compare; a
// match's FP64 fractions (fx, fy) are formed from its position exactly as the
// f32c interp_cell forms them (x * 1/hx, clamped, minus the cell index); the
// contributions and their order are those of k_theta_cells_keyed.
// A workgroup takes a strip of kThetaStrip theta cells of one theta row,
// kThetaLanes lanes per cell; lane l of a cell sweeps every kThetaLanes-th
// slot (from the l-th) of each SPH cell row of the cell's range, the rows in
// order, and a butterfly adds the lanes' sums.  The gather is bound by its
// dependent memory rounds, so a lane reads the cell starts of all its rows at
// once (kLeanRows rows a round), then the keys of up to kLeanKeys of its
// candidates over those rows together, then the positions and galpha of the
// matches among them kLeanBatch at a time.  It holds no LDS, so a launch that
// runs it beside P1 keeps P1's occupancy.  Workgroups are dealt to the scenes
// of a batched state in order (Scene32::tboff).
// (Measured on the 1M block, P1 + gather in one launch: 68.7 us for a version
// that first staged the strip's keys and cell starts in 31 KB of LDS, which
// held P1 at five waves per SIMD; 60.5 us reading keys from global memory
// row by row, four at a time; 57.9 us with the keys of all rows in one
// batch.  Loading the range tables beside the sweep level: no change.)
constexpr int kLeanKeys = 16;
constexpr int kLeanBatch = 4;
constexpr int kLeanRows = 4;
__device__ __forceinline__ void theta_strip(int blk, const ThetaKeyed &A, const Batch &B) {
#pragma clang fp contract(off)
  const int *__restrict__ key = A.key;
  const double2 *__restrict__ X = A.X;
  const float *__restrict__ galpha = A.galpha;
  const int *__restrict__ cstart = A.cstart;
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
  const int lane = threadIdx.x & (kThetaLanes - 1);
  const int i0 = ifirst + (int)threadIdx.x / kThetaLanes;
  if (i0 > ilast) return;          // the kThetaLanes lanes of a group leave together
  // SPH cell ranges holding particles with interp cell (i, j0), tabulated per
  // scene: a quarter cell of margin (theta_cell_range_tight) unless the R pass
  // found a slot further out this step (level 1: one cell, theta_cell_range;
  // level m >= 2: the one-cell range widened by m - 1 cells on each side)
  const int lv = level_of(A.tlev, A.tstamp);
  const int2 *tr = (lv >= 1 ? B.trange : B.trange_tight) + Sc.troff;
  const int wid = lv > 1 ? lv - 1 : 0;
  const int2 rxw = tr[i0], ryw = tr[ncx + j0];
  const int cx0 = max(rxw.x - wid, 0), cx1 = min(rxw.y + wid, g.nx - 1);
  const int cy0 = max(ryw.x - wid, 0), cy1 = min(ryw.y + wid, g.ny - 1);
  const int mykey = tcoff + j0 + i0 * ncy;
  const PhysF32 &P = Sc.P;
  auto frac = [&](double2 x) {
    double gx = x.x * P.inv_hx64, gy = x.y * P.inv_hy64;
    gx = (gx > P.gxclamp) ? P.gxclamp : ((gx < 0.0) ? 0.0 : gx);
    gy = (gy > P.gyclamp) ? P.gyclamp : ((gy < 0.0) ? 0.0 : gy);
    return make_double2(gx - (double)i0, gy - (double)j0);
  };
  double s00 = 0.0, s10 = 0.0, s01 = 0.0, s11 = 0.0;
  auto add = [&](double2 f, double ga) {
    s00 += (ga * (1.0 - f.x)) * (1.0 - f.y);
    s10 += (ga * f.x) * (1.0 - f.y);
    s01 += (ga * (1.0 - f.x)) * f.y;
    s11 += (ga * f.x) * f.y;
  };
  // the lane's candidates of kLeanRows rows, in sweep order: candidate t of
  // row q is slot rb[q] + kThetaLanes t (t < nq[q]), numbered k = pre[q] + t
  for (int ya = cy0; ya <= cy1; ya += kLeanRows) {
    int rb[kLeanRows], pre[kLeanRows];
    int total = 0;
#pragma unroll
    for (int q = 0; q < kLeanRows; ++q) {
      const int y = ya + q;
      int b = 0, e = 0;
      if (y <= cy1) {
        const int base = coff + y * g.nx;
        b = cstart[base + cx0];
        e = cstart[base + cx1 + 1];
      }
      rb[q] = b + lane;
      pre[q] = total;
      total += e > rb[q] ? (e - rb[q] + kThetaLanes - 1) / kThetaLanes : 0;
    }
    auto slot_of = [&](int k) {   // slot of candidate k < total
      int q = 0;
#pragma unroll
      for (int r = 1; r < kLeanRows; ++r) q += (k >= pre[r]) ? 1 : 0;
      return rb[q] + kThetaLanes * (k - pre[q]);
    };
    // kLeanKeys keys at a time, then the matches among them kLeanBatch at a time
    for (int k0 = 0; k0 < total; k0 += kLeanKeys) {
      int kv[kLeanKeys];
#pragma unroll
      for (int u = 0; u < kLeanKeys; ++u) kv[u] = k0 + u < total ? key[slot_of(k0 + u)] : -1;
      uint32_t hits = 0;
#pragma unroll
      for (int u = 0; u < kLeanKeys; ++u) hits |= (kv[u] == mykey ? 1u : 0u) << u;
      while (hits != 0) {
        bool ok[kLeanBatch];
        double2 f[kLeanBatch];
        double ga[kLeanBatch];
#pragma unroll
        for (int u = 0; u < kLeanBatch; ++u) {
          ok[u] = hits != 0;
          if (ok[u]) {
            const int p = slot_of(k0 + __builtin_ctz(hits));
            hits &= hits - 1;
            f[u] = X[p];
            ga[u] = (double)galpha[p];
          }
        }
#pragma unroll
        for (int u = 0; u < kLeanBatch; ++u)
          if (ok[u]) add(frac(f[u]), ga[u]);
      }
    }
  }
#pragma unroll
  for (int m = kThetaLanes / 2; m > 0; m >>= 1) {
    s00 += __shfl_xor(s00, m, kThetaLanes);
    s10 += __shfl_xor(s10, m, kThetaLanes);
    s01 += __shfl_xor(s01, m, kThetaLanes);
    s11 += __shfl_xor(s11, m, kThetaLanes);
  }
  if (lane == 0) {
    double4 &pt = A.part[tcoff + j0 + i0 * ncy];
    if (A.first) {
      pt = make_double4(s00, s10, s01, s11);
    } else {
      const double4 o = pt;
      pt = make_double4(o.x + s00, o.y + s10, o.z + s01, o.w + s11);
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_theta_strip_keyed(ThetaKeyed A, Batch B) {
  theta_strip((int)blockIdx.x, A, B);
}

// adjoint P1: adj_pass1_kernel! + design term + gV += gVs (src/adjoint.jl:30-122,166-185).
// kExact: nm is already the exact list of these positions; otherwise pairs are
// tested and the exact list is written to em for P2.
// The window's three sweep rows are staged and walked one after the other
// (stage_copy_row), each slot walking its rows in the same order as
// for_each_nb: a third of the LDS of staging the whole window at once, so
// that registers, not LDS, bound the occupancy (7 waves per SIMD instead of
// 5: P1 49.8 -> 46.0 us on the 1M block, the same bits).  A pair reads
// (x, pi) and (abar, v) as two 16-byte records.
constexpr int kStageCapP1 = 448;
struct P1Lds {
  StageTab T;
  float4 sxp[kStageCapP1], sav[kStageCapP1];
};
#define P1_PARAMS                                                                           \
  const double2 *__restrict__ X, const float2 *__restrict__ xr,                             \
      const float2 *__restrict__ V, const float2 *__restrict__ PI,                          \
      const float2 *__restrict__ abar, const float *__restrict__ galpha,                    \
      const double *__restrict__ theta, const int *__restrict__ bcell,                      \
      const uint32_t *__restrict__ nm, const int *__restrict__ cstart,                      \
      const int *__restrict__ lane, Batch B, float2 *__restrict__ gXs,                      \
      float2 *__restrict__ gV, float *__restrict__ grho, uint32_t *__restrict__ em
#define P1_ARGS X, xr, V, PI, abar, galpha, theta, bcell, nm, cstart, lane, B, gXs, gV, grho, em

// workgroup blk of P1
template <bool kExact>
__device__ __forceinline__ void adj_p1(P1Lds &L, int blk, P1_PARAMS) {
  StageTab &T = L.T;
  float4 *const sxp = L.sxp, *const sav = L.sav;
  SCENE_WIN_AT(B, lane, blk);
  NB_RSRC;
  // the slot's own loads first: they arrive while the window is staged
  const double2 x64 = X[s];
  const float2 xi = xr[s], vi = V[s], ab = abar[s], pii = PI[s];
  const int bc = bcell[s];
  const float ga = galpha[s];
  stage_plan<true>(T, bcell, cst, w0_, w1_, g, S.coff, B.nostage ? 0 : kStageCapP1);
  const Rsrc pir = rsrc(PI, (uint32_t)B.N * 8u);
  const Rsrc vr = rsrc(V, (uint32_t)B.N * 8u);
  const Rsrc abr = rsrc(abar, (uint32_t)B.N * 8u);
  auto gld = [&](int j) __attribute__((always_inline)) {
    const uint32_t o8 = (uint32_t)j * 8u;   // one offset for the four arrays
    return NbXPAV{ld2(xrr, o8), ld2(pir, o8), ld2(abr, o8), ld2(vr, o8)};
  };
  auto sst = [&](int i, const NbXPAV &d) __attribute__((always_inline)) {
    sxp[i] = make_float4(d.x.x, d.x.y, d.pi.x, d.pi.y);
    sav[i] = make_float4(d.ab.x, d.ab.y, d.v.x, d.v.y);
  };
  auto sld = [&](int i) __attribute__((always_inline)) {
    const float4 q = sxp[i], r = sav[i];
    return NbXPAV{make_float2(q.x, q.y), make_float2(q.z, q.w), make_float2(r.x, r.y),
                  make_float2(r.z, r.w)};
  };
  // (threads past the scene's end take part in the row copies and barriers;
  // they load the head of slot w0_, a slot of the scene, and never walk it)
  const NbHead h = nb_head(nmr, cst, s, bc, g, S.coff, B.N);
  const float pti = pii.x, iri = pii.y;
  const float mpti = P.m * pti, tmi = P.twomumass * iri;
  ASum gx, gy, gvx, gvy, grh, sumS;
  auto pair = [&, P](const NbXPAV &d, float dx, float dy, float r2) __attribute__((always_inline)) {
    if (kExact || r2 < P.r2max) {
      const float rinv = __builtin_amdgcn_rsqf(r2);
      const float u = fmaf(r2 * rinv, P.mhh, 1.f);
      const float u2 = u * u;
      const float F = (P.Fc * u2) * u;
      const float G = P.Gch * u2;       // G/h
      const float2 pij = d.pi;
      const float Pij = fmaf(P.m, pij.x, mpti);
      const float Cbase = tmi * pij.y;
      const float C = Cbase * F;
      const float2 abj = d.ab;
      const float2 vj = d.v;
      float dax = ab.x - abj.x;
      float day = ab.y - abj.y;
      float dvx = vi.x - vj.x;
      float dvy = vi.y - vj.y;
      float ad_d = -fmaf(dax, dx, day * dy);
      float ad_v = fmaf(dax, dvx, day * dvy);
      float Fbar = fmaf(Pij, ad_d, Cbase * ad_v);
      float wgt = (Fbar * G) * rinv;
      float PF = Pij * F;
      gx.add(fmaf(wgt, dx, -(PF * dax)));
      gy.add(fmaf(wgt, dy, -(PF * day)));
      gvx.add(C * dax);
      gvy.add(C * day);
      grh.add(C * ad_v);
      sumS.add(F * ad_d);
      return true;
    }
    return false;
  };
  uint32_t *emw = kExact ? nullptr : em;
  if (T.ok) {
    const RowOff ro = stage_rowoff(T, bc, g);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      if (r) __syncthreads();   // every walk of row r-1 is done with the buffer
      // two elements in flight per thread: the copy's registers are live
      // beside the sums, and a row is ~1.1-1.6 elements per thread
      stage_copy_row<2>(T, r, gld, sst);
      if (valid_) nb_row<false>(r, nmr, emw, h, s, xi, B.N, P, ro, sld, pair);
    }
    if (!valid_) return;
  } else {
    if (!valid_) return;
    for_each_nb(nmr, emw, h, s, xi, B.N, P, RowOff{}, gld, pair);
  }
  float Ssum = sumS.value();
  // the pair sum of -C ad_v ir_i, with ir_i taken out of the sum
  float grv = -(grh.value() * iri) + (-2.f * P.m * pti * iri) * Ssum +
              (P.c2 * P.m * iri * iri) * Ssum;
  float gxv = gx.value() + ab.x * wall_daccel((float)x64.x, P.Lx, P.kw);
  float gyv = gy.value() + ab.y * wall_daccel((float)x64.y, P.Ly, P.kw);
  float dadx, dady;
  interp_alpha(theta + S.toff, interp_cell(x64.x, x64.y, P), P, &dadx, &dady);
  gxv += ga * dadx;
  gyv += ga * dady;
  gXs[s] = make_float2(gxv, gyv);
  grho[s] = grv;
  float2 gvo = gV[s];
  gV[s] = make_float2(gvo.x + gvx.value(), gvo.y + gvy.value());
}

// The exact-list instance fits 72 VGPRs, 7 waves per SIMD, without spills
// when asked to; the other (P1 writing the exact list) stays at its 5.
#define SPH_P1_WAVES __attribute__((amdgpu_waves_per_eu(kExact ? 7 : 5, 8)))
template <bool kExact>
__global__ void __launch_bounds__(kBlock) SPH_P1_WAVES k_adj_p132(P1_PARAMS) {
  __shared__ P1Lds L;
  adj_p1<kExact>(L, xcd_block(), P1_ARGS);
}

// P1 and the gtheta gather of the same step in one launch: both read only
// what the pointwise pass and the tape hold, and neither reads what the other
// writes.  Gather workgroups are spread evenly among the P1 workgroups (block
// b is a gather block when floor((b+1) nt / n) > floor(b nt / n)), so that
// they fill the SIMD time P1 leaves while its waves wait on LDS and memory;
// the gather holds no LDS, so the launch has P1's occupancy.  Each workgroup
// runs the body it would run in its own kernel, so every result is
// unchanged.
template <bool kExact>
__global__ void __launch_bounds__(kBlock) SPH_P1_WAVES k_adj_p1_theta32(P1_PARAMS, ThetaKeyed A) {
  __shared__ P1Lds L;
  const int lb = xcd_block();
  const long long nt = B.ntblk, n = gridDim.x;
  const int tb = (int)((long long)lb * nt / n);
  if ((int)((long long)(lb + 1) * nt / n) > tb)
    theta_strip(tb, A, B);
  else
    adj_p1<kExact>(L, lb - tb, P1_ARGS);
}

// adjoint P2: adj_pass2_kernel! and gX += gXs (src/adjoint.jl:126-156,187-191)
template <bool kFuseR>
__global__ void __launch_bounds__(kBlock)
k_adj_p232(const float2 *__restrict__ xr, const float *__restrict__ grho,
           const float2 *__restrict__ gXs, const int *__restrict__ bcell,
           const uint32_t *__restrict__ nm, const int *__restrict__ cstart,
           const int *__restrict__ lane, Batch B, float2 *__restrict__ gX, RpwArgs R) {
  __shared__ StageTab T;
  __shared__ float2 sx[kStageCap];
  __shared__ float sgr[kStageCap];
  SCENE_WIN(B, lane);
  NB_RSRC;
  // the slot's own loads before the staging plan, its walk's row heads after
  // it: both arrive while the window is copied
  RpwIn rin{};
  if (kFuseR) rin = rpw_load(R, s);
  const float2 xi = xr[s];
  const float gri = grho[s];
  const int bc = bcell[s];
  const float2 gs = gXs[s];
  const float2 go = gX[s];
  stage_plan(T, bcell, cst, w0_, w1_, g, S.coff, B.nostage ? 0 : kStageCap);
  const NbHead h = nb_head(nmr, cst, s, bc, g, S.coff, B.N);
  const Rsrc grr = rsrc(grho, (uint32_t)B.N * 4u);
  if (T.ok)
    stage_copy(T,
               [&](int j) {
                 return NbXG{ld2(xrr, (uint32_t)j * 8u), ld1(grr, (uint32_t)j * 4u)};
               },
               [&](int i, const NbXG &d) {
                 sx[i] = d.x;
                 sgr[i] = d.gr;
               });
  if (!valid_) return;
  ASum gx, gy;
  auto pair = [&, P](const NbXG &d, float dx, float dy, float r2) __attribute__((always_inline)) {
    const float u = fmaf(__builtin_amdgcn_sqrtf(r2), P.mhh, 1.f);
    const float wgt = (P.mFc * (u * u * u)) * (gri + d.gr);
    gx.add(wgt * dx);
    gy.add(wgt * dy);
    return true;
  };
  // nm is the exact list of P1 over these positions
  if (T.ok) {
    for_each_nb<false>(nmr, nullptr, h, s, xi, B.N, P, stage_rowoff(T, bc, g),
                [&](int i) __attribute__((always_inline)) { return NbXG{sx[i], sgr[i]}; }, pair);
  } else {
    for_each_nb(nmr, nullptr, h, s, xi, B.N, P, RowOff{},
                [&](int i) __attribute__((always_inline)) {
                  return NbXG{ld2(xrr, (uint32_t)i * 8u), ld1(grr, (uint32_t)i * 4u)};
                },
                pair);
  }
  const float2 gxn = make_float2(go.x + (gs.x + gx.value()), go.y + (gs.y + gy.value()));
  gX[s] = gxn;
  // the R pass of the previous step (same slot order) on this step's final gX
  if (kFuseR) rpw_finish(R, rin, s, bc, gxn, S, B);
}

}  // namespace

void launch_assign_f32c(hipStream_t s, const double2 *X, const Batch &B, int *cellof,
                        int *counts, int *flags) {
  if (B.N > 0) k_assign32<<<B.nblk, B.bs, 0, s>>>(X, B, cellof, counts, flags);
}
void launch_nm_build_f32c(hipStream_t s, const float2 *xr, const int *bcell, const int *cstart,
                          const Batch &B, uint32_t *nm, int *flags, const int *qs) {
  if (B.N <= 0) return;
  if (qs) k_nm_build32<true><<<B.nblk, B.bs, 0, s>>>(xr, bcell, cstart, qs, B, nm, flags);
  else k_nm_build32<false><<<B.nblk, B.bs, 0, s>>>(xr, bcell, cstart, nullptr, B, nm, flags);
}
void launch_density_build_f32c(hipStream_t s, const float2 *xr, const int *bcell,
                               const int *cstart, const int *lane, const Batch &B, float2 *PI,
                               uint32_t *nm, uint32_t *em, int *flags, const int *qs) {
  if (B.N <= 0) return;
  if (qs)
    k_density_build32<true><<<B.nblk, B.bs, 0, s>>>(xr, bcell, cstart, qs, lane, B, PI, nm, em,
                                                    flags);
  else
    k_density_build32<false><<<B.nblk, B.bs, 0, s>>>(xr, bcell, cstart, nullptr, lane, B, PI, nm,
                                                     em, flags);
}
void launch_density_f32c(hipStream_t s, const float2 *xr, const int *bcell, const uint32_t *nm,
                         const int *cstart, const int *lane, const Batch &B, float2 *PI,
                         double *rho, uint32_t *em) {
  if (B.N > 0) k_density32<<<B.nblk, B.bs, 0, s>>>(xr, bcell, nm, cstart, lane, B, PI, rho, em);
}
void launch_force_f32c(hipStream_t s, const double2 *X, const float2 *xr, const float2 *V,
                       const float2 *PI, const double *theta, const int *bcell, const uint32_t *nm,
                       const int *cstart, const int *lane, const Batch &B, double2 *Xn,
                       float2 *xrn, float2 *Vn, double2 *tapeX, float2 *tapeV, double2 *a_out,
                       double *alpha_out) {
  if (B.N > 0)
    k_force32<<<B.nblk, B.bs, 0, s>>>(X, xr, V, PI, theta, bcell, nm, cstart, lane, B, Xn, xrn,
                                        Vn, tapeX, tapeV, a_out, alpha_out);
}
void launch_xrel_f32c(hipStream_t s, const double2 *X, const int *bcell, const Batch &B,
                      float2 *xr, int *lane) {
  if (lane && B.bs != kLaneWin) throw std::logic_error("the thread order needs window-sized workgroups");
  if (B.N > 0) k_xrel<<<B.nblk, B.bs, 0, s>>>(X, bcell, B, xr, lane);
}
void launch_permute_f32c(hipStream_t s, const int *perm, const int *cellof, const double2 *X,
                         const float2 *V, const double2 *Xref, int check, const Batch &B,
                         double2 *Xn, float2 *Vn, float2 *xr, double2 *Xrefn, int *bcell,
                         int *flags, int *lane) {
  if (lane && B.bs != kLaneWin) throw std::logic_error("the thread order needs window-sized workgroups");
  if (B.N > 0)
    k_permute32<<<B.nblk, B.bs, 0, s>>>(perm, cellof, X, V, Xref, check, B, Xn, Vn, xr, Xrefn,
                                          bcell, flags, lane);
}
void launch_snapshot_f32c(hipStream_t s, bool to_snap, const Batch &B, double2 *X, float2 *V,
                          int *ids, int *scell, int *lane, int *cstart, double2 *Xref,
                          float2 *xr, double2 *qX, float2 *qV, int *qids, int *qscell,
                          int *qlane, int *qcstart) {
  if (B.nscenes != 1) throw std::logic_error("snapshots hold single-scene layouts");
  const long long n = B.N > B.ncell + 1 ? B.N : B.ncell + 1;
  k_snapshot32<<<nblocks(n), kBlock, 0, s>>>(to_snap ? 1 : 0, B, X, V, ids, scell, lane, cstart,
                                             Xref, xr, qX, qV, qids, qscell, qlane, qcstart);
}
void launch_adj_r_f32c(hipStream_t s, const double2 *X, const float2 *xr, const float2 *Vnext,
                       const double *theta, const int *bcell, const uint32_t *nm,
                       const int *cstart, const int *lane, const Batch &B, float2 *PI,
                       float2 *gX, float2 *gV, float2 *abar, float *galpha, uint32_t *em,
                       int *tkey, unsigned long long *tlev, int tstamp) {
  if (B.N > 0)
    k_adj_r32<<<B.nblk, B.bs, 0, s>>>(X, xr, Vnext, theta, bcell, nm, cstart, lane, B, PI, gX,
                                        gV, abar, galpha, em, tkey, tlev, tstamp);
}
void launch_adj_rpw_f32c(hipStream_t s, const double2 *X, const int *bcell, const float2 *Vnext,
                         const double *theta, const Batch &B, float2 *xr, float2 *gX,
                         float2 *gV, float2 *abar, float *galpha, int *tkey,
                         unsigned long long *tlev,
                         int tstamp) {
  const RpwArgs R{X, Vnext, theta, xr, gV, abar, galpha, tkey, tlev, tstamp};
  if (B.N > 0) k_adj_rpw32<<<B.nblk, B.bs, 0, s>>>(R, bcell, gX, B);
}
void launch_remap_rpw_f32c(hipStream_t s, const double2 *X, const int *bcell,
                           const float2 *gXo, const float2 *gVo, const float2 *Vno,
                           const int *inv, const int *ids, const double *theta, const Batch &B,
                           float2 *gXn, float2 *gVn, float2 *xr, float2 *abar, float *galpha,
                           int *tkey, unsigned long long *tlev, int tstamp, bool remap_v) {
  if (gVn == gVo) throw std::logic_error("remapped R pass writes gV over its source");
  const RpwArgs R{X, nullptr, theta, xr, gVn, abar, galpha, tkey, tlev, tstamp};
  if (B.N <= 0) return;
  if (remap_v)
    k_remap_rpw32<true><<<B.nblk, B.bs, 0, s>>>(R, bcell, gXo, gVo, Vno, inv, ids, B, gXn);
  else
    k_remap_rpw32<false><<<B.nblk, B.bs, 0, s>>>(R, bcell, gXo, gVo, Vno, inv, ids, B, gXn);
}
void launch_adj_p1_f32c(hipStream_t s, const double2 *X, const float2 *xr, const float2 *V,
                        const float2 *PI, const float2 *abar, const float *galpha,
                        const double *theta, const int *bcell, const uint32_t *nm, bool exact,
                        const int *cstart, const int *lane, const Batch &B, float2 *gXs,
                        float2 *gV, float *grho, uint32_t *em) {
  if (B.N <= 0) return;
  if (exact)
    k_adj_p132<true><<<B.nblk, B.bs, 0, s>>>(X, xr, V, PI, abar, galpha, theta, bcell, nm,
                                               cstart, lane, B, gXs, gV, grho, nullptr);
  else
    k_adj_p132<false><<<B.nblk, B.bs, 0, s>>>(X, xr, V, PI, abar, galpha, theta, bcell, nm,
                                                cstart, lane, B, gXs, gV, grho, em);
}
void launch_theta_gather_keyed_f32(hipStream_t s, const ThetaKeyed &A, const Batch &B) {
  if (B.ntblk > 0) k_theta_strip_keyed<<<B.ntblk, kBlock, 0, s>>>(A, B);
}
void launch_adj_p1_theta_f32c(hipStream_t s, const double2 *X, const float2 *xr,
                              const float2 *V, const float2 *PI, const float2 *abar,
                              const float *galpha, const double *theta, const int *bcell,
                              const uint32_t *nm, bool exact, const int *cstart, const int *lane,
                              const Batch &B, float2 *gXs, float2 *gV, float *grho, uint32_t *em,
                              const ThetaKeyed &A) {
  if (B.bs != kBlock) throw std::logic_error("P1 + gtheta launch needs 256-thread workgroups");
  if ((long long)B.nblk + B.ntblk > 0x7fffffffLL) throw std::logic_error("P1 + gtheta grid");
  if (B.nblk + B.ntblk == 0) return;
  const int n = B.nblk + B.ntblk;
  if (exact)
    k_adj_p1_theta32<true><<<n, kBlock, 0, s>>>(X, xr, V, PI, abar, galpha, theta, bcell, nm,
                                                  cstart, lane, B, gXs, gV, grho, nullptr, A);
  else
    k_adj_p1_theta32<false><<<n, kBlock, 0, s>>>(X, xr, V, PI, abar, galpha, theta, bcell, nm,
                                                   cstart, lane, B, gXs, gV, grho, em, A);
}
void launch_adj_p2_f32c(hipStream_t s, const float2 *xr, const float *grho, const float2 *gXs,
                        const int *bcell, const uint32_t *nm, const int *cstart, const int *lane,
                        const Batch &B, float2 *gX, const RpwArgs *prev) {
  if (B.N <= 0) return;
  if (prev) {
    if (prev->xr == xr) throw std::logic_error("P2 reads xr while the fused R pass writes it");
    k_adj_p232<true><<<B.nblk, B.bs, 0, s>>>(xr, grho, gXs, bcell, nm, cstart, lane, B, gX, *prev);
  } else {
    k_adj_p232<false><<<B.nblk, B.bs, 0, s>>>(xr, grho, gXs, bcell, nm, cstart, lane, B, gX,
                                                RpwArgs{});
  }
}

}  // namespace sph
