// sph_stage.h -- LDS staging of the neighbour rows of a window of slots,
// shared by the FP64 and f32c pair passes.
//
// A workgroup's slots form one window of consecutive slots, hence a
// consecutive range of binning cells: cells (xa, ya) .. (xb, yb) in row-major
// order.  Every 3x3 sweep of the window reads, for each own cell row y'
// (ya..yb) and each r = 0..2, cells max(x0-1,0)..min(x1+1,nx-1) of row
// y'+r-1, x0..x1 being the window's own cells in row y'; that cell range is
// one slot range (segment (y'-ya)*3 + r).  The segments are copied into LDS
// with coalesced loads and the pair loop then reads its neighbours there.  A
// window spanning more than kStageOwnRows cell rows, or whose segments exceed
// the pass's capacity, reads from global memory instead (same results: only
// the loader differs).
#pragma once
#include "sph_internal.h"

namespace sph {

// (cx, cy) of a 0-based cell; the quotient estimate in FP64 is exact to one
// unit for cells < 2^30 and is corrected once
__device__ __forceinline__ int2 cell_xy(int c, const Grid &g) {
  int cy = (int)((double)c * g.inv_nx);
  int cx = c - cy * g.nx;
  if (cx < 0) { cy -= 1; cx += g.nx; }
  if (cx >= g.nx) { cy += 1; cx -= g.nx; }
  return make_int2(cx, cy);
}

constexpr int kStageOwnRows = 4;
constexpr int kStageSegs = 3 * kStageOwnRows;

struct StageTab {
  int lo[kStageSegs];         // first slot of each segment
  int off[kStageSegs + 1];    // LDS position of each segment; off[kStageSegs] = total
  int delta[kStageSegs];      // lo - off: slot = LDS position + delta of its segment
  int ya, ok, nseg;           // nseg: segments in use (3 per own cell row)
  int ptot[3];                // phased plans: slots of the segments of row r (k % 3 == r)
};

// Plans the staging of the window of slots [w0, w1) (w1 > w0); every thread
// of the workgroup calls it.  Wave 0 forms the segments and their prefix.
// bcell[slot] is the slot's binning cell (ascending over the slots), and
// cst_at(i) returns cstart[i] (the first slot of cell i - coff).  A segment
// of a row outside the grid is empty and has lo = INT_MAX.
// kPhased: the segments of sweep row r (k % 3 == r) share one LDS buffer of
// cap slots, copied one row after the other (stage_copy_row): the positions
// of a row's segments are its own prefix, and T.ok needs each row to fit.
template <bool kPhased = false, class CstAt>
__device__ __forceinline__ void stage_plan_at(StageTab &T, const int *__restrict__ bcell,
                                              CstAt &&cst_at, int w0, int w1, const Grid &g,
                                              int coff, int cap) {
  if (threadIdx.x < 64) {
    const int k = (int)threadIdx.x;
    const int2 a = cell_xy(bcell[w0], g), b = cell_xy(bcell[w1 - 1], g);
    const int nrow = b.y - a.y + 1;
    int len = 0, lo = 0x7fffffff;
    if (k < kStageSegs && k < 3 * nrow) {
      const int yo = a.y + k / 3, y = yo + k % 3 - 1;
      if (y >= 0 && y < g.ny) {
        const int x0 = yo == a.y ? a.x : 0, x1 = yo == b.y ? b.x : g.nx - 1;
        const int xlo = x0 > 0 ? x0 - 1 : 0, xhi = x1 < g.nx - 1 ? x1 + 1 : g.nx - 1;
        const int base = coff + y * g.nx;
        lo = cst_at(base + xlo);
        len = cst_at(base + xhi + 1) - lo;
      }
    }
    int inc = len, total = 0;
    bool fits;
    if constexpr (kPhased) {
#pragma unroll
      for (int d = 3; d < kStageSegs; d += 3) {
        const int t = __shfl_up(len, d, 64);
        if (k >= d) inc += t;
      }
      const int t0 = __shfl(inc, kStageSegs - 3, 64), t1 = __shfl(inc, kStageSegs - 2, 64),
                t2 = __shfl(inc, kStageSegs - 1, 64);
      if (k == 0) {
        T.ptot[0] = t0;
        T.ptot[1] = t1;
        T.ptot[2] = t2;
      }
      fits = t0 <= cap && t1 <= cap && t2 <= cap;
    } else {
#pragma unroll
      for (int d = 1; d < 16; d <<= 1) {
        const int t = __shfl_up(inc, d, 64);
        if (k >= d) inc += t;
      }
      total = __shfl(inc, kStageSegs - 1, 64);
      fits = total <= cap;
    }
    if (k < kStageSegs) {
      T.lo[k] = lo;
      T.off[k] = inc - len;
      T.delta[k] = lo - (inc - len);
    }
    if (k == 0) {
      T.off[kStageSegs] = total;   // unused by phased plans (T.ptot)
      T.ya = a.y;
      T.ok = (nrow <= kStageOwnRows && fits) ? 1 : 0;
      T.nseg = 3 * (nrow < kStageOwnRows ? nrow : kStageOwnRows);
    }
  }
  __syncthreads();
}

// Copies the planned segments: st(i, ld(j)) stores slot j's data at LDS
// index i.  Called by every thread of the workgroup when T.ok.  Each thread
// issues the loads of kStageUnroll elements before storing any of them, so a
// window's copy costs one load latency per kStageUnroll*blockDim elements.
// The segment of an element is the last of the T.nseg segments in use that
// starts at or before it: two compares for a window inside one cell row
// (three segments), one more per further segment, whose offset is read from
// LDS (a loop over an array of offsets in registers, indexed at run time,
// compiles to s_set_gpr_idx per segment and element); its slot offset is
// read from T.delta.
constexpr int kStageUnroll = 4;
template <class Ld, class St>
__device__ __forceinline__ void stage_copy(const StageTab &T, Ld &&ld, St &&st) {
  const int total = T.off[kStageSegs];
  const int nseg = T.nseg;
  // a window inside one cell row (the common case) has three segments: two
  // compares against offsets held in scalar registers; a window over more
  // rows compares against the offsets in LDS
  const int o1 = __builtin_amdgcn_readfirstlane(T.off[1]);
  const int o2 = __builtin_amdgcn_readfirstlane(T.off[2]);
  for (int i0 = (int)threadIdx.x; i0 < total; i0 += kStageUnroll * (int)blockDim.x) {
    decltype(ld(0)) d[kStageUnroll];
#pragma unroll
    for (int u = 0; u < kStageUnroll; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < total) {
        // segment of i: the last one in use starting at or before i (off ascends)
        int k = (i >= o1 ? 1 : 0) + (i >= o2 ? 1 : 0);
        for (int q = 3; q < nseg; ++q) k += (i >= T.off[q]) ? 1 : 0;
        d[u] = ld(i + T.delta[k]);
      }
    }
#pragma unroll
    for (int u = 0; u < kStageUnroll; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < total) st(i, d[u]);
    }
  }
  __syncthreads();
}

// Copies the segments of sweep row r of a phased plan (stage_plan_at<true>)
// to LDS positions 0 .. T.ptot[r]-1, as stage_copy does for a whole plan.
template <int kUnroll = kStageUnroll, class Ld, class St>
__device__ __forceinline__ void stage_copy_row(const StageTab &T, int r, Ld &&ld, St &&st) {
  const int total = T.ptot[r];
  const int nseg = T.nseg;
  int off[kStageOwnRows];
#pragma unroll
  for (int q = 0; q < kStageOwnRows; ++q) off[q] = T.off[3 * q + r];
  for (int i0 = (int)threadIdx.x; i0 < total; i0 += kUnroll * (int)blockDim.x) {
    decltype(ld(0)) d[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < total) {
        // segment of i: the last of row r in use starting at or before i
        int q = 0;
#pragma unroll
        for (int p = 1; p < kStageOwnRows; ++p) {
          if (3 * p + r >= nseg) break;
          q += (i >= off[p]) ? 1 : 0;
        }
        d[u] = ld(i + T.delta[3 * q + r]);
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < total) st(i, d[u]);
    }
  }
  __syncthreads();
}

}  // namespace sph
