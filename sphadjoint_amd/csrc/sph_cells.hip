// sph_cells.hip -- cell list, slot reordering and layout remaps for gfx950.
//
// Rebuild pipeline (the reference's build! + _nl_build!, src/neighbors.jl:97-114,192-245):
//   assign   cell of every slot by floor(x/cs) (a division, as cell_index does),
//            counts by wave-aggregated atomics (one atomic per run of equal cells)
//   scan     exclusive prefix of counts -> cstart (LDS block scan, shuffle wave scan)
//   fill     slot -> position in its cell (wave-aggregated atomic cursor)
//   rank     rank inside the cell by particle id: the result equals the reference's
//            _cl_sort_cells! insertion sort, without a serial per-cell loop
//   permute  gather X, V into the new slot order, check displacement against Xref,
//            record the sweep cell by floor(x*(1/cs)) (a multiplication, as _nl_build! does)
// All of it is integer/byte work bound by HBM and atomics; no MFMA.
#include "sph_internal.h"

namespace sph {

namespace {

inline int nblocks(long long n, int b = kBlock) { return (int)((n + b - 1) / b); }

__global__ void k_fill_i32(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_assign(const double2 *__restrict__ X, Grid g, double cs, int *__restrict__ cellof,
                         int *__restrict__ counts, int *__restrict__ flags) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  int c = -1;
  if (s < g.N) {
    double2 x = X[s];
    // a NaN or infinite position bins into a clamped cell, as the reference's
    // unsafe_trunc does; it is reported (flags[2])
    if (!isfinite(x.x) || !isfinite(x.y)) flags[2] = 1;
    int cx = cell_coord_div(x.x / cs, g.nx);
    int cy = cell_coord_div(x.y / cs, g.ny);
    c = cx + cy * g.nx;
    cellof[s] = c;
  }
  int head_lane, run_len;
  wave_runs(c, head_lane, run_len);
  const int lane = threadIdx.x & 63;
  if (c >= 0 && head_lane == lane) atomicAdd(&counts[c], run_len);
}

// ---- exclusive scan: 256 threads x 4 items per block ---------------------------

constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and sets `total`.
__device__ __forceinline__ int block_excl_scan(int v, int &total) {
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int inc = wave_incl_scan(v);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    int ws = wsum[w];
    off += (w < wid) ? ws : 0;
    tot += ws;
  }
  __syncthreads();
  total = tot;
  return off + inc - v;
}

// the list of cells with more than kRankSmall entries (big[0] = their count,
// big[1..] = the cells, in no particular order): k_rank's extra workgroups
// rank them by a bitmap of their ids instead of the per-entry count
constexpr int kRankSmall = 128;

__global__ void k_scan_partial(const int *__restrict__ in, int n, int *__restrict__ bsum,
                               int *__restrict__ big) {
  if (blockIdx.x == 0 && threadIdx.x == 0) big[0] = 0;   // read by k_scan_final and k_rank
  const int base = blockIdx.x * kScanTile;
  int v = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int i = base + threadIdx.x * kScanItems + k;
    v += (i < n) ? in[i] : 0;
  }
  int total;
  block_excl_scan(v, total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Exclusive scan of one tile; the tile's offset is the sum of the partial sums
// of the tiles before it (each block adds them itself: integers, so any order
// gives the same sum).  The counts are cleared once read, ready for the next
// build's atomics.
__global__ void k_scan_final(int *__restrict__ in, int n, const int *__restrict__ bsum,
                             int *__restrict__ out, int *__restrict__ cursor,
                             int *__restrict__ big) {
  const int base = blockIdx.x * kScanTile;
  int pre = 0;
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += kBlock) pre += bsum[b];
  int tile_off;
  block_excl_scan(pre, tile_off);
  int vals[kScanItems];
  int v = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int i = base + threadIdx.x * kScanItems + k;
    vals[k] = (i < n) ? in[i] : 0;
    if (i < n) in[i] = 0;
    v += vals[k];
  }
  int total;
  int ex = block_excl_scan(v, total) + tile_off;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int i = base + threadIdx.x * kScanItems + k;
    if (i < n) {
      out[i] = ex;
      if (cursor) cursor[i] = ex;
      if (big && vals[k] > kRankSmall) big[1 + atomicAdd(&big[0], 1)] = i;
    }
    ex += vals[k];
    if (i == n - 1) out[n] = ex;
  }
}

// Entries of each cell at their cell's cursor: the rank key (the particle id,
// or with qX the quadrant-major key quadrant * N + id of the half-cell sweep)
// and the slot.
__global__ void k_fill(const int *__restrict__ cellof, const int *__restrict__ ids, int N,
                       const double2 *__restrict__ qX, Grid g, double cs,
                       int *__restrict__ cursor, int *__restrict__ tmp_id,
                       int *__restrict__ tmp_slot) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  int c = (s < N) ? cellof[s] : -1;
  int head_lane, run_len;
  wave_runs(c, head_lane, run_len);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (c >= 0 && head_lane == lane) base = atomicAdd(&cursor[c], run_len);
  base = __shfl(base, head_lane, 64);
  if (c >= 0) {
    int pos = base + (lane - head_lane);
    int key = ids[s];
    if (qX) {
      const int cy = c / g.nx;
      key += cell_quadrant(qX[s], cs, c - cy * g.nx, cy) * N;
    }
    tmp_id[pos] = key;
    tmp_slot[pos] = s;
  }
}

// First slot of each quadrant of each cell, for a layout whose cells hold
// their slots in quadrant order: qs[4c + k] (k = 0..3; qs[4c] = cstart[c]),
// qs[4 ncell] = N.  Thread c binary-searches cell c's slots by the quadrant
// of their position X (the positions the layout was sorted by).
__global__ void k_quad_starts(const double2 *__restrict__ X, const int *__restrict__ cstart,
                              Grid g, double cs, int *__restrict__ qs) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > g.ncell) return;
  if (c == g.ncell) {
    qs[4 * c] = g.N;
    return;
  }
  const int b = cstart[c], e = cstart[c + 1];
  const int cy = c / g.nx, cx = c - cy * g.nx;
  qs[4 * c] = b;
  for (int k = 1; k < 4; ++k) {
    int lo = b, hi = e;
    while (lo < hi) {
      const int mid = lo + ((hi - lo) >> 1);
      if (cell_quadrant(X[mid], cs, cx, cy) < k) lo = mid + 1;
      else hi = mid;
    }
    qs[4 * c + k] = lo;
  }
}

// Rank of one big cell's entries [b, e) by key (distinct: ids, or
// quadrant-major keys): a workgroup sets one bit per key in an LDS bitmap of a
// window of kRankWin keys, prefixes the popcounts of its words, and places
// each entry at b + (keys below it); windows cover the cell's key span
// [lo, hi] in turn.  O(entries + span / 32) per window instead of
// O(entries^2); the same ranks, hence the same order, as the per-entry count.
constexpr int kRankWords = 2048;
constexpr int kRankWin = 32 * kRankWords;
constexpr int kRankWords_t = kRankWords / kBlock;   // words per thread in the prefix
__device__ void rank_big_cell(int b, int e, const int *__restrict__ tmp_id,
                              const int *__restrict__ tmp_slot, const int *__restrict__ ids,
                              int *__restrict__ perm, int *__restrict__ new_ids, unsigned *bits,
                              int *pre) {
  __shared__ int red[2 * (kBlock / 64)];
  const int tid = (int)threadIdx.x, lane = tid & 63, w = tid >> 6;
  int lo = INT_MAX, hi = INT_MIN;
  for (int q = b + tid; q < e; q += kBlock) {
    const int id = tmp_id[q];
    lo = min(lo, id);
    hi = max(hi, id);
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    lo = min(lo, __shfl_xor(lo, d, 64));
    hi = max(hi, __shfl_xor(hi, d, 64));
  }
  if (lane == 0) {
    red[w] = lo;
    red[kBlock / 64 + w] = hi;
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < kBlock / 64; ++v) {
    lo = min(lo, red[v]);
    hi = max(hi, red[kBlock / 64 + v]);
  }
  __syncthreads();
  int base = 0;   // entries with ids below the current window
  for (long long w0 = lo; w0 <= hi; w0 += kRankWin) {
    for (int k = tid; k < kRankWords; k += kBlock) bits[k] = 0u;
    __syncthreads();
    for (int q = b + tid; q < e; q += kBlock) {
      const long long d = (long long)tmp_id[q] - w0;
      if (d >= 0 && d < kRankWin) atomicOr(&bits[d >> 5], 1u << (d & 31));
    }
    __syncthreads();
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < kRankWords_t; ++k) cnt += __popc(bits[tid * kRankWords_t + k]);
    int total;
    int ex = block_excl_scan(cnt, total);
#pragma unroll
    for (int k = 0; k < kRankWords_t; ++k) {
      pre[tid * kRankWords_t + k] = ex;
      ex += __popc(bits[tid * kRankWords_t + k]);
    }
    __syncthreads();
    for (int q = b + tid; q < e; q += kBlock) {
      const int id = tmp_id[q];
      const long long d = (long long)id - w0;
      if (d >= 0 && d < kRankWin) {
        const int wd = (int)(d >> 5);
        const int t = b + base + pre[wd] + __popc(bits[wd] & ((1u << (d & 31)) - 1u));
        const int so = tmp_slot[q];
        perm[t] = so;
        new_ids[t] = ids ? ids[so] : id;
      }
    }
    base += total;
    __syncthreads();
  }
}

// Rank inside the cell by key (the particle id, or a quadrant-major key):
// entry p of a cell with at most kRankSmall entries counts the smaller keys of
// its cell; the last kRankBig workgroups take the big cells listed by the scan
// (big), one at a time.  new_ids receives the ids: the keys themselves, or
// ids[slot] when the keys are not ids.
constexpr int kRankBig = 64;
__global__ void __launch_bounds__(kBlock)
k_rank(const int *__restrict__ cellof, const int *__restrict__ cstart,
       const int *__restrict__ tmp_id, const int *__restrict__ tmp_slot, int N,
       const int *__restrict__ big, const int *__restrict__ ids, int *__restrict__ perm,
       int *__restrict__ new_ids) {
  __shared__ unsigned bits[kRankWords];
  __shared__ int pre[kRankWords];
  const int nb = (N + kBlock - 1) / kBlock;
  if ((int)blockIdx.x >= nb) {
    const int nbig = big[0];
    for (int k = (int)blockIdx.x - nb; k < nbig; k += kRankBig) {
      const int c = big[1 + k];
      rank_big_cell(cstart[c], cstart[c + 1], tmp_id, tmp_slot, ids, perm, new_ids, bits, pre);
    }
    return;
  }
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  int so = tmp_slot[p];
  int id = tmp_id[p];
  int c = cellof[so];
  int b = cstart[c], e = cstart[c + 1];
  if (e - b > kRankSmall) return;   // a big cell
  int rank = 0;
  for (int q = b; q < e; ++q) rank += (tmp_id[q] < id) ? 1 : 0;
  int t = b + rank;
  perm[t] = so;
  new_ids[t] = ids ? ids[so] : id;
}

__global__ void k_permute(const int *__restrict__ perm, const int *__restrict__ cellof,
                          const double2 *__restrict__ X, const double2 *__restrict__ V,
                          const double2 *__restrict__ Xref, int check, Grid g, CellConstF64 k,
                          double2 *__restrict__ Xn, double2 *__restrict__ Vn,
                          double2 *__restrict__ Xrefn, int *__restrict__ scell,
                          int *__restrict__ bcell, int *__restrict__ flags) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int o = perm[t];
  bcell[t] = cellof[o];
  double2 x = X[o];
  if (check) {
    double2 xr = Xref[o];
    double ddx = x.x - xr.x;
    double ddy = x.y - xr.y;
    if (ddx * ddx + ddy * ddy > k.dmax2) flags[1] = 1;
  }
  Xn[t] = x;
  Vn[t] = V[o];
  Xrefn[t] = x;
  int cx = cell_coord_i32(x.x * k.invcs, g.nx);
  int cy = cell_coord_i32(x.y * k.invcs, g.ny);
  scell[t] = cx + cy * g.nx;
}

__global__ void k_materialise(const double2 *__restrict__ Xref, const int *__restrict__ ids,
                              const int *__restrict__ scell, const int *__restrict__ cstart,
                              Grid g, CellConstF64 k, int maxnb, long long sm, long long si,
                              int *__restrict__ indices, int *__restrict__ nbcount,
                              int *__restrict__ flags) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const int id = ids[t];
  const double2 xi = Xref[t];
  Runs r = sweep_runs(cstart, scell[t], g.nx, g.ny);
  int m = 0;
  for (int w = 0; w < 3; ++w) {
    for (int j = r.b[w]; j < r.e[w]; ++j) {
      double2 xj = Xref[j];
      double dx = xi.x - xj.x;
      double dy = xi.y - xj.y;
      double r2 = dx * dx + dy * dy;
      if (k.r2min < r2 && r2 < k.rc2) {
        if (m < maxnb) {
          indices[(long long)m * sm + (long long)id * si] = ids[j] + 1;
          m += 1;
        } else {
          flags[0] = 1;
        }
      }
    }
  }
  nbcount[id] = m;
}

__global__ void k_scatter_i32(const int *__restrict__ src, const int *__restrict__ ids, int N,
                              int *__restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) dst[ids[s]] = src[s];
}
// A tape's record of an epoch's layout, in one launch: thread i copies slot i
// of the per-slot arrays given (ids, sweep cells, binning cells, thread order;
// a null source is skipped) and entry i of the cell starts.
__global__ void k_epoch_record(EpochLayout src, int N, int ncs, EpochLayoutOut dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) {
    dst.ids[i] = src.ids[i];
    dst.scell[i] = src.scell[i];
    if (src.bcell) dst.bcell[i] = src.bcell[i];
    if (src.lane) dst.lane[i] = src.lane[i];
  }
  if (i < ncs) dst.cstart[i] = src.cstart[i];
}

__global__ void k_inverse(const int *__restrict__ ids, int N, int *__restrict__ inv) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) inv[ids[s]] = s;
}
__global__ void k_gather_remap2(const double2 *__restrict__ src, const int *__restrict__ inv_src,
                                const int *__restrict__ ids_dst, int N, double2 *__restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) dst[s] = src[inv_src[ids_dst[s]]];
}
__global__ void k_scatter2(const double2 *__restrict__ src, const int *__restrict__ ids, int N,
                           double2 *__restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) dst[ids[s]] = src[s];
}
__global__ void k_gather2(const double2 *__restrict__ src, const int *__restrict__ ids, int N,
                          double2 *__restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) dst[s] = src[ids[s]];
}
__global__ void k_scatter1(const double *__restrict__ src, const int *__restrict__ ids, int N,
                           double *__restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) dst[ids[s]] = src[s];
}
__global__ void k_copy2(const double2 *__restrict__ src, int N, double2 *__restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) dst[s] = src[s];
}

}  // namespace

void launch_fill_i32(hipStream_t s, int *p, int n, int v) {
  if (n > 0) k_fill_i32<<<nblocks(n), kBlock, 0, s>>>(p, n, v);
}
void launch_assign(hipStream_t s, const double2 *X, Grid g, double cs, int *cellof, int *counts,
                   int *flags) {
  if (g.N > 0) k_assign<<<nblocks(g.N), kBlock, 0, s>>>(X, g, cs, cellof, counts, flags);
}
void launch_scan_excl(hipStream_t s, int *in, int n, int *out, int *cursor, int *bsum,
                      int *big) {
  int nb = (n + kScanTile - 1) / kScanTile;
  k_scan_partial<<<nb, kBlock, 0, s>>>(in, n, bsum, big);
  k_scan_final<<<nb, kBlock, 0, s>>>(in, n, bsum, out, cursor, big);
}
void launch_fill(hipStream_t s, const int *cellof, const int *ids, int N, int *cursor,
                 int *tmp_id, int *tmp_slot, const double2 *qX, Grid g, double cs) {
  if (N > 0)
    k_fill<<<nblocks(N), kBlock, 0, s>>>(cellof, ids, N, qX, g, cs, cursor, tmp_id, tmp_slot);
}
void launch_rank(hipStream_t s, const int *cellof, const int *cstart, const int *tmp_id,
                 const int *tmp_slot, int N, const int *big, int *perm, int *new_ids,
                 const int *ids) {
  if (N > 0)
    k_rank<<<nblocks(N) + kRankBig, kBlock, 0, s>>>(cellof, cstart, tmp_id, tmp_slot, N, big,
                                                    ids, perm, new_ids);
}
void launch_quad_starts(hipStream_t s, const double2 *X, const int *cstart, Grid g, double cs,
                        int *qs) {
  k_quad_starts<<<nblocks(g.ncell + 1), kBlock, 0, s>>>(X, cstart, g, cs, qs);
}
void launch_permute(hipStream_t s, const int *perm, const int *cellof, const double2 *X,
                    const double2 *V, const double2 *Xref, int check, Grid g, CellConstF64 k,
                    double2 *Xn, double2 *Vn, double2 *Xrefn, int *scell, int *bcell,
                    int *flags) {
  if (g.N > 0)
    k_permute<<<nblocks(g.N), kBlock, 0, s>>>(perm, cellof, X, V, Xref, check, g, k, Xn, Vn,
                                              Xrefn, scell, bcell, flags);
}
void launch_materialise(hipStream_t s, const double2 *Xref, const int *ids, const int *scell,
                        const int *cstart, Grid g, CellConstF64 k, int maxnb, long long sm,
                        long long si, int *indices, int *nbcount, int *flags) {
  if (g.N > 0)
    k_materialise<<<nblocks(g.N), kBlock, 0, s>>>(Xref, ids, scell, cstart, g, k, maxnb, sm, si,
                                                  indices, nbcount, flags);
}
void launch_scatter_i32(hipStream_t s, const int *src, const int *ids, int N, int *dst) {
  if (N > 0) k_scatter_i32<<<nblocks(N), kBlock, 0, s>>>(src, ids, N, dst);
}
void launch_epoch_record(hipStream_t s, const EpochLayout &src, int N, int ncs,
                         const EpochLayoutOut &dst) {
  const int n = N > ncs ? N : ncs;
  if (n > 0) k_epoch_record<<<nblocks(n), kBlock, 0, s>>>(src, N, ncs, dst);
}
void launch_inverse(hipStream_t s, const int *ids, int N, int *inv) {
  if (N > 0) k_inverse<<<nblocks(N), kBlock, 0, s>>>(ids, N, inv);
}
void launch_gather_remap2(hipStream_t s, const double2 *src, const int *inv_src,
                          const int *ids_dst, int N, double2 *dst) {
  if (N > 0) k_gather_remap2<<<nblocks(N), kBlock, 0, s>>>(src, inv_src, ids_dst, N, dst);
}
void launch_scatter2(hipStream_t s, const double2 *src, const int *ids, int N, double2 *dst) {
  if (N > 0) k_scatter2<<<nblocks(N), kBlock, 0, s>>>(src, ids, N, dst);
}
void launch_gather2(hipStream_t s, const double2 *src, const int *ids, int N, double2 *dst) {
  if (N > 0) k_gather2<<<nblocks(N), kBlock, 0, s>>>(src, ids, N, dst);
}
void launch_scatter1(hipStream_t s, const double *src, const int *ids, int N, double *dst) {
  if (N > 0) k_scatter1<<<nblocks(N), kBlock, 0, s>>>(src, ids, N, dst);
}
void launch_copy2(hipStream_t s, const double2 *src, int N, double2 *dst) {
  if (N > 0) k_copy2<<<nblocks(N), kBlock, 0, s>>>(src, N, dst);
}

}  // namespace sph

// ---- float2 variants for the f32c form --------------------------------------
#include "sph_f32c.h"

namespace sph {
namespace {
__global__ void k_f2_to_d2(const float2 *__restrict__ src, int N, double2 *__restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) dst[s] = make_double2(src[s].x, src[s].y);
}
__global__ void k_d2_to_f2(const double2 *__restrict__ src, int N, float2 *__restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) dst[s] = make_float2((float)src[s].x, (float)src[s].y);
}
__global__ void k_gather_remap2f(const float2 *__restrict__ src, const int *__restrict__ inv_src,
                                 const int *__restrict__ ids_dst, int N, float2 *__restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) dst[s] = src[inv_src[ids_dst[s]]];
}
// three float2 arrays into a new slot order at once (one index lookup per slot)
__global__ void k_gather_remap3f(const float2 *__restrict__ a, const float2 *__restrict__ b,
                                 const float2 *__restrict__ c, const int *__restrict__ inv_src,
                                 const int *__restrict__ ids_dst, int N, float2 *__restrict__ da,
                                 float2 *__restrict__ db, float2 *__restrict__ dc) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) {
    const int j = inv_src[ids_dst[s]];
    da[s] = a[j];
    db[s] = b[j];
    dc[s] = c[j];
  }
}
// dst[s] = (float2) src[ids[s]]: id-order FP64 host data into the slot order
__global__ void k_gather2f(const double2 *__restrict__ src, const int *__restrict__ ids, int N,
                           float2 *__restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) {
    double2 v = src[ids[s]];
    dst[s] = make_float2((float)v.x, (float)v.y);
  }
}
__global__ void k_scatter2f(const float2 *__restrict__ src, const int *__restrict__ ids, int N,
                            double2 *__restrict__ dst) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) dst[ids[s]] = make_double2(src[s].x, src[s].y);
}
}  // namespace

void launch_f2_to_d2(hipStream_t s, const float2 *src, int N, double2 *dst) {
  if (N > 0) k_f2_to_d2<<<nblocks(N), kBlock, 0, s>>>(src, N, dst);
}
void launch_d2_to_f2(hipStream_t s, const double2 *src, int N, float2 *dst) {
  if (N > 0) k_d2_to_f2<<<nblocks(N), kBlock, 0, s>>>(src, N, dst);
}
void launch_gather_remap3f(hipStream_t s, const float2 *a, const float2 *b, const float2 *c,
                           const int *inv_src, const int *ids_dst, int N, float2 *da, float2 *db,
                           float2 *dc) {
  if (N > 0)
    k_gather_remap3f<<<nblocks(N), kBlock, 0, s>>>(a, b, c, inv_src, ids_dst, N, da, db, dc);
}
void launch_gather_remap2f(hipStream_t s, const float2 *src, const int *inv_src,
                           const int *ids_dst, int N, float2 *dst) {
  if (N > 0) k_gather_remap2f<<<nblocks(N), kBlock, 0, s>>>(src, inv_src, ids_dst, N, dst);
}
void launch_gather2f(hipStream_t s, const double2 *src, const int *ids, int N, float2 *dst) {
  if (N > 0) k_gather2f<<<nblocks(N), kBlock, 0, s>>>(src, ids, N, dst);
}
void launch_scatter2f(hipStream_t s, const float2 *src, const int *ids, int N, double2 *dst) {
  if (N > 0) k_scatter2f<<<nblocks(N), kBlock, 0, s>>>(src, ids, N, dst);
}
void launch_scatter_pi(hipStream_t s, const float2 *src, const int *ids, int N, double2 *dst) {
  launch_scatter2f(s, src, ids, N, dst);
}
}  // namespace sph
