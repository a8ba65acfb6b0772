// sph_api.cpp -- host side of libsphadj: the C ABI of include/sphadj.h.
//
// Maps the reference's drivers (src/driver.jl, src/neighbors.jl:252-290) onto
// one in-order HIP stream per context.  Host logic that the reference keeps on
// the host (rebuild age counter, tape capacity, flag reporting) stays here;
// everything per particle runs in the kernels of sph_cells.hip / sph_f64.hip.
#include "sphadj.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sph_f32c.h"
#include "sph_internal.h"

using namespace sph;

namespace {

thread_local std::string g_err;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(SPH_E_HIP, "%s: %s", what, hipGetErrorString(e));
}
#define HIPC(x) hip_check((x), #x)

template <class F>
int guarded(F &&f) {
  try {
    f();
    return SPH_OK;
  } catch (const Error &e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_err = "host allocation failed";
    return SPH_E_NOMEM;
  } catch (const std::exception &e) {
    g_err = e.what();
    return SPH_E_HIP;
  }
}

template <class T>
T *dalloc(size_t n) {
  void *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess) fail(SPH_E_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  return static_cast<T *>(p);
}
template <class T>
void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

constexpr int64_t kAgeNever = INT64_MAX;
// state flags: [0] neighbour-list row overflow, [1] displacement beyond
// skin*h, [2] a position that is not finite at a rebuild
constexpr int kNumFlags = 3;

// Derived FP64 constants with the reference's operation order in T = Float64.
PhysF64 make_phys(const sph_params &p) {
  PhysF64 P{};
  const double h = p.h;
  P.m = p.m;
  P.rho0 = p.rho0;
  P.c2 = p.c * p.c;                                   // p.c^2
  P.mu = p.mu;
  P.gx = p.gx;
  P.gy = p.gy;
  P.dt = p.dt;
  P.Lx = p.Lx;
  P.Ly = p.Ly;
  P.kw = p.kw;
  P.A = 7.0 / ((4.0 * M_PI) * (h * h));               // wnorm (src/kernels.jl:15)
  P.invh = 1.0 / h;
  P.r2max = (4.0 * h) * h;                            // T(4)*h*h
  P.r2min = (DBL_EPSILON * h) * h;                    // eps(T)*p.h*p.h
  P.Fc = ((-5.0 * P.A) * P.invh) * P.invh;            // src/forward.jl:87
  P.Gc = ((7.5 * P.A) * P.invh) * P.invh;             // src/adjoint.jl:44
  P.twomumass = (2.0 * p.m) * p.mu;
  P.hx = p.Lx / (double)(p.ngx - 1);                  // dxg (src/params.jl:31)
  P.hy = p.Ly / (double)(p.ngy - 1);
  P.gxmax = (double)(p.ngx - 1);
  P.gymax = (double)(p.ngy - 1);
  P.gxclamp = (double)(p.ngx - 1) - 1e-6;
  P.gyclamp = (double)(p.ngy - 1) - 1e-6;
  P.ngx = (int)p.ngx;
  P.ngy = (int)p.ngy;
  return P;
}

// Constants of the f32c form: formed in FP64 and rounded once to FP32.
PhysF32 make_phys32(const sph_params &p, double rc) {
  PhysF32 P{};
  const double h = p.h;
  const double A = 7.0 / ((4.0 * M_PI) * (h * h));
  P.m = (float)p.m;
  P.rho0 = (float)p.rho0;
  P.c2 = (float)(p.c * p.c);
  P.mu = (float)p.mu;
  P.gx = (float)p.gx;
  P.gy = (float)p.gy;
  P.dt = (float)p.dt;
  P.Lx = (float)p.Lx;
  P.Ly = (float)p.Ly;
  P.kw = (float)p.kw;
  P.mA = (float)(p.m * A);
  P.invh = (float)(1.0 / h);
  P.r2max = (float)(4.0 * h * h);
  P.r2min = (float)(FLT_EPSILON * h * h);
  P.Fc = (float)(-5.0 * A / (h * h));
  P.Gc = (float)(7.5 * A / (h * h));
  P.twomumass = (float)(2.0 * p.m * p.mu);
  P.cs = (float)rc;
  P.dt64 = p.dt;
  P.hx = p.Lx / (double)(p.ngx - 1);
  P.hy = p.Ly / (double)(p.ngy - 1);
  P.inv_hx = (float)(1.0 / P.hx);
  P.inv_hy = (float)(1.0 / P.hy);
  P.inv_hx64 = 1.0 / P.hx;
  P.inv_hy64 = 1.0 / P.hy;
  P.mhh = (float)(-0.5 / h);
  P.twoinvh = (float)(2.0 / h);
  P.mFc = (float)(p.m * (-5.0 * A / (h * h)));
  P.Gch = (float)(7.5 * A / (h * h * h));
  P.gxmax = (double)(p.ngx - 1);
  P.gymax = (double)(p.ngy - 1);
  P.gxclamp = (double)(p.ngx - 1) - 1e-6;
  P.gyclamp = (double)(p.ngy - 1) - 1e-6;
  P.ngx = (int)p.ngx;
  P.ngy = (int)p.ngy;
  return P;
}

ThetaGrid make_theta_grid(const sph_params &p, double rc) {
  ThetaGrid T{};
  T.ngx = (int)p.ngx;
  T.ngy = (int)p.ngy;
  T.hx = p.Lx / (double)(p.ngx - 1);
  T.hy = p.Ly / (double)(p.ngy - 1);
  T.inv_hx = 1.0 / T.hx;
  T.inv_hy = 1.0 / T.hy;
  T.gxclamp = (double)(p.ngx - 1) - 1e-6;
  T.gyclamp = (double)(p.ngy - 1) - 1e-6;
  T.cs = rc;
  return T;
}

// Device copy of an f32c state's scene table and workgroup -> scene maps, for
// each per-slot workgroup size, and the size each tuned kernel launches with.
struct BatchStore {
  Batch B{};              // the 256-thread geometry (untuned kernels)
  // per launch configuration: kSlotBlocks entry v with LDS staging (v) and
  // without it (v + 3, the pair passes then read neighbours from global memory)
  Batch Bv[6]{};
  int cfg[TK_COUNT] = {2, 2, 2, 2, 2, 2};
  std::vector<Scene32> host;
  Scene32 *sc = nullptr;
  int *b2s[3] = {}, *boff[3] = {}, *tb2s = nullptr, *nb2s = nullptr;
  int2 *trange = nullptr, *trange_tight = nullptr;
  const Batch &at(int k) const { return Bv[cfg[k]]; }
};

int blocks_of(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// One scene of an f32c state: its grid, constants and design grid, with the
// offsets left at zero (batch_build lays the scenes out).
Scene32 make_scene32(const sph_params &p, int64_t N, double skin) {
  Scene32 c{};
  const double rc = (2.0 * p.h) * (1.0 + skin);
  long long nx = (long long)std::ceil(p.Lx / rc) + 2;
  long long ny = (long long)std::ceil(p.Ly / rc) + 2;
  nx = nx < 1 ? 1 : nx;
  ny = ny < 1 ? 1 : ny;
  if (nx * ny > (1LL << 30)) fail(SPH_E_ARG, "cell grid %lldx%lld too large", nx, ny);
  if (nx >= (1LL << 24) || ny >= (1LL << 24)) fail(SPH_E_ARG, "cell grid too wide");
  c.g.nx = (int)nx;
  c.g.ny = (int)ny;
  c.g.ncell = (int)(nx * ny);
  c.g.N = (int)N;
  c.g.inv_nx = 1.0 / (double)nx;
  c.P = make_phys32(p, rc);
  c.T = make_theta_grid(p, rc);
  c.cs64 = rc;
  c.dmax2 = (p.h * skin) * (p.h * skin);
  return c;
}

void batch_free(BatchStore &bs) {
  dfree(bs.sc);
  for (int v = 0; v < 3; ++v) {
    dfree(bs.b2s[v]);
    dfree(bs.boff[v]);
  }
  dfree(bs.tb2s);
  dfree(bs.nb2s);
  dfree(bs.trange);
  dfree(bs.trange_tight);
}

// Lays the scenes out (slots, cells, design nodes and cells, workgroups in
// order) and uploads the table and the workgroup maps; cfg keeps the tuned
// workgroup sizes.
void batch_build(BatchStore &bs, std::vector<Scene32> scenes) {
  long long off = 0, coff = 0, toff = 0, tcoff = 0, tboff = 0, nboff = 0;
  std::vector<int2> trange, trange_tight;
  for (auto &c : scenes) {
    // SPH cell ranges of the theta columns, then of the theta rows
    c.troff = (int)trange.size();
    const int ncx = c.T.ngx - 1, ncy = c.T.ngy - 1;
    const double ex = 0.25 * c.T.cs;
    for (int i = 0; i < ncx; ++i) {
      trange.push_back(theta_cell_range(i, ncx, c.T.hx, c.T.cs, c.g.nx));
      trange_tight.push_back(theta_cell_range_tight(i, ncx, c.T.hx, c.T.cs, c.g.nx, ex));
    }
    for (int j = 0; j < ncy; ++j) {
      trange.push_back(theta_cell_range(j, ncy, c.T.hy, c.T.cs, c.g.ny));
      trange_tight.push_back(theta_cell_range_tight(j, ncy, c.T.hy, c.T.cs, c.g.ny, ex));
    }
    const long long nodes = (long long)c.T.ngx * c.T.ngy;
    const long long tcells = (long long)(c.T.ngx - 1) * (c.T.ngy - 1);
    c.off = (int)off;
    c.coff = (int)coff;
    c.toff = (int)toff;
    c.tcoff = (int)tcoff;
    c.tboff = (int)tboff;
    c.nboff = (int)nboff;
    off += c.g.N;
    coff += c.g.ncell;
    toff += nodes;
    tcoff += tcells;
    tboff += theta_strip_blocks(c.T.ngx, c.T.ngy);
    nboff += blocks_of(nodes);
  }
  if (off > kMaxSlotsF32C)
    fail(SPH_E_ARG, "the f32c form addresses its mask lists with 32-bit offsets: N=%lld above %lld",
         off, kMaxSlotsF32C);
  if (coff >= (1LL << 30)) fail(SPH_E_ARG, "too many cells in all scenes (%lld)", coff);
  const int nsc = (int)scenes.size();
  std::vector<int> tb2s((size_t)tboff), nb2s((size_t)nboff);
  for (int k = 0; k < nsc; ++k) {
    const Scene32 &c = scenes[k];
    for (long long b = 0; b < theta_strip_blocks(c.T.ngx, c.T.ngy); ++b) tb2s[c.tboff + b] = k;
    for (int b = 0; b < blocks_of((long long)c.T.ngx * c.T.ngy); ++b) nb2s[c.nboff + b] = k;
  }
  int cfg[TK_COUNT];
  std::copy(bs.cfg, bs.cfg + TK_COUNT, cfg);
  batch_free(bs);
  std::copy(cfg, cfg + TK_COUNT, bs.cfg);
  bs.host = scenes;
  bs.sc = dalloc<Scene32>(scenes.size());
  HIPC(hipMemcpy(bs.sc, scenes.data(), sizeof(Scene32) * scenes.size(), hipMemcpyHostToDevice));
  bs.tb2s = dalloc<int>(tb2s.size() + 1);
  bs.nb2s = dalloc<int>(nb2s.size() + 1);
  bs.trange = dalloc<int2>(trange.size() + 1);
  bs.trange_tight = dalloc<int2>(trange_tight.size() + 1);
  if (!trange.empty()) {
    HIPC(hipMemcpy(bs.trange, trange.data(), sizeof(int2) * trange.size(), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(bs.trange_tight, trange_tight.data(), sizeof(int2) * trange_tight.size(),
                   hipMemcpyHostToDevice));
  }
  if (!tb2s.empty())
    HIPC(hipMemcpy(bs.tb2s, tb2s.data(), sizeof(int) * tb2s.size(), hipMemcpyHostToDevice));
  if (!nb2s.empty())
    HIPC(hipMemcpy(bs.nb2s, nb2s.data(), sizeof(int) * nb2s.size(), hipMemcpyHostToDevice));
  for (int v = 0; v < 3; ++v) {
    const int wg = kSlotBlocks[v];
    std::vector<int> b2s, boff(nsc);
    for (int k = 0; k < nsc; ++k) {
      boff[k] = (int)b2s.size();
      const int nb = (scenes[k].g.N + wg - 1) / wg;
      for (int b = 0; b < nb; ++b) b2s.push_back(k);
    }
    bs.b2s[v] = dalloc<int>(b2s.size() + 1);
    bs.boff[v] = dalloc<int>((size_t)nsc);
    if (!b2s.empty())
      HIPC(hipMemcpy(bs.b2s[v], b2s.data(), sizeof(int) * b2s.size(), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(bs.boff[v], boff.data(), sizeof(int) * nsc, hipMemcpyHostToDevice));
    Batch &B = bs.Bv[v];
    B.sc = bs.sc;
    B.b2s = bs.b2s[v];
    B.boff = bs.boff[v];
    B.tb2s = bs.tb2s;
    B.nb2s = bs.nb2s;
    B.trange = bs.trange;
    B.trange_tight = bs.trange_tight;
    B.bs = wg;
    B.nscenes = nsc;
    B.nblk = (int)b2s.size();
    B.ntblk = (int)tboff;
    B.nnblk = (int)nboff;
    B.N = (int)off;
    B.ncell = (int)coff;
    B.ntnode = (int)toff;
    B.ntcell = (int)tcoff;
  }
  for (int v = 0; v < 3; ++v) {
    bs.Bv[v + 3] = bs.Bv[v];
    bs.Bv[v + 3].nostage = 1;
  }
  bs.B = bs.Bv[2];
}

// the tape keeps its own copy of a state's scene table (same tuned sizes)
void batch_copy(BatchStore &dst, const BatchStore &src) {
  std::copy(src.cfg, src.cfg + TK_COUNT, dst.cfg);
  batch_build(dst, src.host);
}

void check_params(const sph_params *p) {
  if (!p) fail(SPH_E_ARG, "params is NULL");
  if (!(p->h > 0) || !(p->dt > 0) || !(p->Lx > 0) || !(p->Ly > 0))
    fail(SPH_E_ARG, "params: h, dt, Lx, Ly must be positive");
  if (p->ngx < 2 || p->ngy < 2) fail(SPH_E_ARG, "params: ngx, ngy must be >= 2");
}

}  // namespace

// ---------------------------------------------------------------------------

// Kernels timed by the profiler (sph_profile_*).
enum KernelId {
  K_ASSIGN, K_SCAN, K_FILL, K_RANK, K_PERMUTE, K_DENSITY, K_FORCE, K_ADJ_R, K_ADJ_P1, K_ADJ_P2,
  K_REMAP, K_THETA, K_NLBUILD, K_COUNT
};
const char *const kKernelNames[K_COUNT] = {
  "cell_assign", "cell_scan", "cell_fill", "cell_rank", "slot_permute", "density", "force_integrate",
  "adjoint_r", "adjoint_p1", "adjoint_p2", "layout_remap", "theta_gather", "nl_build"};

struct ProfRec {
  int kid;
  hipEvent_t a, b;
};

struct sph_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;  // false: the caller's stream (sph_ctx_create_stream)
  double *scratch = nullptr;  // this device's copy of `scratch` (scratch_address)
  // objects created on this context plus the handle itself; the context is
  // freed when the handle has been destroyed and the last object is gone
  int refs = 1;
  // optional per-kernel timing with events on this stream
  bool prof = false;
  int prof_only = -1;  // >= 0: time only this kernel id
  std::vector<ProfRec> pending;
  std::vector<hipEvent_t> evpool;
  double tot_ms[K_COUNT] = {};
  int64_t launches[K_COUNT] = {};
};

// Scratch memory of the allocation-free entry point
// (sph_target_objective_noalloc): the objective's per-workgroup partial sums,
// kScratchPartials of them at most (N <= kScratchPartials * 256 particles).
// Each GPU the library runs on holds its own copy until the process ends;
// every call on a GPU shares that copy, so calls that overlap on one GPU
// race.  Host code never names the array after set-up: scratch_address()
// gives the current device's copy, which the context keeps.
constexpr int kScratchPartials = 16384;
static __device__ __align__(256) double scratch[kScratchPartials + 2];

namespace {

// the address of the current device's copy of `scratch`
double *scratch_address() {
  void *p = nullptr;
  HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) fail(SPH_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  return static_cast<double *>(p);
}

void ctx_retain(sph_ctx *c) { c->refs += 1; }

void ctx_release(sph_ctx *c) {
  if (--c->refs > 0) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto &r : c->pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : c->evpool) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

hipEvent_t prof_event(sph_ctx *c) {
  if (!c->evpool.empty()) {
    hipEvent_t e = c->evpool.back();
    c->evpool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIPC(hipEventCreate(&e));
  return e;
}

// Runs f (one or more launches on c->stream); with profiling on, brackets it
// with events and books the interval under kernel id kid.
template <class F>
void timed(sph_ctx *c, int kid, F &&f) {
  if (!c->prof || (c->prof_only >= 0 && c->prof_only != kid)) {
    f();
    return;
  }
  hipEvent_t a = prof_event(c), b = prof_event(c);
  HIPC(hipEventRecord(a, c->stream));
  f();
  HIPC(hipEventRecord(b, c->stream));
  c->pending.push_back({kid, a, b});
}

// after a stream synchronisation
void prof_collect(sph_ctx *c) {
  for (auto &r : c->pending) {
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, r.a, r.b));
    c->tot_ms[r.kid] += ms;
    c->launches[r.kid] += 1;
    c->evpool.push_back(r.a);
    c->evpool.push_back(r.b);
  }
  c->pending.clear();
}

}  // namespace

struct sph_field {
  sph_ctx *ctx = nullptr;
  int64_t ngy = 0, ngx = 0;
  double *d = nullptr;
};

struct sph_buffer {
  sph_ctx *ctx = nullptr;
  int64_t nbytes = 0;
  unsigned char *d = nullptr;
};

struct sph_state {
  sph_ctx *ctx = nullptr;
  sph_params p{};
  int precision = SPH_F64;
  int64_t N = 0;
  Grid g{};
  PhysF64 P{};
  PhysF32 P32{};
  BatchStore bs;          // f32c form: scene table (one entry unless batched)
  int64_t ntnode = 0;     // design-field nodes over all scenes
  CellConstF64 kc{};
  double skin = 0.2;
  int64_t interval = 4, maxnb = 64, sm = 0, si = 0;
  int layout = SPH_LAYOUT_SLOT;
  int64_t age = kAgeNever;
  uint64_t epoch = 0;  // serial number of the current slot layout
  // f32c form: the layout of this epoch is built but its mask list is not;
  // the next step's density pass builds it (set by the checkpointer, which
  // rebuilds before a snapshot and restores the rebuilt layout)
  bool pending_list = false;
  // f32c form, one scene: sweep 5 has each rebuild order every cell's slots by
  // quadrant (cell_quadrant) and build the mask list by the 5x5 half-cell
  // sweep; sweep 3 (the default) keeps the reference's (cell, id) order and
  // the 3x3 sweep.  layout_quad: the current layout is quadrant-ordered, and
  // qs holds its quadrant starts (4 ncell + 1 entries).
  int sweep = 3;
  bool layout_quad = false;
  int *qs = nullptr;
  // current layout; for the f32c form scell holds the binning cell of each
  // slot, for the FP64 form bcell does (scell is the sweep cell there)
  int *ids = nullptr, *scell = nullptr, *cstart = nullptr, *bcell = nullptr;
  float2 *V32 = nullptr, *V32n = nullptr, *xr = nullptr, *xrn = nullptr, *PI32 = nullptr;
  // neighbour list of the current layout: FP64 form transposed (entry m of
  // slot s at nl[m*N + s]); f32c form row bitmasks (nm)
  int *nl = nullptr, *nbcnt = nullptr;
  uint32_t *nm = nullptr;
  uint32_t *em = nullptr;  // f32c form: exact list of the last density pass
  int *lane = nullptr;     // f32c form: thread -> slot order of the pair passes
  double2 *X = nullptr, *V = nullptr, *Xn = nullptr, *Vn = nullptr;
  double2 *Xref = nullptr, *Xref2 = nullptr;
  double2 *PI = nullptr;
  // build! buffers (cell list of the last build)
  int *cellof = nullptr, *counts = nullptr, *cl_cstart = nullptr, *cursor = nullptr;
  int *tmp_id = nullptr, *tmp_slot = nullptr, *perm = nullptr, *new_ids = nullptr;
  int *bsum = nullptr;
  int *bigcells = nullptr;  // cells of the last build with more than 128 entries (rank)
  int cl_view = 0;  // where the last build's (starts, order) live: 0 none, 1 cl buffers, 2 layout
  int *flags = nullptr;
  // auxiliary outputs: 0 = zeros (no step yet), 1 = recomputable from the
  // pre-step buffers (Xn, Vn) of the last step, 2 = materialised in id order
  int aux_state = 0;
  double2 *a_id = nullptr, *pi_id = nullptr, *a_slot = nullptr;
  double *alpha_id = nullptr, *rho_id = nullptr, *alpha_slot = nullptr, *rho_slot = nullptr;
  double *theta_last = nullptr;  // copy of the design field of the last step (ngy*ngx)
  // materialised neighbour list (built on request)
  int *nl_idx = nullptr, *nl_cnt = nullptr;
  // staging
  double2 *tmp2 = nullptr, *seed_slot = nullptr;
  // captured forward cycle (two rebuild epochs) of sph_simulate_graph
  hipGraphExec_t gexec = nullptr;
  const double *gtheta_d = nullptr;   // design field the graph reads
  std::vector<const void *> groles;   // buffer roles at the start of the captured cycle
  double *tmp1 = nullptr;
  double *obj_part = nullptr, *obj_J = nullptr;
  int obj_nb = 0;
};

struct TapeEpoch {
  uint64_t serial = 0;
  const sph_state *owner = nullptr;
  int64_t first = 0;  // first tape entry of the epoch
  int *ids = nullptr, *scell = nullptr, *cstart = nullptr;
  int *bcell = nullptr;  // FP64 form: binning cells (the f32c form's scell is its binning cell)
  int *lane = nullptr;   // f32c form: the pair passes' thread order of the epoch
};

struct sph_tape {
  sph_ctx *ctx = nullptr;
  int64_t N = 0, cap = 0, n = 0;
  int precision = SPH_F64;
  double2 *X = nullptr, *V = nullptr;  // (cap+1) entries of N; entry k (1-based) at (k-1)*N
  float2 *V32 = nullptr;                // f32c form: FP32 velocities instead of V
  // optional (pterm, invrho) of every step, so that the reverse sweep skips
  // the density recompute (used by the checkpointed driver's epoch tapes)
  bool keep_pi = false;
  double2 *PI = nullptr;
  float2 *PI32 = nullptr;
  // f32c form with keep_pi: each step's exact neighbour masks (the density
  // pass's em list, 3*kMaskWords words per slot), so that the reverse sweep's
  // P1 and P2 walk them without testing r^2
  uint32_t *EM = nullptr;
  std::vector<int> entry_epoch;         // epoch record of each entry (cap+1)
  std::vector<TapeEpoch> epochs;        // records in use
  std::vector<TapeEpoch> pool;          // retained allocations
  bool has_geom = false;
  Grid g{};
  PhysF64 P{};
  double rc = 0.0;  // cell width of the recorded epochs
  CellConstF64 kc{};
  int ncell = 0;
  BatchStore bs;  // f32c form: the scene table of the recorded state
};

struct sph_workspace {
  sph_ctx *ctx = nullptr;
  int64_t N = 0;
  sph_params p{};
  int precision = SPH_F64;
  double2 *gX = nullptr, *gV = nullptr, *gXs = nullptr, *abar = nullptr, *PI = nullptr;
  double2 *Vtmp = nullptr, *t2 = nullptr, *gX_id = nullptr, *gV_id = nullptr;
  double *galpha = nullptr, *grho = nullptr, *gtheta = nullptr;
  int *inv = nullptr;
  double4 *tpart = nullptr;  // per theta cell partial sums of the gtheta gather
  int *nl = nullptr, *nbcnt = nullptr, *flags = nullptr;  // neighbour list of the epoch
  uint32_t *nm = nullptr;                                  // f32c form: row bitmasks
  uint32_t *em = nullptr;                                  // f32c form: exact list of a step
  int *lane = nullptr;                                     // f32c form: pair-pass thread order
  int *tkey = nullptr;        // interp-cell key of each slot (keyed gtheta gather)
  unsigned long long *tlev = nullptr;  // gtheta sweep level of the reverse steps (raise_level)
  int tstamp = 0;             // stamp of the current reverse step (1, 2, ...)
  // FP64 form: the one-cell gtheta sweep ranges (theta columns, then rows) of
  // the cell grid they were made for
  int2 *trange64 = nullptr;
  int tr_nx = -1, tr_ny = -1;
  double tr_cs = 0.0;
  int64_t nl_cap = 0;                                      // maxnb the list was sized for
  int64_t ntnode = 0, ntcell = 0;  // design-field nodes and cells (all scenes)
  int *ids_cur = nullptr;  // slot order of the adjoint state between reverse sweeps
  int *iota = nullptr;     // particle-id order as a slot order
  // f32c form
  float2 *gX32 = nullptr, *gV32 = nullptr, *gXs32 = nullptr, *abar32 = nullptr, *PI32 = nullptr;
  float2 *Vtmp32 = nullptr, *t2f = nullptr, *xr32 = nullptr;
  float2 *xr32b = nullptr;   // relative coordinates of the next step (R pass fused into P2)
  float *galpha32 = nullptr, *grho32 = nullptr;
};

namespace {

void state_free(sph_state *st) {
  int **ints[] = {&st->ids, &st->scell, &st->cstart, &st->cellof, &st->counts, &st->cl_cstart,
                  &st->cursor, &st->tmp_id, &st->tmp_slot, &st->perm, &st->new_ids, &st->bsum,
                  &st->bigcells,
                  &st->flags, &st->nl_idx, &st->nl_cnt, &st->nl, &st->nbcnt, &st->lane,
                  &st->bcell, &st->qs};
  for (auto pp : ints) dfree(*pp);
  double2 **d2[] = {&st->X, &st->V, &st->Xn, &st->Vn, &st->Xref, &st->Xref2, &st->PI, &st->a_id,
                    &st->pi_id, &st->a_slot, &st->tmp2, &st->seed_slot};
  for (auto pp : d2) dfree(*pp);
  float2 **f2[] = {&st->V32, &st->V32n, &st->xr, &st->xrn, &st->PI32};
  for (auto pp : f2) dfree(*pp);
  dfree(st->nm);
  dfree(st->em);
  batch_free(st->bs);
  if (st->gexec) (void)hipGraphExecDestroy(st->gexec);
  st->gexec = nullptr;
  double **d1[] = {&st->alpha_id, &st->rho_id, &st->alpha_slot, &st->rho_slot, &st->theta_last,
                   &st->tmp1, &st->obj_part, &st->obj_J};
  for (auto pp : d1) dfree(*pp);
}

// build! (src/neighbors.jl:97-114): cell list of the current X into the cl buffers.
// build! (src/neighbors.jl:97-114); with quad_keys each cell's entries are
// ranked by (quadrant, id) instead of id, for a quadrant-ordered layout
void do_build(sph_state *st, bool quad_keys = false) {
  hipStream_t s = st->ctx->stream;
  const int N = (int)st->N;
  sph_ctx *c = st->ctx;
  timed(c, K_ASSIGN, [&] {
    // counts are zero here: cleared at creation, then by each build's scan
    if (st->precision == SPH_F32C)
      launch_assign_f32c(s, st->X, st->bs.B, st->cellof, st->counts, st->flags);
    else launch_assign(s, st->X, st->g, st->kc.cs, st->cellof, st->counts, st->flags);
  });
  timed(c, K_SCAN, [&] {
    launch_scan_excl(s, st->counts, st->g.ncell, st->cl_cstart, st->cursor, st->bsum,
                     st->bigcells);
  });
  timed(c, K_FILL, [&] {
    if (quad_keys)
      launch_fill(s, st->cellof, st->ids, N, st->cursor, st->tmp_id, st->tmp_slot, st->X, st->g,
                  st->kc.cs);
    else launch_fill(s, st->cellof, st->ids, N, st->cursor, st->tmp_id, st->tmp_slot);
  });
  timed(c, K_RANK, [&] {
    launch_rank(s, st->cellof, st->cl_cstart, st->tmp_id, st->tmp_slot, N, st->bigcells,
                st->perm, st->new_ids, quad_keys ? st->ids : nullptr);
  });
  st->cl_view = 1;  // the cell list of this build is (cl_cstart, new_ids)
}

// st.a, st.alpha, st.rho, st.pterm, st.invrho of the last step, recomputed
// from that step's input state (still held in the spare buffers Xn, Vn) and
// stored in particle-id order.
void materialise_aux(sph_state *st) {
  if (st->aux_state != 1) return;
  hipStream_t s = st->ctx->stream;
  const int N = (int)st->N;
  if (st->precision == SPH_F32C) {
    launch_density_f32c(s, st->xrn, st->scell, st->nm, st->cstart, st->lane,
                        st->bs.at(TK_DENSITY), st->PI32, st->rho_slot, st->em);
    launch_force_f32c(s, st->Xn, st->xrn, st->V32n, st->PI32, st->theta_last, st->scell, st->em,
                      st->cstart, st->lane, st->bs.at(TK_FORCE), nullptr, nullptr, nullptr,
                      nullptr, nullptr, st->a_slot, st->alpha_slot);
    launch_scatter2f(s, st->PI32, st->ids, N, st->pi_id);
  } else {
    const EpochCells ec{st->bcell, st->scell, st->cstart};
    launch_density_f64(s, st->Xn, st->nl, st->nbcnt, st->g, st->P, ec, st->PI, st->rho_slot,
                       nullptr);
    launch_force_f64(s, st->Xn, st->Vn, st->PI, st->theta_last, st->nl, st->nbcnt, st->g, st->P,
                     ec, nullptr, nullptr, nullptr, st->a_slot, st->alpha_slot);
    launch_scatter2(s, st->PI, st->ids, N, st->pi_id);
  }
  launch_scatter2(s, st->a_slot, st->ids, N, st->a_id);
  launch_scatter1(s, st->alpha_slot, st->ids, N, st->alpha_id);
  launch_scatter1(s, st->rho_slot, st->ids, N, st->rho_id);
  st->aux_state = 2;
}

// build_neighbors! (src/neighbors.jl:252-263): build, then move every slot to
// its (cell, id) position and start a new layout epoch.  With defer_list (f32c
// form) the mask list is left to the step's fused density pass.
void do_build_neighbors(sph_state *st, bool defer_list = false) {
  hipStream_t s = st->ctx->stream;
  const bool quad = st->sweep == 5;
  do_build(st, quad);
  int check = (st->age != kAgeNever) ? 1 : 0;
  if (st->precision == SPH_F32C) {
    timed(st->ctx, K_PERMUTE, [&] {
      launch_permute_f32c(s, st->perm, st->cellof, st->X, st->V32, st->Xref, check, st->bs.B,
                          st->Xn, st->V32n, st->xr, st->Xref2, st->scell, st->flags, st->lane);
    });
    std::swap(st->V32, st->V32n);
  } else {
    timed(st->ctx, K_PERMUTE, [&] {
      launch_permute(s, st->perm, st->cellof, st->X, st->V, st->Xref, check, st->g, st->kc,
                     st->Xn, st->Vn, st->Xref2, st->scell, st->bcell, st->flags);
    });
    std::swap(st->V, st->Vn);
  }
  std::swap(st->X, st->Xn);
  std::swap(st->Xref, st->Xref2);
  std::swap(st->ids, st->new_ids);
  std::swap(st->cstart, st->cl_cstart);
  st->cl_view = 2;  // the cell list of this build is now (cstart, ids) of the layout
  st->layout_quad = quad;
  if (quad) launch_quad_starts(s, st->X, st->cstart, st->g, st->kc.cs, st->qs);
  timed(st->ctx, K_NLBUILD, [&] {
    if (st->precision == SPH_F32C) {
      // (the permute above has dealt the new thread order)
      if (!defer_list)
        launch_nm_build_f32c(s, st->xr, st->scell, st->cstart, st->bs.at(TK_NLBUILD), st->nm,
                             st->flags, quad ? st->qs : nullptr);
    } else
      launch_nl_build_f64(s, st->X, st->scell, st->cstart, st->g, st->kc, (int)st->maxnb, st->nl,
                          st->nbcnt, st->flags);
  });
  st->epoch += 1;
  st->age = 0;
}

}  // namespace

namespace {

TapeEpoch &tape_epoch_for(sph_tape *tp, const sph_state *st) {
  if (!tp->epochs.empty() && tp->epochs.back().owner == st && tp->epochs.back().serial == st->epoch)
    return tp->epochs.back();
  if (!tp->has_geom) {
    tp->g = st->g;
    tp->P = st->P;
    tp->rc = st->kc.cs;
    tp->kc = st->kc;
    tp->ncell = st->g.ncell;
    if (st->precision == SPH_F32C) batch_copy(tp->bs, st->bs);
    tp->has_geom = true;
  } else if (tp->ncell != st->g.ncell || tp->g.nx != st->g.nx) {
    fail(SPH_E_ARG, "tape already holds steps of a state with another cell grid");
  }
  TapeEpoch e;
  if (!tp->pool.empty()) {
    e = tp->pool.back();
    tp->pool.pop_back();
  } else {
    e.ids = dalloc<int>(tp->N);
    e.scell = dalloc<int>(tp->N);
    e.cstart = dalloc<int>(tp->ncell + 1);
  }
  if (st->bcell && !e.bcell) e.bcell = dalloc<int>(tp->N);
  if (st->precision == SPH_F32C && !e.lane) e.lane = dalloc<int>(tp->N);
  e.owner = st;
  e.serial = st->epoch;
  e.first = tp->n + 1;
  const bool f32c = st->precision == SPH_F32C;
  launch_epoch_record(st->ctx->stream,
                      EpochLayout{st->ids, st->scell, st->bcell, f32c ? st->lane : nullptr,
                                  st->cstart},
                      (int)tp->N, tp->ncell + 1,
                      EpochLayoutOut{e.ids, e.scell, e.bcell, e.lane, e.cstart});
  tp->epochs.push_back(e);
  return tp->epochs.back();
}

// step! (src/forward.jl:161-173); with a tape, entry k (1-based) receives the
// state before the step (written by the kernels themselves).
void do_step(sph_state *st, const double *theta, sph_tape *tp) {
  hipStream_t s = st->ctx->stream;
  // maybe_rebuild!; the f32c form builds the new mask list in the step's
  // density pass (the same list, one sweep fewer)
  const bool rebuild = st->age >= st->interval;
  const bool fused = st->precision == SPH_F32C && (rebuild || st->pending_list);
  if (rebuild) do_build_neighbors(st, fused);
  st->pending_list = false;
  st->age += 1;
  double2 *tX = nullptr, *tV = nullptr;
  float2 *tV32 = nullptr;
  double2 *PI = st->PI;
  float2 *PI32 = st->PI32;
  uint32_t *em = st->em;
  if (tp) {
    tape_epoch_for(tp, st);
    tp->n += 1;
    tp->entry_epoch[tp->n] = (int)tp->epochs.size() - 1;
    const size_t off = (size_t)(tp->n - 1) * tp->N;
    tX = tp->X + off;
    if (st->precision == SPH_F32C) tV32 = tp->V32 + off;
    else tV = tp->V + off;
    if (tp->keep_pi) {
      if (st->precision == SPH_F32C) PI32 = tp->PI32 + off;
      else PI = tp->PI + off;
    }
    if (tp->EM) em = tp->EM + off * (3 * kMaskWords);
  }
  if (st->precision == SPH_F32C) {
    timed(st->ctx, K_DENSITY, [&] {
      if (fused)
        launch_density_build_f32c(s, st->xr, st->scell, st->cstart, st->lane,
                                  st->bs.at(TK_DENSITY), PI32, st->nm, em, st->flags,
                                  st->layout_quad ? st->qs : nullptr);
      else
        launch_density_f32c(s, st->xr, st->scell, st->nm, st->cstart, st->lane,
                            st->bs.at(TK_DENSITY), PI32, nullptr, em);
    });
    timed(st->ctx, K_FORCE, [&] {
      launch_force_f32c(s, st->X, st->xr, st->V32, PI32, theta, st->scell, em, st->cstart,
                        st->lane, st->bs.at(TK_FORCE), st->Xn, st->xrn, st->V32n, tX, tV32,
                        nullptr, nullptr);
    });
    std::swap(st->X, st->Xn);
    std::swap(st->xr, st->xrn);
    std::swap(st->V32, st->V32n);
    st->aux_state = 1;
    return;
  }
  const EpochCells ec{st->bcell, st->scell, st->cstart};
  timed(st->ctx, K_DENSITY, [&] {
    launch_density_f64(s, st->X, st->nl, st->nbcnt, st->g, st->P, ec, PI, nullptr, tX);
  });
  timed(st->ctx, K_FORCE, [&] {
    launch_force_f64(s, st->X, st->V, PI, theta, st->nl, st->nbcnt, st->g, st->P, ec, st->Xn,
                     st->Vn, tV, nullptr, nullptr);
  });
  std::swap(st->X, st->Xn);
  std::swap(st->V, st->Vn);
  st->aux_state = 1;
}

bool is_batched(const sph_state *st) {
  return st->precision == SPH_F32C && st->bs.B.nscenes > 1;
}

void need_precision(int precision) {
  if (precision != SPH_F64 && precision != SPH_F32C) fail(SPH_E_ARG, "unknown precision %d", precision);
}

void copy_theta(sph_state *st, const sph_field *th) {
  if (!th) fail(SPH_E_ARG, "theta is NULL");
  if (st->bs.B.nscenes > 1) {
    if (th->ngy * th->ngx != st->ntnode)
      fail(SPH_E_ARG, "batched theta has %lld nodes, the scenes have %lld",
           (long long)(th->ngy * th->ngx), (long long)st->ntnode);
  } else if (th->ngy != st->p.ngy || th->ngx != st->p.ngx) {
    fail(SPH_E_ARG, "theta is (%lld, %lld) but params want (%lld, %lld)", (long long)th->ngy,
         (long long)th->ngx, (long long)st->p.ngy, (long long)st->p.ngx);
  }
  HIPC(hipMemcpyAsync(st->theta_last, th->d, sizeof(double) * th->ngy * th->ngx,
                      hipMemcpyDeviceToDevice, st->ctx->stream));
}

}  // namespace

extern "C" {

const char *sph_last_error(void) { return g_err.c_str(); }

int sph_ctx_create(int device, sph_ctx **out) {
  return guarded([&] {
    if (!out) fail(SPH_E_ARG, "out is NULL");
    int n = 0;
    HIPC(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) fail(SPH_E_ARG, "device %d not present (%d devices)", device, n);
    HIPC(hipSetDevice(device));
    double *scr = scratch_address();
    auto *c = new sph_ctx;
    c->device = device;
    c->scratch = scr;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      hip_check(e, "hipStreamCreateWithFlags");
    }
    *out = c;
  });
}

int sph_ctx_create_stream(int device, void *stream, sph_ctx **out) {
  return guarded([&] {
    if (!out) fail(SPH_E_ARG, "out is NULL");
    if (!stream) fail(SPH_E_ARG, "stream is NULL (use sph_ctx_create for a stream of the library's own)");
    int n = 0;
    HIPC(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) fail(SPH_E_ARG, "device %d not present (%d devices)", device, n);
    HIPC(hipSetDevice(device));
    hipStream_t hs = (hipStream_t)stream;
    int sdev = -1;
    HIPC(hipStreamGetDevice(hs, &sdev));
    if (sdev != device) fail(SPH_E_ARG, "the stream belongs to device %d, not %d", sdev, device);
    double *scr = scratch_address();
    auto *c = new sph_ctx;
    c->device = device;
    c->scratch = scr;
    c->stream = hs;
    c->own_stream = false;
    *out = c;
  });
}

int sph_mem_info(sph_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes) {
  return guarded([&] {
    if (!ctx || !free_bytes || !total_bytes) fail(SPH_E_ARG, "NULL argument");
    HIPC(hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    HIPC(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
  });
}

int sph_ctx_destroy(sph_ctx *ctx) {
  return guarded([&] {
    if (!ctx) return;
    ctx_release(ctx);
  });
}

int sph_sync(sph_ctx *ctx) {
  return guarded([&] {
    if (!ctx) fail(SPH_E_ARG, "ctx is NULL");
    HIPC(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
  });
}

int sph_profile_enable(sph_ctx *ctx, int on) {
  return guarded([&] {
    if (!ctx) fail(SPH_E_ARG, "ctx is NULL");
    HIPC(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    if (on < 0 || on >= 2 + K_COUNT) fail(SPH_E_ARG, "profile mode %d out of range", on);
    ctx->prof = on != 0;
    ctx->prof_only = on >= 2 ? on - 2 : -1;
    for (int k = 0; k < K_COUNT; ++k) {
      ctx->tot_ms[k] = 0.0;
      ctx->launches[k] = 0;
    }
  });
}

int sph_profile_get(sph_ctx *ctx, int kernel, int64_t *launches, double *total_ms) {
  return guarded([&] {
    if (!ctx || !launches || !total_ms) fail(SPH_E_ARG, "NULL argument");
    if (kernel < 0 || kernel >= K_COUNT) fail(SPH_E_ARG, "kernel id %d out of range", kernel);
    HIPC(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    *launches = ctx->launches[kernel];
    *total_ms = ctx->tot_ms[kernel];
  });
}

const char *sph_profile_name(int kernel) {
  return (kernel >= 0 && kernel < K_COUNT) ? kKernelNames[kernel] : nullptr;
}

int sph_field_create(sph_ctx *ctx, int64_t ngy, int64_t ngx, const double *host, sph_field **out) {
  return guarded([&] {
    if (!ctx || !out) fail(SPH_E_ARG, "NULL argument");
    if (ngy < 1 || ngx < 1 || ngy * ngx < 4) fail(SPH_E_ARG, "field must hold at least 4 nodes");
    HIPC(hipSetDevice(ctx->device));
    auto *f = new sph_field;
    f->ctx = ctx;
    f->ngy = ngy;
    f->ngx = ngx;
    f->d = dalloc<double>(ngy * ngx);
    if (host) HIPC(hipMemcpy(f->d, host, sizeof(double) * ngy * ngx, hipMemcpyHostToDevice));
    else HIPC(hipMemset(f->d, 0, sizeof(double) * ngy * ngx));
    // the context's stream does not order itself after the null stream
    HIPC(hipDeviceSynchronize());
    ctx_retain(ctx);
    *out = f;
  });
}

int sph_field_set(sph_field *f, const double *host) {
  return guarded([&] {
    if (!f || !host) fail(SPH_E_ARG, "NULL argument");
    HIPC(hipStreamSynchronize(f->ctx->stream));
    HIPC(hipMemcpy(f->d, host, sizeof(double) * f->ngy * f->ngx, hipMemcpyHostToDevice));
  });
}

int sph_field_get(const sph_field *f, double *host) {
  return guarded([&] {
    if (!f || !host) fail(SPH_E_ARG, "NULL argument");
    HIPC(hipStreamSynchronize(f->ctx->stream));
    HIPC(hipMemcpy(host, f->d, sizeof(double) * f->ngy * f->ngx, hipMemcpyDeviceToHost));
  });
}

int sph_field_destroy(sph_field *f) {
  return guarded([&] {
    if (!f) return;
    (void)hipStreamSynchronize(f->ctx->stream);
    dfree(f->d);
    sph_ctx *c = f->ctx;
    delete f;
    ctx_release(c);
  });
}

// ---- design-field chain (scripts/03_optimize.jl:142-244) -------------------

struct sph_design {
  sph_ctx *ctx = nullptr;
  DesignGrid D{};
  double *wtab = nullptr;   // (r+1) - sqrt(k) for squared grid distance k
  double *snorm = nullptr;  // normalisation of each node's weight list
  double *grdf = nullptr;   // scratch: gtheta .* dalpha(rdf)
  double *gmax = nullptr;   // scratch: max|g| of the last candidate
};

namespace {
void need_grid(const sph_design *d, const sph_field *f, const char *what) {
  if (!f) fail(SPH_E_ARG, "%s is NULL", what);
  if (f->ngy != d->D.ngy || f->ngx != d->D.ngx)
    fail(SPH_E_ARG, "%s is (%lld, %lld), the design grid is (%d, %d)", what, (long long)f->ngy,
         (long long)f->ngx, d->D.ngy, d->D.ngx);
}
}  // namespace

int sph_design_create(sph_ctx *ctx, int64_t ngy, int64_t ngx, int64_t radius, double alpha_max,
                      double qp, sph_design **out) {
  return guarded([&] {
    if (!ctx || !out) fail(SPH_E_ARG, "NULL argument");
    if (ngy < 1 || ngx < 1 || ngy * ngx >= (1LL << 30)) fail(SPH_E_ARG, "bad design grid");
    if (radius < 0 || radius > 64) fail(SPH_E_ARG, "filter radius must be in [0, 64]");
    if (!(qp > 0.0)) fail(SPH_E_ARG, "qp must be positive");
    HIPC(hipSetDevice(ctx->device));
    auto *d = new sph_design;
    d->ctx = ctx;
    d->D = DesignGrid{(int)ngy, (int)ngx, (int)radius, alpha_max, qp};
    // build_filter (scripts/03_optimize.jl:156-164): w = (r+1) - sqrt(di^2 + dj^2),
    // kept when w > 0; s = sum of the kept w over the box, jj outer, ii inner
    const int r = (int)radius;
    std::vector<double> wtab((size_t)(2 * r * r + 1));
    for (int k = 0; k <= 2 * r * r; ++k) wtab[k] = (double)(r + 1) - std::sqrt((double)k);
    std::vector<double> snorm((size_t)(ngy * ngx));
    for (int64_t j = 0; j < ngy; ++j)
      for (int64_t i = 0; i < ngx; ++i) {
        double sum = 0.0;
        for (int64_t jj = std::max<int64_t>(0, j - r); jj <= std::min<int64_t>(ngy - 1, j + r); ++jj)
          for (int64_t ii = std::max<int64_t>(0, i - r); ii <= std::min<int64_t>(ngx - 1, i + r);
               ++ii) {
            const double w = wtab[(size_t)((ii - i) * (ii - i) + (jj - j) * (jj - j))];
            if (w > 0.0) sum += w;
          }
        snorm[(size_t)(j + i * ngy)] = sum;
      }
    d->wtab = dalloc<double>(wtab.size());
    d->snorm = dalloc<double>(snorm.size());
    d->grdf = dalloc<double>(snorm.size());
    d->gmax = dalloc<double>(1);
    HIPC(hipMemcpy(d->wtab, wtab.data(), sizeof(double) * wtab.size(), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d->snorm, snorm.data(), sizeof(double) * snorm.size(), hipMemcpyHostToDevice));
    ctx_retain(ctx);
    *out = d;
  });
}

int sph_design_destroy(sph_design *d) {
  return guarded([&] {
    if (!d) return;
    (void)hipStreamSynchronize(d->ctx->stream);
    dfree(d->wtab);
    dfree(d->snorm);
    dfree(d->grdf);
    dfree(d->gmax);
    sph_ctx *c = d->ctx;
    delete d;
    ctx_release(c);
  });
}

int sph_design_theta(sph_design *d, const sph_field *rd, sph_field *rdf, sph_field *theta,
                     int64_t offset) {
  return guarded([&] {
    if (!d) fail(SPH_E_ARG, "design is NULL");
    need_grid(d, rd, "rd");
    need_grid(d, rdf, "rdf");
    if (rd == rdf) fail(SPH_E_ARG, "rd and rdf must be distinct fields");
    if (!theta) fail(SPH_E_ARG, "theta is NULL");
    const int64_t n = (int64_t)d->D.ngy * d->D.ngx;
    if (offset < 0 || offset + n > theta->ngy * theta->ngx)
      fail(SPH_E_ARG, "theta holds %lld nodes, need %lld", (long long)(theta->ngy * theta->ngx),
           (long long)(offset + n));
    HIPC(hipSetDevice(d->ctx->device));
    hipStream_t s = d->ctx->stream;
    launch_design_filter(s, rd->d, d->wtab, d->snorm, d->D, rdf->d);
    launch_design_alpha(s, rdf->d, d->D, theta->d + offset);
    HIPC(hipGetLastError());
  });
}

int sph_design_grad(sph_design *d, const sph_field *rdf, const sph_workspace *ws, int64_t offset,
                    sph_field *grd) {
  return guarded([&] {
    if (!d || !ws) fail(SPH_E_ARG, "NULL argument");
    need_grid(d, rdf, "rdf");
    need_grid(d, grd, "grd");
    if (rdf == grd) fail(SPH_E_ARG, "rdf and grd must be distinct fields");
    const int64_t n = (int64_t)d->D.ngy * d->D.ngx;
    const int64_t wn = ws->ntnode > 0 ? ws->ntnode : (int64_t)ws->p.ngy * ws->p.ngx;
    if (offset < 0 || offset + n > wn)
      fail(SPH_E_ARG, "the workspace's gtheta holds %lld nodes, need %lld", (long long)wn,
           (long long)(offset + n));
    HIPC(hipSetDevice(d->ctx->device));
    hipStream_t s = d->ctx->stream;
    launch_design_chain(s, ws->gtheta + offset, rdf->d, d->D, d->grdf);
    launch_design_filter_adjoint(s, d->grdf, d->wtab, d->snorm, d->D, grd->d);
    HIPC(hipGetLastError());
  });
}

int sph_design_project(sph_design *d, const sph_field *x, double volfrac, sph_field *out) {
  return guarded([&] {
    if (!d) fail(SPH_E_ARG, "design is NULL");
    need_grid(d, x, "x");
    need_grid(d, out, "out");
    if (x == out) fail(SPH_E_ARG, "x and out must be distinct fields");
    HIPC(hipSetDevice(d->ctx->device));
    launch_design_project(d->ctx->stream, x->d, nullptr, 0.0, d->D.ngy * d->D.ngx, volfrac,
                          out->d, nullptr);
    HIPC(hipGetLastError());
  });
}

int sph_design_candidate(sph_design *d, const sph_field *rd, const sph_field *g, double step,
                         double volfrac, sph_field *out, double *gmax) {
  return guarded([&] {
    if (!d) fail(SPH_E_ARG, "design is NULL");
    need_grid(d, rd, "rd");
    need_grid(d, g, "g");
    need_grid(d, out, "out");
    if (out == rd || out == g) fail(SPH_E_ARG, "out must be distinct from rd and g");
    HIPC(hipSetDevice(d->ctx->device));
    hipStream_t s = d->ctx->stream;
    launch_design_project(s, rd->d, g->d, step, d->D.ngy * d->D.ngx, volfrac, out->d, d->gmax);
    HIPC(hipGetLastError());
    if (gmax) {
      HIPC(hipMemcpyAsync(gmax, d->gmax, sizeof(double), hipMemcpyDeviceToHost, s));
      HIPC(hipStreamSynchronize(s));
    }
  });
}

int sph_buffer_create(sph_ctx *ctx, int64_t nbytes, const void *host, sph_buffer **out) {
  return guarded([&] {
    if (!ctx || !out) fail(SPH_E_ARG, "NULL argument");
    if (nbytes < 1) fail(SPH_E_ARG, "buffer size must be >= 1 byte");
    HIPC(hipSetDevice(ctx->device));
    auto *b = new sph_buffer;
    b->ctx = ctx;
    b->nbytes = nbytes;
    try {
      b->d = dalloc<unsigned char>((size_t)nbytes);
      hipStream_t s = ctx->stream;
      if (host) HIPC(hipMemcpyAsync(b->d, host, (size_t)nbytes, hipMemcpyHostToDevice, s));
      else HIPC(hipMemsetAsync(b->d, 0, (size_t)nbytes, s));
      HIPC(hipStreamSynchronize(s));
    } catch (...) {
      dfree(b->d);
      delete b;
      throw;
    }
    ctx_retain(ctx);
    *out = b;
  });
}

int sph_buffer_destroy(sph_buffer *b) {
  return guarded([&] {
    if (!b) return;
    (void)hipStreamSynchronize(b->ctx->stream);
    dfree(b->d);
    sph_ctx *c = b->ctx;
    delete b;
    ctx_release(c);
  });
}

int sph_buffer_ptr(const sph_buffer *b, void **ptr) {
  return guarded([&] {
    if (!b || !ptr) fail(SPH_E_ARG, "NULL argument");
    *ptr = b->d;
  });
}

int sph_buffer_read(const sph_buffer *b, void *host, int64_t nbytes) {
  return guarded([&] {
    if (!b || !host) fail(SPH_E_ARG, "NULL argument");
    if (nbytes < 0 || nbytes > b->nbytes) fail(SPH_E_ARG, "read of %lld bytes from a %lld-byte buffer",
                                                 (long long)nbytes, (long long)b->nbytes);
    hipStream_t s = b->ctx->stream;
    HIPC(hipMemcpyAsync(host, b->d, (size_t)nbytes, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
  });
}

int sph_buffer_write(sph_buffer *b, const void *host, int64_t nbytes) {
  return guarded([&] {
    if (!b || !host) fail(SPH_E_ARG, "NULL argument");
    if (nbytes < 0 || nbytes > b->nbytes) fail(SPH_E_ARG, "write of %lld bytes to a %lld-byte buffer",
                                                 (long long)nbytes, (long long)b->nbytes);
    hipStream_t s = b->ctx->stream;
    HIPC(hipMemcpyAsync(b->d, host, (size_t)nbytes, hipMemcpyHostToDevice, s));
    HIPC(hipStreamSynchronize(s));
  });
}

}  // extern "C"

namespace {

// State(backend, X0, V0, p; ...) for nsc scenes side by side (nsc > 1: the
// f32c form's batched state); X0, V0 hold the scenes' particles in order.
void state_create_impl(sph_ctx *ctx, const sph_params *ps, int64_t nsc, const int64_t *Ns,
                       int precision, const double *X0, const double *V0, double skin,
                       int64_t interval, int64_t maxnb, int layout, sph_state **out) {
  {
    if (!ctx || !out || !ps || !Ns) fail(SPH_E_ARG, "NULL argument");
    if (nsc < 1) fail(SPH_E_ARG, "a state needs at least one scene");
    need_precision(precision);
    if (nsc > 1 && precision != SPH_F32C) fail(SPH_E_ARG, "batched states use the f32c form");
    int64_t N = 0;
    for (int64_t k = 0; k < nsc; ++k) {
      check_params(&ps[k]);
      if (Ns[k] < 1) fail(SPH_E_ARG, "scene %lld: N=%lld out of range", (long long)k,
                          (long long)Ns[k]);
      N += Ns[k];
    }
    const sph_params *p = &ps[0];
    if (N < 1 || N > (int64_t)1 << 30) fail(SPH_E_ARG, "N=%lld out of range", (long long)N);
    if (precision == SPH_F32C && N > kMaxSlotsF32C)
      fail(SPH_E_ARG, "the f32c form addresses its mask lists (48 bytes per particle) with 32-bit "
                      "offsets: N=%lld above %lld", (long long)N, kMaxSlotsF32C);
    if (!X0 || !V0) fail(SPH_E_ARG, "X0 and V0 are required");
    if (!(skin >= 0)) fail(SPH_E_ARG, "skin must be >= 0");
    if (interval < 1) fail(SPH_E_ARG, "interval must be >= 1");
    if (maxnb < 1) fail(SPH_E_ARG, "maxnb must be >= 1");
    if (layout != SPH_LAYOUT_SLOT && layout != SPH_LAYOUT_PARTICLE)
      fail(SPH_E_ARG, "layout must be :slot or :particle");
    HIPC(hipSetDevice(ctx->device));
    auto *st = new sph_state;
    try {
      st->ctx = ctx;
      st->p = *p;
      st->precision = precision;
      st->N = N;
      st->skin = skin;
      st->interval = interval;
      st->maxnb = maxnb;
      st->layout = layout;
      if (layout == SPH_LAYOUT_SLOT) { st->sm = N; st->si = 1; }
      else { st->sm = 1; st->si = maxnb; }
      st->P = make_phys(*p);
      st->P32 = make_phys32(*p, (2.0 * p->h) * (1.0 + skin));
      // NeighborList ctor: rc = Float64(support(p))*(1+skin), dmax = Float64(p.h)*skin;
      // CellList with cs = rc (src/neighbors.jl:28-36,174-190; src/driver.jl:33-34)
      const double rc = (2.0 * p->h) * (1.0 + skin);
      const double dmax = p->h * skin;
      long long nx = (long long)std::ceil(p->Lx / rc) + 2;
      long long ny = (long long)std::ceil(p->Ly / rc) + 2;
      nx = nx < 1 ? 1 : nx;
      ny = ny < 1 ? 1 : ny;
      if (nx * ny > (1LL << 30)) fail(SPH_E_ARG, "cell grid %lldx%lld too large", nx, ny);
      st->g.nx = (int)nx;
      st->g.inv_nx = 1.0 / (double)nx;
      st->g.ny = (int)ny;
      st->g.ncell = (int)(nx * ny);
      st->g.N = (int)N;
      st->kc.cs = rc;
      st->kc.invcs = 1.0 / rc;
      st->kc.rc2 = rc * rc;
      st->kc.r2min = (DBL_EPSILON * p->h) * p->h;
      st->kc.dmax2 = dmax * dmax;
      st->ntnode = p->ngx * p->ngy;
      if (precision == SPH_F32C) {
        std::vector<Scene32> scenes;
        for (int64_t k = 0; k < nsc; ++k) scenes.push_back(make_scene32(ps[k], Ns[k], skin));
        batch_build(st->bs, scenes);
        if (nsc > 1) {
          // the global arrays span every scene; per-scene grids live in the table
          st->g.nx = st->g.ny = 0;
          st->g.inv_nx = 0.0;
          st->g.ncell = st->bs.B.ncell;
          st->ntnode = st->bs.B.ntnode;
        }
      }
      const size_t n = (size_t)N, nc = (size_t)st->g.ncell;
      st->tmp2 = dalloc<double2>(n);
      st->ids = dalloc<int>(n);
      st->scell = dalloc<int>(n);
      st->cstart = dalloc<int>(nc + 1);
      const bool f32c = precision == SPH_F32C;
      if (!f32c) st->bcell = dalloc<int>(n);
      st->X = dalloc<double2>(n);
      st->Xn = dalloc<double2>(n);
      st->Xref = dalloc<double2>(n);
      st->Xref2 = dalloc<double2>(n);
      if (f32c) {
        st->V32 = dalloc<float2>(n);
        st->V32n = dalloc<float2>(n);
        st->xr = dalloc<float2>(n);
        st->xrn = dalloc<float2>(n);
        st->PI32 = dalloc<float2>(n);
      } else {
        st->V = dalloc<double2>(n);
        st->Vn = dalloc<double2>(n);
        st->PI = dalloc<double2>(n);
      }
      if (f32c) {
        st->nm = dalloc<uint32_t>(n * 3 * kMaskWords);
        st->em = dalloc<uint32_t>(n * 3 * kMaskWords);
        st->lane = dalloc<int>(n);
      } else {
        st->nl = dalloc<int>(n * (size_t)maxnb);
        st->nbcnt = dalloc<int>(n);
      }
      st->cellof = dalloc<int>(n);
      st->counts = dalloc<int>(nc);
      st->cl_cstart = dalloc<int>(nc + 1);
      st->cursor = dalloc<int>(nc);
      st->tmp_id = dalloc<int>(n);
      st->tmp_slot = dalloc<int>(n);
      st->perm = dalloc<int>(n);
      st->new_ids = dalloc<int>(n);
      st->bsum = dalloc<int>(nc / 1024 + 2);
      st->bigcells = dalloc<int>(nc + 1);
      st->flags = dalloc<int>(kNumFlags);
      st->a_id = dalloc<double2>(n);
      st->pi_id = dalloc<double2>(n);
      st->a_slot = dalloc<double2>(n);
      st->alpha_id = dalloc<double>(n);
      st->rho_id = dalloc<double>(n);
      st->alpha_slot = dalloc<double>(n);
      st->rho_slot = dalloc<double>(n);
      st->theta_last = dalloc<double>((size_t)st->ntnode);
      st->seed_slot = dalloc<double2>(n);
      st->tmp1 = dalloc<double>(n);
      st->obj_nb = objective_blocks((int)N);
      st->obj_part = dalloc<double>(st->obj_nb + 2);
      st->obj_J = dalloc<double>((size_t)nsc);
      hipStream_t s = ctx->stream;
      // State ctor: X, V copied in; a, rho, pterm, invrho, alpha start at zero
      std::vector<int> iota(n);
      for (size_t i = 0; i < n; ++i) iota[i] = (int)i;
      HIPC(hipMemcpyAsync(st->ids, iota.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
      HIPC(hipMemcpyAsync(st->X, X0, sizeof(double2) * n, hipMemcpyHostToDevice, s));
      if (f32c) {
        HIPC(hipMemcpyAsync(st->tmp2, V0, sizeof(double2) * n, hipMemcpyHostToDevice, s));
        launch_d2_to_f2(s, st->tmp2, (int)N, st->V32);
        HIPC(hipMemsetAsync(st->PI32, 0, sizeof(float2) * n, s));
      } else {
        HIPC(hipMemcpyAsync(st->V, V0, sizeof(double2) * n, hipMemcpyHostToDevice, s));
        HIPC(hipMemsetAsync(st->PI, 0, sizeof(double2) * n, s));
      }
      HIPC(hipMemsetAsync(st->flags, 0, sizeof(int) * kNumFlags, s));
      HIPC(hipMemsetAsync(st->a_id, 0, sizeof(double2) * n, s));
      HIPC(hipMemsetAsync(st->pi_id, 0, sizeof(double2) * n, s));
      HIPC(hipMemsetAsync(st->alpha_id, 0, sizeof(double) * n, s));
      HIPC(hipMemsetAsync(st->rho_id, 0, sizeof(double) * n, s));
      HIPC(hipMemsetAsync(st->counts, 0, sizeof(int) * nc, s));
      HIPC(hipStreamSynchronize(s));
    } catch (...) {
      state_free(st);
      delete st;
      throw;
    }
    ctx_retain(ctx);
    *out = st;
  }
}

}  // namespace

extern "C" {

int sph_state_create(sph_ctx *ctx, const sph_params *p, int precision, int64_t N, const double *X0,
                     const double *V0, double skin, int64_t interval, int64_t maxnb, int layout,
                     sph_state **out) {
  return guarded([&] {
    if (!p) fail(SPH_E_ARG, "NULL argument");
    state_create_impl(ctx, p, 1, &N, precision, X0, V0, skin, interval, maxnb, layout, out);
  });
}

int sph_batch_create(sph_ctx *ctx, int64_t nscenes, const sph_params *params, const int64_t *N,
                     const double *X0, const double *V0, double skin, int64_t interval,
                     int64_t maxnb, sph_state **out) {
  return guarded([&] {
    state_create_impl(ctx, params, nscenes, N, SPH_F32C, X0, V0, skin, interval, maxnb,
                      SPH_LAYOUT_SLOT, out);
  });
}

int sph_state_destroy(sph_state *st) {
  return guarded([&] {
    if (!st) return;
    (void)hipStreamSynchronize(st->ctx->stream);
    state_free(st);
    sph_ctx *c = st->ctx;
    delete st;
    ctx_release(c);
  });
}

int sph_state_info(const sph_state *st, int64_t *o) {
  return guarded([&] {
    if (!st || !o) fail(SPH_E_ARG, "NULL argument");
    o[0] = st->g.nx; o[1] = st->g.ny; o[2] = st->g.ncell; o[3] = st->maxnb;
    o[4] = st->sm; o[5] = st->si; o[6] = st->interval; o[7] = st->age;
  });
}

int sph_state_get(sph_state *st, int field, double *host) {
  return guarded([&] {
    if (!st || !host) fail(SPH_E_ARG, "NULL argument");
    hipStream_t s = st->ctx->stream;
    const int N = (int)st->N;
    const size_t n = (size_t)N;
    std::vector<double2> buf;
    switch (field) {
      case SPH_X:
      case SPH_V:
        if (field == SPH_V && st->precision == SPH_F32C)
          launch_scatter2f(s, st->V32, st->ids, N, st->tmp2);
        else
          launch_scatter2(s, field == SPH_X ? st->X : st->V, st->ids, N, st->tmp2);
        HIPC(hipMemcpyAsync(host, st->tmp2, sizeof(double2) * n, hipMemcpyDeviceToHost, s));
        break;
      case SPH_A:
        materialise_aux(st);
        HIPC(hipMemcpyAsync(host, st->a_id, sizeof(double2) * n, hipMemcpyDeviceToHost, s));
        break;
      case SPH_RHO:
      case SPH_ALPHA:
        materialise_aux(st);
        HIPC(hipMemcpyAsync(host, field == SPH_RHO ? st->rho_id : st->alpha_id, sizeof(double) * n,
                            hipMemcpyDeviceToHost, s));
        break;
      case SPH_PTERM:
      case SPH_INVRHO:
        materialise_aux(st);
        buf.resize(n);
        HIPC(hipMemcpyAsync(buf.data(), st->pi_id, sizeof(double2) * n, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        for (size_t i = 0; i < n; ++i) host[i] = field == SPH_PTERM ? buf[i].x : buf[i].y;
        break;
      default:
        fail(SPH_E_ARG, "unknown field %d", field);
    }
    HIPC(hipStreamSynchronize(s));
  });
}

int sph_state_set(sph_state *st, int field, const double *host) {
  return guarded([&] {
    if (!st || !host) fail(SPH_E_ARG, "NULL argument");
    if (field != SPH_X && field != SPH_V) fail(SPH_E_ARG, "only X and V can be set");
    hipStream_t s = st->ctx->stream;
    materialise_aux(st);  // the spare buffers may be reused below
    HIPC(hipMemcpyAsync(st->tmp2, host, sizeof(double2) * st->N, hipMemcpyHostToDevice, s));
    if (st->precision == SPH_F32C) {
      if (field == SPH_V) {
        launch_gather2f(s, st->tmp2, st->ids, (int)st->N, st->V32);
      } else {
        launch_gather2(s, st->tmp2, st->ids, (int)st->N, st->X);
        if (st->age != kAgeNever) launch_xrel_f32c(s, st->X, st->scell, st->bs.B, st->xr);
      }
    } else {
      launch_gather2(s, st->tmp2, st->ids, (int)st->N, field == SPH_X ? st->X : st->V);
    }
    HIPC(hipStreamSynchronize(s));
  });
}

int sph_build(sph_state *st) {
  return guarded([&] {
    if (!st) fail(SPH_E_ARG, "NULL argument");
    do_build(st);
  });
}

int sph_build_neighbors(sph_state *st) {
  return guarded([&] {
    if (!st) fail(SPH_E_ARG, "NULL argument");
    materialise_aux(st);
    do_build_neighbors(st);
  });
}

int sph_cell_list_get(sph_state *st, int32_t *cellof, int32_t *counts, int32_t *starts,
                      int32_t *order) {
  return guarded([&] {
    if (!st) fail(SPH_E_ARG, "NULL argument");
    if (is_batched(st)) fail(SPH_E_ARG, "cell lists of batched states are not exported");
    if (st->cl_view == 0) fail(SPH_E_STATE, "no cell list has been built");
    hipStream_t s = st->ctx->stream;
    const size_t n = (size_t)st->N, nc = (size_t)st->g.ncell;
    std::vector<int> cnt(nc), stt(nc + 1), ord(n);
    const int *dstart = st->cl_view == 1 ? st->cl_cstart : st->cstart;
    const int *dorder = st->cl_view == 1 ? st->new_ids : st->ids;
    HIPC(hipMemcpyAsync(stt.data(), dstart, sizeof(int) * (nc + 1), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(ord.data(), dorder, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    for (size_t c = 0; c < nc; ++c) cnt[c] = stt[c + 1] - stt[c];
    // a quadrant-ordered layout holds each cell's ids by quadrant: the
    // reference's order is ascending id
    if (st->cl_view == 2 && st->layout_quad)
      for (size_t c = 0; c < nc; ++c) std::sort(ord.begin() + stt[c], ord.begin() + stt[c + 1]);
    if (counts) std::memcpy(counts, cnt.data(), sizeof(int) * nc);
    if (starts) std::memcpy(starts, stt.data(), sizeof(int) * nc);
    if (order)
      for (size_t t = 0; t < n; ++t) order[t] = ord[t] + 1;
    if (cellof)
      for (size_t c = 0; c < nc; ++c)
        for (int t = stt[c]; t < stt[c] + cnt[c]; ++t) cellof[ord[t]] = (int)c + 1;
  });
}

int sph_neighbors_get(sph_state *st, int32_t *indices, int32_t *counts) {
  return guarded([&] {
    if (!st) fail(SPH_E_ARG, "NULL argument");
    if (is_batched(st)) fail(SPH_E_ARG, "neighbour lists of batched states are not exported");
    if (st->age == kAgeNever) fail(SPH_E_STATE, "the neighbour list has not been built");
    hipStream_t s = st->ctx->stream;
    const size_t n = (size_t)st->N;
    if (!st->nl_idx) {
      st->nl_idx = dalloc<int>(n * st->maxnb);
      st->nl_cnt = dalloc<int>(n);
    }
    HIPC(hipMemsetAsync(st->nl_idx, 0, sizeof(int) * n * st->maxnb, s));
    launch_materialise(s, st->Xref, st->ids, st->scell, st->cstart, st->g, st->kc, (int)st->maxnb,
                       st->sm, st->si, st->nl_idx, st->nl_cnt, st->flags);
    std::vector<int> cnt_h;
    if (indices)
      HIPC(hipMemcpyAsync(indices, st->nl_idx, sizeof(int) * n * st->maxnb, hipMemcpyDeviceToHost, s));
    if (counts || (indices && st->layout_quad)) {
      cnt_h.resize(n);
      HIPC(hipMemcpyAsync(cnt_h.data(), st->nl_cnt, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    }
    HIPC(hipStreamSynchronize(s));
    if (counts) std::memcpy(counts, cnt_h.data(), sizeof(int) * n);
    if (indices && st->layout_quad) {
      // a quadrant-ordered layout lists each cell's entries by quadrant; the
      // reference's sweep visits cells in ascending index and ids ascending in
      // each, i.e. the order of (cell, id).  (When a list overflows maxnb the
      // entries kept are the first maxnb of the quadrant order.)
      std::vector<int> sc(n), idv(n), cell_of(n);
      HIPC(hipMemcpy(sc.data(), st->scell, sizeof(int) * n, hipMemcpyDeviceToHost));
      HIPC(hipMemcpy(idv.data(), st->ids, sizeof(int) * n, hipMemcpyDeviceToHost));
      for (size_t t = 0; t < n; ++t) cell_of[idv[t]] = sc[t];
      std::vector<int> e;
      for (size_t i = 0; i < n; ++i) {
        e.resize(cnt_h[i]);
        for (int m = 0; m < cnt_h[i]; ++m) e[m] = indices[m * st->sm + (int64_t)i * st->si];
        std::sort(e.begin(), e.end(), [&](int a, int b) {
          const int ca = cell_of[a - 1], cb = cell_of[b - 1];
          return ca != cb ? ca < cb : a < b;
        });
        for (int m = 0; m < cnt_h[i]; ++m) indices[m * st->sm + (int64_t)i * st->si] = e[m];
      }
    }
  });
}

int sph_flags_get(sph_state *st, int32_t *flags2) {
  return guarded([&] {
    if (!st || !flags2) fail(SPH_E_ARG, "NULL argument");
    HIPC(hipMemcpyAsync(flags2, st->flags, sizeof(int) * 2, hipMemcpyDeviceToHost, st->ctx->stream));
    HIPC(hipStreamSynchronize(st->ctx->stream));
  });
}

int sph_flags_get_all(sph_state *st, int32_t *flags3) {
  return guarded([&] {
    if (!st || !flags3) fail(SPH_E_ARG, "NULL argument");
    HIPC(hipMemcpyAsync(flags3, st->flags, sizeof(int) * kNumFlags, hipMemcpyDeviceToHost,
                        st->ctx->stream));
    HIPC(hipStreamSynchronize(st->ctx->stream));
  });
}

int sph_flags_clear(sph_state *st) {
  return guarded([&] {
    if (!st) fail(SPH_E_ARG, "NULL argument");
    launch_fill_i32(st->ctx->stream, st->flags, kNumFlags, 0);
  });
}

int sph_step(sph_state *st, const sph_field *theta) {
  return guarded([&] {
    if (!st) fail(SPH_E_ARG, "NULL argument");
    copy_theta(st, theta);
    do_step(st, theta->d, nullptr);
  });
}

// simulate! without a tape, the steps replayed from a captured HIP graph.
// Two rebuild epochs (2*interval steps, the first one a rebuild) bring every
// buffer role back to where it started, so one captured cycle replays as is;
// the steps before the first rebuild and after the last whole cycle run
// eagerly.  The graph holds kernel launches only (no memset or copy nodes) and
// is kept with the state until the design field or the buffer roles change.
int sph_simulate_graph(sph_state *st, const sph_field *theta, int64_t nsteps, int64_t *replays) {
  return guarded([&] {
    if (!st) fail(SPH_E_ARG, "NULL argument");
    if (nsteps < 0) fail(SPH_E_ARG, "nsteps must be >= 0");
    copy_theta(st, theta);
    hipStream_t s = st->ctx->stream;
    int64_t nrep = 0;
    auto eager = [&](int64_t n) {
      for (int64_t k = 0; k < n; ++k) do_step(st, theta->d, nullptr);
      nsteps -= n;
    };
    // up to a rebuild that checks displacements (a first build does not)
    if (st->age == kAgeNever) eager(std::min<int64_t>(nsteps, st->interval));
    if (st->age < st->interval) eager(std::min<int64_t>(nsteps, st->interval - st->age));
    const int64_t cyc = 2 * st->interval;
    // every buffer whose role a step or a rebuild swaps
    auto roles = [&]() {
      return std::vector<const void *>{st->X, st->Xn, st->V, st->Vn, st->V32, st->V32n, st->xr,
                                       st->xrn, st->Xref, st->Xref2, st->ids, st->new_ids,
                                       st->cstart, st->cl_cstart};
    };
    if (nsteps >= cyc && !st->ctx->prof) {
      if (!st->gexec || st->gtheta_d != theta->d || st->groles != roles()) {
        if (st->gexec) HIPC(hipGraphExecDestroy(st->gexec));
        st->gexec = nullptr;
        hipGraph_t graph = nullptr;
        const std::vector<const void *> start = roles();
        HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          for (int64_t k = 0; k < cyc; ++k) do_step(st, theta->d, nullptr);
        } catch (...) {
          (void)hipStreamEndCapture(s, &graph);
          if (graph) (void)hipGraphDestroy(graph);
          throw;
        }
        HIPC(hipStreamEndCapture(s, &graph));
        if (roles() != start) {
          (void)hipGraphDestroy(graph);
          fail(SPH_E_STATE, "a captured cycle did not return the buffers to their roles");
        }
        hipError_t e = hipGraphInstantiate(&st->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPC(e);
        st->gtheta_d = theta->d;
        st->groles = start;
        // the capture advanced the host-side bookkeeping by one cycle; the
        // first launch brings the device there
        HIPC(hipGraphLaunch(st->gexec, s));
        nsteps -= cyc;
        nrep += 1;
      }
      while (nsteps >= cyc) {
        HIPC(hipGraphLaunch(st->gexec, s));
        st->epoch += 2;
        nsteps -= cyc;
        nrep += 1;
      }
    }
    eager(nsteps);
    if (replays) *replays = nrep;
    HIPC(hipStreamSynchronize(s));
    prof_collect(st->ctx);
  });
}

int sph_simulate(sph_state *st, const sph_field *theta, int64_t nsteps, sph_tape *tape) {
  return guarded([&] {
    if (!st) fail(SPH_E_ARG, "NULL argument");
    if (nsteps < 0) fail(SPH_E_ARG, "nsteps must be >= 0");
    if (tape) {
      if (tape->N != st->N) fail(SPH_E_ARG, "tape holds N=%lld, state has N=%lld",
                                 (long long)tape->N, (long long)st->N);
      if (tape->precision != st->precision) fail(SPH_E_ARG, "tape and state precision differ");
      if (tape->n + nsteps > tape->cap)
        fail(SPH_E_CAPACITY, "tape capacity %lld steps, %lld needed; allocate a larger Tape or reset!",
             (long long)tape->cap, (long long)(tape->n + nsteps));
    }
    copy_theta(st, theta);
    for (int64_t k = 0; k < nsteps; ++k) do_step(st, theta->d, tape);
    if (tape && nsteps > 0) {
      // entry n+1 holds the final state: the reverse sweep reads V_{n+1} from it
      hipStream_t s = st->ctx->stream;
      const size_t off = (size_t)tape->n * tape->N;
      HIPC(hipMemcpyAsync(tape->X + off, st->X, sizeof(double2) * tape->N, hipMemcpyDeviceToDevice, s));
      if (st->precision == SPH_F32C)
        HIPC(hipMemcpyAsync(tape->V32 + off, st->V32, sizeof(float2) * tape->N,
                            hipMemcpyDeviceToDevice, s));
      else
        HIPC(hipMemcpyAsync(tape->V + off, st->V, sizeof(double2) * tape->N,
                            hipMemcpyDeviceToDevice, s));
      tape->entry_epoch[tape->n + 1] = tape->entry_epoch[tape->n];
    }
    HIPC(hipStreamSynchronize(st->ctx->stream));
    prof_collect(st->ctx);
  });
}

int sph_tape_create_keep(sph_ctx *ctx, int64_t N, int64_t capacity, int precision, int keep,
                         sph_tape **out) {
  return guarded([&] {
    if (!ctx || !out) fail(SPH_E_ARG, "NULL argument");
    need_precision(precision);
    if (N < 1 || capacity < 0) fail(SPH_E_ARG, "bad tape size");
    if (keep & ~(SPH_TAPE_KEEP_PI | SPH_TAPE_KEEP_LISTS)) fail(SPH_E_ARG, "unknown keep flags %d", keep);
    if ((keep & SPH_TAPE_KEEP_LISTS) && precision != SPH_F32C)
      fail(SPH_E_ARG, "the FP64 form keeps no neighbour masks");
    HIPC(hipSetDevice(ctx->device));
    auto *tp = new sph_tape;
    try {
      tp->ctx = ctx;
      tp->N = N;
      tp->cap = capacity;
      tp->precision = precision;
      const size_t tot = (size_t)(capacity + 1) * (size_t)N;
      tp->X = dalloc<double2>(tot);
      if (precision == SPH_F32C) tp->V32 = dalloc<float2>(tot);
      else tp->V = dalloc<double2>(tot);
      if (keep & SPH_TAPE_KEEP_PI) {
        tp->keep_pi = true;
        if (precision == SPH_F32C) tp->PI32 = dalloc<float2>(tot);
        else tp->PI = dalloc<double2>(tot);
      }
      if (keep & SPH_TAPE_KEEP_LISTS)
        tp->EM = dalloc<uint32_t>((size_t)capacity * (size_t)N * (3 * kMaskWords));
      tp->entry_epoch.assign((size_t)capacity + 2, -1);
    } catch (...) {
      dfree(tp->X);
      dfree(tp->V);
      dfree(tp->V32);
      dfree(tp->PI);
      dfree(tp->PI32);
      dfree(tp->EM);
      delete tp;
      throw;
    }
    ctx_retain(ctx);
    *out = tp;
  });
}

int sph_tape_create(sph_ctx *ctx, int64_t N, int64_t capacity, int precision, sph_tape **out) {
  return sph_tape_create_keep(ctx, N, capacity, precision, 0, out);
}

int sph_tape_destroy(sph_tape *tp) {
  return guarded([&] {
    if (!tp) return;
    (void)hipStreamSynchronize(tp->ctx->stream);
    for (auto *v : {&tp->epochs, &tp->pool})
      for (auto &e : *v) {
        dfree(e.ids);
        dfree(e.scell);
        dfree(e.bcell);
        dfree(e.lane);
        dfree(e.cstart);
      }
    dfree(tp->X);
    dfree(tp->V);
    dfree(tp->V32);
    dfree(tp->PI);
    dfree(tp->PI32);
    dfree(tp->EM);
    batch_free(tp->bs);
    sph_ctx *c = tp->ctx;
    delete tp;
    ctx_release(c);
  });
}

int sph_tape_reset(sph_tape *tp) {
  return guarded([&] {
    if (!tp) fail(SPH_E_ARG, "NULL argument");
    for (auto &e : tp->epochs) tp->pool.push_back(e);
    tp->epochs.clear();
    tp->n = 0;
    std::fill(tp->entry_epoch.begin(), tp->entry_epoch.end(), -1);
  });
}

int sph_tape_length(const sph_tape *tp, int64_t *n) {
  return guarded([&] {
    if (!tp || !n) fail(SPH_E_ARG, "NULL argument");
    *n = tp->n;
  });
}

int sph_tape_capacity(const sph_tape *tp, int64_t *cap) {
  return guarded([&] {
    if (!tp || !cap) fail(SPH_E_ARG, "NULL argument");
    *cap = tp->cap;
  });
}

int sph_tape_get(sph_tape *tp, int64_t k, double *X, double *V) {
  return guarded([&] {
    if (!tp) fail(SPH_E_ARG, "NULL argument");
    if (k < 1 || k > tp->n) fail(SPH_E_ARG, "tape entry %lld outside 1..%lld", (long long)k, (long long)tp->n);
    hipStream_t s = tp->ctx->stream;
    const TapeEpoch &e = tp->epochs[tp->entry_epoch[k]];
    const size_t off = (size_t)(k - 1) * tp->N;
    double2 *tmp = dalloc<double2>(tp->N);
    try {
      if (X) {
        launch_scatter2(s, tp->X + off, e.ids, (int)tp->N, tmp);
        HIPC(hipMemcpyAsync(X, tmp, sizeof(double2) * tp->N, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
      }
      if (V) {
        if (tp->precision == SPH_F32C) launch_scatter2f(s, tp->V32 + off, e.ids, (int)tp->N, tmp);
        else launch_scatter2(s, tp->V + off, e.ids, (int)tp->N, tmp);
        HIPC(hipMemcpyAsync(V, tmp, sizeof(double2) * tp->N, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
      }
    } catch (...) {
      dfree(tmp);
      throw;
    }
    dfree(tmp);
  });
}

void ws_free(sph_workspace *ws) {
  double2 **d2[] = {&ws->gX, &ws->gV, &ws->gXs, &ws->abar, &ws->PI, &ws->Vtmp, &ws->t2,
                    &ws->gX_id, &ws->gV_id};
  for (auto pp : d2) dfree(*pp);
  float2 **f2[] = {&ws->gX32, &ws->gV32, &ws->gXs32, &ws->abar32, &ws->PI32, &ws->Vtmp32,
                   &ws->t2f, &ws->xr32, &ws->xr32b};
  for (auto pp : f2) dfree(*pp);
  dfree(ws->galpha); dfree(ws->grho); dfree(ws->gtheta); dfree(ws->inv);
  dfree(ws->galpha32); dfree(ws->grho32); dfree(ws->tpart);
  dfree(ws->nl); dfree(ws->nbcnt); dfree(ws->flags); dfree(ws->ids_cur); dfree(ws->iota);
  dfree(ws->nm);
  dfree(ws->em);
  dfree(ws->lane);
  dfree(ws->tkey);
  dfree(ws->tlev);
  dfree(ws->trange64);
}

}  // extern "C"

namespace {

// AdjointWorkspace for N particles; ntnode, ntcell > 0 size the design-field
// gradient for the scenes of a batched state
void workspace_create_impl(sph_ctx *ctx, int64_t N, const sph_params *p, int precision,
                           int64_t ntnode, int64_t ntcell, sph_workspace **out) {
  {
    if (!ctx || !out) fail(SPH_E_ARG, "NULL argument");
    check_params(p);
    need_precision(precision);
    if (N < 1) fail(SPH_E_ARG, "bad N");
    HIPC(hipSetDevice(ctx->device));
    auto *ws = new sph_workspace;
    try {
      ws->ctx = ctx;
      ws->N = N;
      ws->p = *p;
      ws->precision = precision;
      ws->ntnode = ntnode;
      ws->ntcell = ntcell;
      const size_t n = (size_t)N;
      ws->t2 = dalloc<double2>(n);
      ws->gX_id = dalloc<double2>(n);
      ws->gV_id = dalloc<double2>(n);
      if (ws->ntnode == 0) {
        ws->ntnode = p->ngx * p->ngy;
        ws->ntcell = (p->ngx - 1) * (p->ngy - 1);
      }
      ws->gtheta = dalloc<double>((size_t)ws->ntnode);
      ws->inv = dalloc<int>(n);
      ws->tpart = dalloc<double4>((size_t)ws->ntcell);
      ws->nbcnt = dalloc<int>(n);
      ws->flags = dalloc<int>(2);
      HIPC(hipMemset(ws->flags, 0, sizeof(int) * 2));
      ws->ids_cur = dalloc<int>(n);
      ws->iota = dalloc<int>(n);
      {
        std::vector<int> io(n);
        for (size_t i = 0; i < n; ++i) io[i] = (int)i;
        HIPC(hipMemcpy(ws->iota, io.data(), sizeof(int) * n, hipMemcpyHostToDevice));
      }
      if (precision == SPH_F32C) {
        ws->gX32 = dalloc<float2>(n);
        ws->gV32 = dalloc<float2>(n);
        ws->gXs32 = dalloc<float2>(n);
        ws->abar32 = dalloc<float2>(n);
        ws->PI32 = dalloc<float2>(n);
        ws->Vtmp32 = dalloc<float2>(n);
        ws->t2f = dalloc<float2>(n);
        ws->xr32 = dalloc<float2>(n);
        ws->xr32b = dalloc<float2>(n);
        ws->galpha32 = dalloc<float>(n);
        ws->grho32 = dalloc<float>(n);
      } else {
        ws->gX = dalloc<double2>(n);
        ws->gV = dalloc<double2>(n);
        ws->gXs = dalloc<double2>(n);
        ws->abar = dalloc<double2>(n);
        ws->PI = dalloc<double2>(n);
        ws->Vtmp = dalloc<double2>(n);
        ws->galpha = dalloc<double>(n);
        ws->grho = dalloc<double>(n);
      }
      HIPC(hipMemset(ws->gX_id, 0, sizeof(double2) * n));
      HIPC(hipMemset(ws->gV_id, 0, sizeof(double2) * n));
      HIPC(hipMemset(ws->gtheta, 0, sizeof(double) * ws->ntnode));
      HIPC(hipDeviceSynchronize());
    } catch (...) {
      ws_free(ws);
      delete ws;
      throw;
    }
    ctx_retain(ctx);
    *out = ws;
  }
}

}  // namespace

extern "C" {

int sph_workspace_create(sph_ctx *ctx, int64_t N, const sph_params *p, int precision,
                         sph_workspace **out) {
  return guarded([&] { workspace_create_impl(ctx, N, p, precision, 0, 0, out); });
}

int sph_batch_workspace_create(sph_ctx *ctx, const sph_state *st, sph_workspace **out) {
  return guarded([&] {
    if (!st) fail(SPH_E_ARG, "NULL argument");
    if (st->precision != SPH_F32C) fail(SPH_E_ARG, "batched workspaces use the f32c form");
    workspace_create_impl(ctx, st->N, &st->p, SPH_F32C, st->bs.B.ntnode, st->bs.B.ntcell, out);
  });
}

// per-scene layout of a batched state: for each scene k, out[4k..4k+3] =
// {first particle, particle count, first design node, design nodes}
int sph_batch_info(const sph_state *st, int64_t *nscenes, int64_t *out) {
  return guarded([&] {
    if (!st || !nscenes) fail(SPH_E_ARG, "NULL argument");
    const int nsc = st->precision == SPH_F32C ? st->bs.B.nscenes : 1;
    *nscenes = nsc;
    if (!out) return;
    if (st->precision != SPH_F32C) {
      out[0] = 0; out[1] = st->N; out[2] = 0; out[3] = st->ntnode;
      return;
    }
    for (int k = 0; k < nsc; ++k) {
      const Scene32 &c = st->bs.host[k];
      out[4 * k + 0] = c.off;
      out[4 * k + 1] = c.g.N;
      out[4 * k + 2] = c.toff;
      out[4 * k + 3] = (int64_t)c.T.ngx * c.T.ngy;
    }
  });
}

int sph_workspace_destroy(sph_workspace *ws) {
  return guarded([&] {
    if (!ws) return;
    (void)hipStreamSynchronize(ws->ctx->stream);
    ws_free(ws);
    sph_ctx *c = ws->ctx;
    delete ws;
    ctx_release(c);
  });
}

int sph_workspace_get(sph_workspace *ws, int which, double *host) {
  return guarded([&] {
    if (!ws || !host) fail(SPH_E_ARG, "NULL argument");
    hipStream_t s = ws->ctx->stream;
    switch (which) {
      case SPH_GX:
        HIPC(hipMemcpyAsync(host, ws->gX_id, sizeof(double2) * ws->N, hipMemcpyDeviceToHost, s));
        break;
      case SPH_GV:
        HIPC(hipMemcpyAsync(host, ws->gV_id, sizeof(double2) * ws->N, hipMemcpyDeviceToHost, s));
        break;
      case SPH_GTHETA:
        HIPC(hipMemcpyAsync(host, ws->gtheta, sizeof(double) * ws->ntnode,
                            hipMemcpyDeviceToHost, s));
        break;
      default:
        fail(SPH_E_ARG, "unknown gradient %d", which);
    }
    HIPC(hipStreamSynchronize(s));
  });
}

// backward! (src/driver.jl:141-207).  The reverse sweep reuses the neighbour
// epochs the tape recorded: the reference's own comments note that any
// rebuild schedule gives the same interaction set while displacements stay
// within skin*h (src/neighbors.jl:133-144; src/driver.jl:162-167), and reusing
// the forward's keeps the recomputed densities bit-identical to the forward's.
// skin, interval and layout are validated for interface parity; maxnb sizes
// the rebuilt lists as in the reference.
namespace {

void ensure_ws_list(sph_workspace *ws, int64_t maxnb) {
  if (!ws->tlev) {
    ws->tlev = dalloc<unsigned long long>(1);
    HIPC(hipMemsetAsync(ws->tlev, 0, sizeof(unsigned long long), ws->ctx->stream));
  }
  if (!ws->tkey) ws->tkey = dalloc<int>((size_t)ws->N);   // the gtheta gather's keys
  if (ws->precision == SPH_F32C) {
    if (!ws->nm) ws->nm = dalloc<uint32_t>((size_t)ws->N * 3 * kMaskWords);
    if (!ws->em) ws->em = dalloc<uint32_t>((size_t)ws->N * 3 * kMaskWords);
    if (!ws->lane) ws->lane = dalloc<int>((size_t)ws->N);
    return;
  }
  if (ws->nl_cap < maxnb) {
    dfree(ws->nl);
    ws->nl = dalloc<int>((size_t)ws->N * (size_t)maxnb);
    ws->nl_cap = maxnb;
  }
}

// Reverse sweep over every entry of `tape`.  The adjoint state (gX, gV) enters
// in the slot order whose ids are `ids_in` and leaves in the slot order of the
// tape's first entry, whose ids are copied to ws->ids_cur.  When `lnl` is
// given it is the already built neighbour list of the tape's last epoch.
void bw_run(sph_workspace *ws, sph_tape *tape, const sph_field *theta, const int *ids_in,
            int maxnb, const int *lnl, const int *lcnt, const uint32_t *lnm, const int *llane) {
  hipStream_t s = ws->ctx->stream;
  sph_ctx *c = ws->ctx;
  const int N = (int)ws->N;
  const size_t n = (size_t)N;
  const bool f32c = ws->precision == SPH_F32C;
  const Grid g = tape->g;
  const PhysF64 P = make_phys(ws->p);
  const Batch &B = tape->bs.B;   // f32c form: untuned launches
  const BatchStore &BS = tape->bs;
  const ThetaGrid T = make_theta_grid(ws->p, tape->rc);
  if (!f32c && (ws->tr_nx != g.nx || ws->tr_ny != g.ny || ws->tr_cs != T.cs)) {
    // the FP64 R passes compare each slot's binning cell with these ranges,
    // and the gtheta gather sweeps them: the one-cell ranges of the theta
    // columns, then of the theta rows, then the quarter-cell ranges likewise
    std::vector<int2> tr;
    const int ncx = T.ngx - 1, ncy = T.ngy - 1;
    const double ex = 0.25 * T.cs;
    for (int i = 0; i < ncx; ++i) tr.push_back(theta_cell_range(i, ncx, T.hx, T.cs, g.nx));
    for (int j = 0; j < ncy; ++j) tr.push_back(theta_cell_range(j, ncy, T.hy, T.cs, g.ny));
    for (int i = 0; i < ncx; ++i) tr.push_back(theta_cell_range_tight(i, ncx, T.hx, T.cs, g.nx, ex));
    for (int j = 0; j < ncy; ++j) tr.push_back(theta_cell_range_tight(j, ncy, T.hy, T.cs, g.ny, ex));
    dfree(ws->trange64);
    ws->trange64 = dalloc<int2>(tr.size());
    HIPC(hipMemcpyAsync(ws->trange64, tr.data(), sizeof(int2) * tr.size(), hipMemcpyHostToDevice, s));
    HIPC(hipStreamSynchronize(s));
    ws->tr_nx = g.nx;
    ws->tr_ny = g.ny;
    ws->tr_cs = T.cs;
  }
  int cur = tape->entry_epoch[tape->n];
  // gtheta: the theta-cell partials of an epoch's steps add up in tpart; its
  // node sums go to gtheta when the sweep leaves the epoch
  int tepoch = -1;
  auto theta_nodes = [&] {
    timed(c, K_THETA, [&] {
      if (f32c) launch_theta_nodes_keyed_f32(s, ws->tpart, B, ws->gtheta);
      else launch_theta_nodes_f64(s, ws->tpart, T, ws->gtheta);
    });
  };
  // the adjoint state into the slot order of the final entry; with stored
  // (pterm, invrho) and exact masks (no list to rebuild over the epoch's first
  // positions, which would overwrite the relative coordinates) the final
  // step's pointwise R pass runs in the same launch
  bool r_done = false;   // f32c form: the R pass of step k already ran (in P2 of step k+1)
  int r_stamp = 0;
  timed(c, K_REMAP, [&] {
    launch_inverse(s, ids_in, N, ws->inv);
    if (f32c && tape->keep_pi && tape->EM && tape->n >= 1) {
      const TapeEpoch &E = tape->epochs[cur];
      const size_t off = (size_t)(tape->n - 1) * n;
      r_stamp = ws->tstamp % 2000000000 + 1;   // never the initial 0
      launch_remap_rpw_f32c(s, tape->X + off, E.scell, ws->gX32, ws->gV32,
                            tape->V32 + (size_t)tape->n * n, ws->inv, E.ids, theta->d,
                            BS.at(TK_ADJ_R), ws->t2f, ws->gXs32, ws->xr32, ws->abar32,
                            ws->galpha32, ws->tkey, ws->tlev, r_stamp, false);
      std::swap(ws->gX32, ws->t2f);
      std::swap(ws->gV32, ws->gXs32);
      r_done = true;
    } else if (f32c) {
      launch_gather_remap2f(s, ws->gX32, ws->inv, tape->epochs[cur].ids, N, ws->t2f);
      std::swap(ws->gX32, ws->t2f);
      launch_gather_remap2f(s, ws->gV32, ws->inv, tape->epochs[cur].ids, N, ws->t2f);
      std::swap(ws->gV32, ws->t2f);
    } else {
      launch_gather_remap2(s, ws->gX, ws->inv, tape->epochs[cur].ids, N, ws->t2);
      std::swap(ws->gX, ws->t2);
      launch_gather_remap2(s, ws->gV, ws->inv, tape->epochs[cur].ids, N, ws->t2);
      std::swap(ws->gV, ws->t2);
    }
  });
  // the given list belongs to the tape's last epoch
  const bool use_given = f32c ? lnm != nullptr : lnl != nullptr;
  const int *nl = use_given ? lnl : ws->nl;
  const int *nbcnt = use_given ? lcnt : ws->nbcnt;
  const uint32_t *nm = use_given ? lnm : ws->nm;
  const int *lane = use_given ? llane : ws->lane;
  int listed = use_given ? (int)tape->epochs.size() - 1 : -1;  // epoch whose list `nl` holds
  for (int64_t k = tape->n; k >= 1; --k) {
    const int e = tape->entry_epoch[k];
    const TapeEpoch &E = tape->epochs[e];
    const size_t off = (size_t)(k - 1) * n;
    const double2 *Xk = tape->X + off;
    const bool tfirst = e != tepoch;
    if (tfirst && tepoch >= 0) theta_nodes();
    tepoch = e;
    if (e != listed) {
      // the epoch's list, rebuilt from the positions of its first entry
      timed(c, K_NLBUILD, [&] {
        const double2 *X0 = tape->X + (size_t)(E.first - 1) * n;
        if (f32c) {
          // a tape with each step's exact masks and (pterm, invrho) needs no
          // mask list and takes the epoch's thread order from its record;
          // otherwise the thread order with the relative coordinates of the
          // epoch's first entry, and the mask list over them (without stored
          // (pterm, invrho) the R pass recomputes the density over it)
          if (!(tape->EM && tape->keep_pi)) {
            launch_xrel_f32c(s, X0, E.scell, B, ws->xr32, ws->lane);
            launch_nm_build_f32c(s, ws->xr32, E.scell, E.cstart, BS.at(TK_NLBUILD), ws->nm,
                                 ws->flags);
          }
        } else {
          launch_nl_build_f64(s, X0, E.scell, E.cstart, g, tape->kc, maxnb, ws->nl, ws->nbcnt,
                              ws->flags);
        }
      });
      listed = e;
      nl = ws->nl;
      nbcnt = ws->nbcnt;
      nm = ws->nm;
      lane = f32c && tape->EM && tape->keep_pi ? E.lane : ws->lane;
    }
    if (f32c) {
      const float2 *Vk = tape->V32 + off;
      const float2 *Vnext = tape->V32 + (size_t)k * n;
      const float2 *PIk = ws->PI32;
      bool remap_r = false;   // this step's R pass ran with the remap
      if (e != cur) {
        // entry k+1 was stored in the next epoch's slot order: move the
        // adjoint state and V_{k+1} into this epoch's order (gXs32 is free
        // between steps: P1 writes it before P2 reads it); with stored
        // (pterm, invrho) the step's pointwise R pass runs in the same launch
        // (an epoch's last step never had its R pass fused into a P2)
        remap_r = tape->keep_pi;
        if (remap_r) {
          ws->tstamp = ws->tstamp % 2000000000 + 1;   // never the initial 0
          PIk = tape->PI32 + off;
        }
        timed(c, K_REMAP, [&] {
          launch_inverse(s, tape->epochs[cur].ids, N, ws->inv);
          if (remap_r) {
            launch_remap_rpw_f32c(s, Xk, E.scell, ws->gX32, ws->gV32, Vnext, ws->inv, E.ids,
                                  theta->d, BS.at(TK_ADJ_R), ws->t2f, ws->gXs32, ws->xr32,
                                  ws->abar32, ws->galpha32, ws->tkey, ws->tlev, ws->tstamp,
                                  true);
          } else {
            launch_gather_remap3f(s, ws->gX32, ws->gV32, Vnext, ws->inv, E.ids, N, ws->t2f,
                                  ws->gXs32, ws->Vtmp32);
          }
          std::swap(ws->gX32, ws->t2f);
          std::swap(ws->gV32, ws->gXs32);
        });
        if (!remap_r) Vnext = ws->Vtmp32;
        cur = e;
      }
      if (remap_r) {
        // done above
      } else if (r_done) {
        // this step's R pass ran in the previous step's P2 launch
        ws->tstamp = r_stamp;
        PIk = tape->PI32 + off;
      } else {
        ws->tstamp = ws->tstamp % 2000000000 + 1;   // never the initial 0
        timed(c, K_ADJ_R, [&] {
          if (tape->keep_pi) {
            PIk = tape->PI32 + off;
            launch_adj_rpw_f32c(s, Xk, E.scell, Vnext, theta->d, BS.at(TK_ADJ_R), ws->xr32,
                                ws->gX32, ws->gV32, ws->abar32, ws->galpha32, ws->tkey, ws->tlev,
                                ws->tstamp);
          } else {
            launch_xrel_f32c(s, Xk, E.scell, B, ws->xr32);
            launch_adj_r_f32c(s, Xk, ws->xr32, Vnext, theta->d, E.scell, nm, E.cstart, lane,
                              BS.at(TK_ADJ_R), ws->PI32, ws->gX32, ws->gV32, ws->abar32,
                              ws->galpha32, ws->em, ws->tkey, ws->tlev, ws->tstamp);
          }
        });
      }
      const ThetaKeyed TA{ws->tkey, Xk, ws->galpha32, E.cstart, ws->tlev, ws->tstamp,
                          tfirst ? 1 : 0, ws->tpart};
      // the exact list of these positions: the forward's density pass kept it
      // on the tape, or the density recompute just wrote it; with stored
      // (pterm, invrho) and no stored list, P1 tests the pairs and writes it
      const uint32_t *exact = tape->EM ? tape->EM + off * (3 * kMaskWords) : ws->em;
      const bool p1_exact = !(tape->keep_pi && !tape->EM);
      const uint32_t *p1_nm = p1_exact ? exact : nm;
      uint32_t *p1_em = p1_exact ? nullptr : ws->em;
      const Batch &BP1 = BS.at(TK_ADJ_P1);
      if (BP1.bs == kBlock) {
        // the gtheta gather runs in P1's launch (timed with P1)
        timed(c, K_ADJ_P1, [&] {
          launch_adj_p1_theta_f32c(s, Xk, ws->xr32, Vk, PIk, ws->abar32, ws->galpha32, theta->d,
                                   E.scell, p1_nm, p1_exact, E.cstart, lane, BP1, ws->gXs32,
                                   ws->gV32, ws->grho32, p1_em, TA);
        });
      } else {
        timed(c, K_THETA, [&] { launch_theta_gather_keyed_f32(s, TA, B); });
        timed(c, K_ADJ_P1, [&] {
          launch_adj_p1_f32c(s, Xk, ws->xr32, Vk, PIk, ws->abar32, ws->galpha32, theta->d,
                             E.scell, p1_nm, p1_exact, E.cstart, lane, BP1, ws->gXs32, ws->gV32,
                             ws->grho32, p1_em);
        });
      }
      // P2; when the next step of the sweep (k-1) is in this epoch and its
      // R pass needs no density recompute, each slot runs that R pass on its
      // final gX in the same launch (its xr goes to the spare buffer)
      const bool fuse = tape->keep_pi && k >= 2 && tape->entry_epoch[k - 1] == e;
      RpwArgs prev{};
      if (fuse) {
        r_stamp = ws->tstamp % 2000000000 + 1;
        prev = RpwArgs{Xk - n, Vk, theta->d, ws->xr32b, ws->gV32, ws->abar32,
                       ws->galpha32, ws->tkey, ws->tlev, r_stamp};
      }
      timed(c, K_ADJ_P2, [&] {
        launch_adj_p2_f32c(s, ws->xr32, ws->grho32, ws->gXs32, E.scell, exact, E.cstart, lane,
                           BS.at(TK_ADJ_P2), ws->gX32, fuse ? &prev : nullptr);
      });
      if (fuse) std::swap(ws->xr32, ws->xr32b);
      r_done = fuse;
    } else {
      const double2 *Vk = tape->V + off;
      const double2 *Vnext = tape->V + (size_t)k * n;
      if (e != cur) {
        timed(c, K_REMAP, [&] {
          launch_inverse(s, tape->epochs[cur].ids, N, ws->inv);
          launch_gather_remap2(s, ws->gX, ws->inv, E.ids, N, ws->t2);
          std::swap(ws->gX, ws->t2);
          launch_gather_remap2(s, ws->gV, ws->inv, E.ids, N, ws->t2);
          std::swap(ws->gV, ws->t2);
          launch_gather_remap2(s, Vnext, ws->inv, E.ids, N, ws->Vtmp);
        });
        Vnext = ws->Vtmp;
        cur = e;
      }
      const double2 *PIk = ws->PI;
      ws->tstamp = ws->tstamp % 2000000000 + 1;   // never the initial 0
      const LevelArgs LA{E.bcell, ws->trange64, ws->tlev, ws->tstamp, T.ngx - 1, T.ngy,
                         ws->tkey};
      const EpochCells EC{E.bcell, E.scell, E.cstart};
      timed(c, K_ADJ_R, [&] {
        if (tape->keep_pi) {
          PIk = tape->PI + off;
          launch_adj_rpw_f64(s, Xk, Vnext, theta->d, g, P, ws->gX, ws->gV, ws->abar, ws->galpha,
                             LA);
        } else {
          launch_adj_r_f64(s, Xk, Vnext, theta->d, nl, nbcnt, g, P, EC, ws->PI, ws->gX, ws->gV,
                           ws->abar, ws->galpha, LA);
        }
      });
      timed(c, K_THETA, [&] {
        launch_theta_cells_keyed_f64(s, Xk, ws->galpha, ws->tkey, E.cstart, ws->trange64, g, T,
                                     tfirst, ws->tpart, ws->tlev, ws->tstamp);
      });
      timed(c, K_ADJ_P1, [&] {
        launch_adj_p1_f64(s, Xk, Vk, PIk, ws->abar, ws->galpha, theta->d, nl, nbcnt, g, P, EC,
                          ws->gXs, ws->gV, ws->grho);
      });
      timed(c, K_ADJ_P2, [&] {
        launch_adj_p2_f64(s, Xk, ws->grho, ws->gXs, nl, nbcnt, g, P, EC, ws->gX);
      });
    }
  }
  if (tepoch >= 0) theta_nodes();
  HIPC(hipMemcpyAsync(ws->ids_cur, tape->epochs[cur].ids, sizeof(int) * n,
                      hipMemcpyDeviceToDevice, s));
}

// seeds (particle-id order; host memory, or device memory when `dev`) into
// the adjoint state, in id order
void bw_seed(sph_workspace *ws, const double *seedX, const double *seedV, bool dev) {
  hipStream_t s = ws->ctx->stream;
  const int N = (int)ws->N;
  const size_t n = (size_t)N;
  HIPC(hipMemsetAsync(ws->gtheta, 0, sizeof(double) * ws->ntnode, s));
  if (ws->precision == SPH_F32C && dev) {
    launch_d2_to_f2(s, reinterpret_cast<const double2 *>(seedX), N, ws->gX32);
    launch_d2_to_f2(s, reinterpret_cast<const double2 *>(seedV), N, ws->gV32);
  } else if (ws->precision == SPH_F32C) {
    HIPC(hipMemcpyAsync(ws->t2, seedX, sizeof(double2) * n, hipMemcpyHostToDevice, s));
    launch_d2_to_f2(s, ws->t2, N, ws->gX32);
    HIPC(hipStreamSynchronize(s));
    HIPC(hipMemcpyAsync(ws->t2, seedV, sizeof(double2) * n, hipMemcpyHostToDevice, s));
    launch_d2_to_f2(s, ws->t2, N, ws->gV32);
  } else {
    const hipMemcpyKind kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIPC(hipMemcpyAsync(ws->gX, seedX, sizeof(double2) * n, kind, s));
    HIPC(hipMemcpyAsync(ws->gV, seedV, sizeof(double2) * n, kind, s));
  }
  HIPC(hipMemcpyAsync(ws->ids_cur, ws->iota, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
}

// the adjoint state, whose slot order is ws->ids_cur, into id order (gX_id, gV_id)
void bw_finish(sph_workspace *ws) {
  hipStream_t s = ws->ctx->stream;
  const int N = (int)ws->N;
  if (ws->precision == SPH_F32C) {
    launch_scatter2f(s, ws->gX32, ws->ids_cur, N, ws->gX_id);
    launch_scatter2f(s, ws->gV32, ws->ids_cur, N, ws->gV_id);
  } else {
    launch_scatter2(s, ws->gX, ws->ids_cur, N, ws->gX_id);
    launch_scatter2(s, ws->gV, ws->ids_cur, N, ws->gV_id);
  }
}

void check_bw_args(sph_workspace *ws, const sph_field *theta, const double *seedX,
                   const double *seedV) {
  if (!ws || !theta || !seedX || !seedV) fail(SPH_E_ARG, "NULL argument");
  if (theta->ngy * theta->ngx != ws->ntnode) fail(SPH_E_ARG, "theta shape mismatch");
}

}  // namespace

}  // extern "C"

namespace {

int backward_impl(sph_workspace *ws, sph_tape *tape, const sph_field *theta, const double *seedX,
                  const double *seedV, double skin, int64_t interval, int64_t maxnb, int layout,
                  bool dev) {
  return guarded([&] {
    check_bw_args(ws, theta, seedX, seedV);
    if (!tape) fail(SPH_E_ARG, "NULL argument");
    if (tape->N != ws->N) fail(SPH_E_ARG, "tape and workspace N differ");
    if (tape->precision != ws->precision) fail(SPH_E_ARG, "tape and workspace precision differ");
    if (!(skin >= 0) || interval < 1 || maxnb < 1) fail(SPH_E_ARG, "bad neighbour-list settings");
    if (layout != SPH_LAYOUT_SLOT && layout != SPH_LAYOUT_PARTICLE) fail(SPH_E_ARG, "bad layout");
    hipStream_t s = ws->ctx->stream;
    ensure_ws_list(ws, maxnb);
    bw_seed(ws, seedX, seedV, dev);
    if (tape->n > 0)
      bw_run(ws, tape, theta, ws->ids_cur, (int)maxnb, nullptr, nullptr, nullptr, nullptr);
    bw_finish(ws);
    HIPC(hipStreamSynchronize(s));
    prof_collect(ws->ctx);
  });
}

int objective_impl(sph_state *st, double xc, double yc, double sigma, double *J, double *seed,
                   bool dev) {
  return guarded([&] {
    if (!st || !J) fail(SPH_E_ARG, "NULL argument");
    if (is_batched(st)) fail(SPH_E_ARG, "batched states have one J per scene: sph_batch_objective");
    hipStream_t s = st->ctx->stream;
    const int N = (int)st->N;
    launch_objective_f64(s, st->X, N, xc, yc, sigma, st->obj_part, st->obj_nb, st->obj_J,
                         seed ? st->seed_slot : nullptr);
    HIPC(hipMemcpyAsync(J, st->obj_J, sizeof(double), hipMemcpyDeviceToHost, s));
    if (seed && dev) {
      launch_scatter2(s, st->seed_slot, st->ids, N, reinterpret_cast<double2 *>(seed));
    } else if (seed) {
      launch_scatter2(s, st->seed_slot, st->ids, N, st->tmp2);
      HIPC(hipMemcpyAsync(seed, st->tmp2, sizeof(double2) * N, hipMemcpyDeviceToHost, s));
    }
    HIPC(hipStreamSynchronize(s));
  });
}

}  // namespace

extern "C" {

int sph_backward(sph_workspace *ws, sph_tape *tape, const sph_field *theta, const double *seedX,
                 const double *seedV, double skin, int64_t interval, int64_t maxnb, int layout) {
  return backward_impl(ws, tape, theta, seedX, seedV, skin, interval, maxnb, layout, false);
}

int sph_backward_dev(sph_workspace *ws, sph_tape *tape, const sph_field *theta,
                     const double *seedX, const double *seedV, double skin, int64_t interval,
                     int64_t maxnb, int layout) {
  return backward_impl(ws, tape, theta, seedX, seedV, skin, interval, maxnb, layout, true);
}

int sph_target_objective(sph_state *st, double xc, double yc, double sigma, double *J,
                         double *seed) {
  return objective_impl(st, xc, yc, sigma, J, seed, false);
}

int sph_target_objective_dev(sph_state *st, double xc, double yc, double sigma, double *J,
                             double *seed) {
  return objective_impl(st, xc, yc, sigma, J, seed, true);
}

// sph_target_objective_dev without library memory: the partial sums go to this
// device's scratch array, J and the id-order seed straight to the caller's
// device memory (the partial kernel writes each slot's seed at its id), and
// nothing is waited for.  No allocation, no memset, no host transfer.
int sph_target_objective_noalloc(sph_state *st, double xc, double yc, double sigma, double *J,
                                 double *seed) {
  return guarded([&] {
    if (!st || !J) fail(SPH_E_ARG, "NULL argument");
    if (is_batched(st)) fail(SPH_E_ARG, "batched states have one J per scene: sph_batch_objective");
    const int nb = objective_blocks((int)st->N);
    if (nb > kScratchPartials)
      fail(SPH_E_ARG, "N=%lld needs %d objective partials; the scratch array holds %d "
                      "(N <= %d)", (long long)st->N, nb, kScratchPartials,
           kScratchPartials * kBlock);
    if (!st->ctx->scratch) fail(SPH_E_STATE, "the context has no scratch address");
    launch_objective_f64(st->ctx->stream, st->X, (int)st->N, xc, yc, sigma, st->ctx->scratch, nb,
                         J, reinterpret_cast<double2 *>(seed), seed ? st->ids : nullptr);
  });
}

}  // extern "C"

namespace {

// target_objective per scene of a (batched) f32c state: J[k] in FP64 and the
// seeds of every scene, scene by scene as the single-scene objective forms them
int batch_objective_impl(sph_state *st, double xc, double yc, double sigma, double *J,
                         double *seed, bool dev) {
  return guarded([&] {
    if (!st || !J) fail(SPH_E_ARG, "NULL argument");
    if (st->precision != SPH_F32C) fail(SPH_E_ARG, "batched objectives use the f32c form");
    hipStream_t s = st->ctx->stream;
    const int nsc = st->bs.B.nscenes;
    for (int k = 0; k < nsc; ++k) {
      const Scene32 &c = st->bs.host[k];
      launch_objective_f64(s, st->X + c.off, c.g.N, xc, yc, sigma, st->obj_part,
                           objective_blocks(c.g.N), st->obj_J + k,
                           seed ? st->seed_slot + c.off : nullptr);
    }
    HIPC(hipMemcpyAsync(J, st->obj_J, sizeof(double) * nsc, hipMemcpyDeviceToHost, s));
    if (seed && dev) {
      launch_scatter2(s, st->seed_slot, st->ids, (int)st->N, reinterpret_cast<double2 *>(seed));
    } else if (seed) {
      launch_scatter2(s, st->seed_slot, st->ids, (int)st->N, st->tmp2);
      HIPC(hipMemcpyAsync(seed, st->tmp2, sizeof(double2) * st->N, hipMemcpyDeviceToHost, s));
    }
    HIPC(hipStreamSynchronize(s));
  });
}

}  // namespace

extern "C" {

int sph_batch_objective(sph_state *st, double xc, double yc, double sigma, double *J,
                        double *seed) {
  return batch_objective_impl(st, xc, yc, sigma, J, seed, false);
}

int sph_batch_objective_dev(sph_state *st, double xc, double yc, double sigma, double *J,
                            double *seed) {
  return batch_objective_impl(st, xc, yc, sigma, J, seed, true);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Binomial checkpointing of the time loop.  The reference keeps every step on
// its Tape (src/driver.jl:38-65); here only snapshots at neighbour-list epoch
// starts are kept, at positions chosen by the binomial (revolve) schedule of
// Griewank & Walther, and each epoch is recomputed into a one-epoch tape right
// before its reverse sweep.  Snapshots sit at rebuild steps, where the rebuild
// is a pure function of the positions and ids, so a restored run repeats the
// original forward bit for bit and the gradients equal the full-tape ones.

namespace {

// C(c + t, t), saturating
int64_t beta(int64_t c, int64_t t) {
  long double v = 1.0L;
  for (int64_t k = 1; k <= t; ++k) {
    v = v * (long double)(c + k) / (long double)k;
    if (v > 4.0e18L) return (int64_t)4e18;
  }
  return (int64_t)(v + 0.5L);
}

// Fewest epoch advances that reverse l epochs whose first state is held, with
// c more snapshots free (s = c + 1 in Griewank's count, which includes the
// held one): with t the least repetition number such that beta(s, t) >= l,
// the minimum is t*l - C(s+t, t-1).
int64_t revolve_cost(int64_t l, int64_t c) {
  if (l <= 1) return 0;
  const int64_t s = c + 1;
  int64_t t = 0;
  while (beta(s, t) < l) ++t;
  // C(s+t, t-1) = beta(s+1, t-1)
  return t * l - beta(s + 1, t - 1);
}

// First advance j of an optimal schedule: the largest j with
// j + cost(l-j, c-1) + cost(j, c) equal to cost(l, c).
int64_t revolve_split(int64_t l, int64_t c) {
  const int64_t best = revolve_cost(l, c);
  for (int64_t j = l - 1; j >= 1; --j)
    if (j + revolve_cost(l - j, c - 1) + revolve_cost(j, c) == best) return j;
  return 1;
}

// The schedule as a sequence of operations on a runner.  With `dry` set only
// the counters move (used by sph_ckpt_plan and the CPU tests).
struct Schedule {
  int64_t E = 0, nsnap = 0;
  bool dry = true;
  int64_t advanced = 0, taped = 0, live = 0, peak = 0;
  std::vector<std::pair<int64_t, int64_t>> chain;  // (a, free slots) of the forward's levels
  // hooks (unused when dry)
  std::function<void(int64_t)> restore, store, drop, tape_epoch, adjoint;
  std::function<void(int64_t, int64_t)> advance;

  void op_restore(int64_t a) { if (!dry) restore(a); }
  void op_advance(int64_t a, int64_t b) {
    advanced += b - a;
    if (!dry && b > a) advance(a, b);
  }
  void op_store(int64_t a) {
    live += 1;
    if (live > peak) peak = live;
    if (!dry) store(a);
  }
  void op_drop(int64_t a) {
    live -= 1;
    if (!dry) drop(a);
  }
  void op_taped(int64_t e) {
    taped += 1;
    if (!dry) tape_epoch(e);
  }
  void op_adjoint(int64_t e) { if (!dry) adjoint(e); }

  // forward: the leftmost descent of the recursion, ending with the last
  // epoch taped (its tape serves the first reverse step)
  void forward() {
    chain.clear();
    int64_t a = 0, c = nsnap;
    for (;;) {
      chain.emplace_back(a, c);
      const int64_t l = E - a;
      if (l == 1) {
        op_taped(a);
        return;
      }
      if (c == 0) {
        op_advance(a, E - 1);
        op_taped(E - 1);
        return;
      }
      const int64_t j = revolve_split(l, c);
      op_advance(a, a + j);
      op_store(a + j);
      a += j;
      c -= 1;
    }
  }

  void reverse(int64_t a, int64_t b, int64_t c) {
    const int64_t l = b - a;
    if (l == 1) {
      op_restore(a);
      op_taped(a);
      op_adjoint(a);
      return;
    }
    if (c == 0) {
      for (int64_t e = b - 1; e >= a; --e) {
        op_restore(a);
        op_advance(a, e);
        op_taped(e);
        op_adjoint(e);
      }
      return;
    }
    const int64_t j = revolve_split(l, c);
    op_restore(a);
    op_advance(a, a + j);
    op_store(a + j);
    reverse(a + j, b, c - 1);
    op_drop(a + j);
    reverse(a, a + j, c);
  }

  void backward() {
    for (int64_t k = (int64_t)chain.size() - 1; k >= 0; --k) {
      const int64_t a = chain[k].first, c = chain[k].second;
      if (k == (int64_t)chain.size() - 1) {
        // the deepest level: its last epoch is on the tape already
        op_adjoint(E - 1);
        for (int64_t e = E - 2; e >= a; --e) {
          op_restore(a);
          op_advance(a, e);
          op_taped(e);
          op_adjoint(e);
        }
      } else {
        op_drop(chain[k + 1].first);
        reverse(a, chain[k + 1].first, c);
      }
    }
  }
};

}  // namespace

struct sph_ckpt {
  sph_ctx *ctx = nullptr;
  sph_params p{};
  int precision = SPH_F64;
  int64_t N = 0, nsteps = 0, nsnap = 0, interval = 4, maxnb = 64, E = 0;
  int64_t seg = 4;                      // steps per schedule unit (whole epochs)
  // tape budget: the last `tail` units of the horizon are taped in the
  // forward, together with the schedule's last unit, and never recomputed;
  // the binomial schedule covers the first Es = E - tail units
  int64_t tail = 0, Es = 0;
  double skin = 0.2;
  sph_state *work = nullptr;
  sph_tape *mini = nullptr;
  struct Snap {
    double2 *X = nullptr, *V = nullptr;
    float2 *V32 = nullptr;
    int *ids = nullptr;
    int *scell = nullptr, *cstart = nullptr, *lane = nullptr;   // f32c form: the layout
    bool quad = false;                                           // ... quadrant-ordered
  };
  std::vector<Snap> snaps;              // slot 0: the initial state; 1..nsnap: schedule
  std::vector<int> slot_of;             // epoch -> slot (-1: none)
  std::vector<int> free_slots;
  int64_t at = -1;                      // the work state sits at the start of this epoch
  int64_t on_tape = -1;                 // epoch held by the one-epoch tape
  bool forward_done = false;
  Schedule sched;
  const sph_field *theta = nullptr;
  sph_workspace *ws = nullptr;
  int64_t fwd_advanced = 0, fwd_taped = 0;
  int *iota = nullptr;                  // particle-id order as a slot order
};

namespace {

// steps of schedule unit e (the last one may be short; the schedule's last
// unit runs to the end of the horizon, through the taped tail)
int64_t ck_steps(const sph_ckpt *ck, int64_t e) {
  int64_t rest = ck->nsteps - e * ck->seg;
  if (e == ck->Es - 1) return rest;
  return rest < ck->seg ? rest : ck->seg;
}

// Snapshots sit at the start of schedule units, where the next step rebuilds.
// FP64 form: (X, V, ids) before that rebuild; a restored state rebuilds as
// the original run did.  f32c form: the rebuild is done before the snapshot
// (do_build_neighbors leaves the mask list to the next step's density pass),
// and the snapshot holds the rebuilt layout (X, V, ids, binning cells, cell
// starts, thread order), so that a restored state skips the cell-list build;
// the layout is a pure function of the positions and ids, so either way the
// restored run repeats the original bit for bit.
void ck_copy(sph_ckpt *ck, bool to_snap, int slot) {
  hipStream_t s = ck->ctx->stream;
  sph_state *st = ck->work;
  sph_ckpt::Snap &q = ck->snaps[slot];
  const size_t n = (size_t)ck->N;
  const bool f32c = ck->precision == SPH_F32C;
  auto cp = [&](void *dst, const void *src, size_t bytes) {
    HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
  };
  if (to_snap) {
    if (f32c && !st->pending_list) {
      do_build_neighbors(st, true);
      st->pending_list = true;
    }
    if (f32c) {
      launch_snapshot_f32c(s, true, st->bs.B, st->X, st->V32, st->ids, st->scell, st->lane,
                           st->cstart, st->Xref, st->xr, q.X, q.V32, q.ids, q.scell, q.lane,
                           q.cstart);
      q.quad = st->layout_quad;
    } else {
      cp(q.X, st->X, sizeof(double2) * n);
      cp(q.V, st->V, sizeof(double2) * n);
      cp(q.ids, st->ids, sizeof(int) * n);
    }
  } else {
    if (f32c) {
      // the rebuilt layout of the snapshot: reference positions and relative
      // coordinates follow from X; the next step builds the mask list
      launch_snapshot_f32c(s, false, st->bs.B, st->X, st->V32, st->ids, st->scell, st->lane,
                           st->cstart, st->Xref, st->xr, q.X, q.V32, q.ids, q.scell, q.lane,
                           q.cstart);
      st->layout_quad = q.quad;
      if (q.quad) launch_quad_starts(s, st->X, st->cstart, st->g, st->kc.cs, st->qs);
      st->age = 0;
      st->pending_list = true;
      st->cl_view = 2;
    } else {
      cp(st->X, q.X, sizeof(double2) * n);
      cp(st->V, q.V, sizeof(double2) * n);
      cp(st->ids, q.ids, sizeof(int) * n);
      // the next step rebuilds the cell list and neighbour list from these
      // positions, exactly as the original run did at this epoch start
      st->age = kAgeNever;
    }
    st->aux_state = 0;
    st->epoch += 1;
  }
}

void ck_bind(sph_ckpt *ck) {
  Schedule &S = ck->sched;
  S.dry = false;
  S.restore = [ck](int64_t a) {
    if (ck->at == a) return;
    int slot = ck->slot_of[a];
    if (slot < 0) fail(SPH_E_STATE, "checkpoint schedule lost the snapshot of epoch %lld", (long long)a);
    ck_copy(ck, false, slot);
    ck->at = a;
  };
  S.store = [ck](int64_t a) {
    if (ck->at != a) fail(SPH_E_STATE, "snapshot of epoch %lld requested away from it", (long long)a);
    if (ck->free_slots.empty()) fail(SPH_E_STATE, "checkpoint schedule ran out of snapshots");
    int slot = ck->free_slots.back();
    ck->free_slots.pop_back();
    ck_copy(ck, true, slot);
    ck->slot_of[a] = slot;
  };
  S.drop = [ck](int64_t a) {
    int slot = ck->slot_of[a];
    if (slot > 0) ck->free_slots.push_back(slot);
    ck->slot_of[a] = -1;
  };
  S.advance = [ck](int64_t a, int64_t b) {
    if (ck->at != a) fail(SPH_E_STATE, "advance from epoch %lld, state at %lld", (long long)a,
                          (long long)ck->at);
    for (int64_t e = a; e < b; ++e)
      for (int64_t k = 0; k < ck_steps(ck, e); ++k) do_step(ck->work, ck->theta->d, nullptr);
    ck->at = b;
  };
  S.tape_epoch = [ck](int64_t e) {
    if (ck->at != e) fail(SPH_E_STATE, "taped epoch %lld, state at %lld", (long long)e,
                          (long long)ck->at);
    sph_tape *tp = ck->mini;
    const int64_t ns = ck_steps(ck, e);
    if (ns > tp->cap)
      fail(SPH_E_STATE, "segment of %lld steps on a tape of %lld", (long long)ns,
           (long long)tp->cap);
    for (auto &r : tp->epochs) tp->pool.push_back(r);
    tp->epochs.clear();
    tp->n = 0;
    for (int64_t k = 0; k < ns; ++k) do_step(ck->work, ck->theta->d, tp);
    hipStream_t s = ck->ctx->stream;
    const size_t off = (size_t)tp->n * tp->N;
    HIPC(hipMemcpyAsync(tp->X + off, ck->work->X, sizeof(double2) * tp->N,
                        hipMemcpyDeviceToDevice, s));
    if (ck->precision == SPH_F32C)
      HIPC(hipMemcpyAsync(tp->V32 + off, ck->work->V32, sizeof(float2) * tp->N,
                          hipMemcpyDeviceToDevice, s));
    else
      HIPC(hipMemcpyAsync(tp->V + off, ck->work->V, sizeof(double2) * tp->N,
                          hipMemcpyDeviceToDevice, s));
    tp->entry_epoch[tp->n + 1] = tp->entry_epoch[tp->n];
    ck->at = e + 1;
    ck->on_tape = e;
  };
  S.adjoint = [ck](int64_t e) {
    if (ck->on_tape != e) fail(SPH_E_STATE, "adjoint of epoch %lld without its tape", (long long)e);
    // the work state has not stepped since the epoch was taped: its list is
    // the epoch's list
    bw_run(ck->ws, ck->mini, ck->theta, ck->ws->ids_cur, (int)ck->maxnb, ck->work->nl,
           ck->work->nbcnt, ck->work->nm, ck->work->lane);
  };
}

void ck_free(sph_ckpt *ck) {
  for (auto &q : ck->snaps) {
    dfree(q.X);
    dfree(q.V);
    dfree(q.V32);
    dfree(q.ids);
    dfree(q.scell);
    dfree(q.cstart);
    dfree(q.lane);
  }
  dfree(ck->iota);
  if (ck->work) sph_state_destroy(ck->work);
  if (ck->mini) sph_tape_destroy(ck->mini);
}

}  // namespace

extern "C" {

// the tuned launch configuration of an f32c state (TK_COUNT entries): the
// workgroup size, negated when the kernel reads neighbours without LDS staging
int sph_tuning_get(const sph_state *st, int32_t *sizes) {
  return guarded([&] {
    if (!st || !sizes) fail(SPH_E_ARG, "NULL argument");
    if (st->precision != SPH_F32C) fail(SPH_E_ARG, "the f32c form has tuned kernels");
    for (int k = 0; k < TK_COUNT; ++k) {
      const int v = st->bs.cfg[k];
      sizes[k] = kSlotBlocks[v % 3] * (v >= 3 ? -1 : 1);
    }
  });
}

int sph_tuning_set(sph_state *st, const int32_t *sizes) {
  return guarded([&] {
    if (!st || !sizes) fail(SPH_E_ARG, "NULL argument");
    if (st->precision != SPH_F32C) fail(SPH_E_ARG, "the f32c form has tuned kernels");
    int v[TK_COUNT];
    for (int k = 0; k < TK_COUNT; ++k) {
      v[k] = -1;
      for (int j = 0; j < 3; ++j) {
        if (sizes[k] == kSlotBlocks[j]) v[k] = j;
        if (sizes[k] == -kSlotBlocks[j]) v[k] = j + 3;
      }
      if (v[k] < 0)
        fail(SPH_E_ARG, "launch configuration %d is not +-64, +-128 or +-256", (int)sizes[k]);
    }
    for (int k = 0; k < TK_COUNT; ++k) st->bs.cfg[k] = v[k];
  });
}

int sph_state_set_sweep(sph_state *st, int sweep) {
  return guarded([&] {
    if (!st) fail(SPH_E_ARG, "NULL argument");
    if (sweep != 3 && sweep != 5) fail(SPH_E_ARG, "sweep must be 3 or 5, not %d", sweep);
    if (sweep == 5) {
      if (st->precision != SPH_F32C) fail(SPH_E_ARG, "the 5x5 half-cell sweep is an f32c build");
      if (is_batched(st)) fail(SPH_E_ARG, "the 5x5 half-cell sweep takes one scene");
      if (st->N >= (1LL << 29)) fail(SPH_E_ARG, "the 5x5 half-cell sweep takes N < 2^29");
      if (!st->qs) st->qs = dalloc<int>(4 * (size_t)st->g.ncell + 1);
    }
    st->sweep = sweep;
  });
}

// Launch configuration of each tuned f32c kernel: for every candidate (three
// workgroup sizes, each with and without LDS staging of the neighbour rows), a
// warm-up pass and a timed pass of `steps` steps with a tape of the kind the
// checkpointed and batched paths record (keeping (pterm, invrho) and the exact
// masks) plus the reverse sweep over them (zero seeds: the same instruction
// stream), every kernel timed with events; each kernel keeps its fastest
// configuration.  Advances st by 12*steps steps.  chosen (may be NULL)
// receives the TK_COUNT configurations as sph_tuning_get reports them.
int sph_autotune(sph_state *st, const sph_field *theta, int64_t steps, int32_t *chosen) {
  sph_tape *tape = nullptr;
  sph_workspace *ws = nullptr;
  int rc = guarded([&] {
    if (!st || !theta) fail(SPH_E_ARG, "NULL argument");
    if (st->precision != SPH_F32C) fail(SPH_E_ARG, "the autotuner tunes the f32c kernels");
    if (steps < 1) fail(SPH_E_ARG, "steps must be >= 1");
    sph_ctx *c = st->ctx;
    if (c->prof) fail(SPH_E_STATE, "turn profiling off before tuning");
    int r = sph_tape_create_keep(c, st->N, steps, SPH_F32C,
                                 SPH_TAPE_KEEP_PI | SPH_TAPE_KEEP_LISTS, &tape);
    if (r != SPH_OK) throw Error(r, g_err);
    workspace_create_impl(c, st->N, &st->p, SPH_F32C, st->ntnode,
                          st->bs.B.nscenes > 1 ? st->bs.B.ntcell
                                               : (st->p.ngx - 1) * (st->p.ngy - 1),
                          &ws);
    std::vector<double> zeros((size_t)st->N * 2, 0.0);
    const int kid[TK_COUNT] = {K_NLBUILD, K_DENSITY, K_FORCE, K_ADJ_R, K_ADJ_P1, K_ADJ_P2};
    double best[TK_COUNT];
    int pick[TK_COUNT];
    for (int k = 0; k < TK_COUNT; ++k) {
      best[k] = 1e300;
      pick[k] = 2;
    }
    for (int v = 0; v < 6; ++v) {
      for (int k = 0; k < TK_COUNT; ++k) st->bs.cfg[k] = v;
      for (int pass = 0; pass < 2; ++pass) {
        HIPC(hipStreamSynchronize(c->stream));
        prof_collect(c);
        for (int k = 0; k < K_COUNT; ++k) {
          c->tot_ms[k] = 0.0;
          c->launches[k] = 0;
        }
        c->prof = pass == 1;
        // the reverse sweep launches with the tape's copy of the sizes
        if (tape->has_geom) batch_copy(tape->bs, st->bs);
        r = sph_tape_reset(tape);
        if (r == SPH_OK) r = sph_simulate(st, theta, steps, tape);
        if (r == SPH_OK)
          r = sph_backward(ws, tape, theta, zeros.data(), zeros.data(), st->skin, st->interval,
                           st->maxnb, SPH_LAYOUT_SLOT);
        c->prof = false;
        if (r != SPH_OK) throw Error(r, g_err);
      }
      for (int k = 0; k < TK_COUNT; ++k) {
        const int64_t n = c->launches[kid[k]];
        // (on this tape the R pass runs inside P2 and the epoch remaps, timed
        // with them: adjoint_r, whose own launch serves other tapes, keeps its
        // default size)
        if (n == 0) continue;
        // P1 at 256 threads carries the gtheta gather in its launch: compare
        // P1 + gather at every size
        const double t =
            (c->tot_ms[kid[k]] + (k == TK_ADJ_P1 ? c->tot_ms[K_THETA] : 0.0)) / (double)n;
        if (t < best[k]) {
          best[k] = t;
          pick[k] = v;
        }
      }
    }
    for (int k = 0; k < TK_COUNT; ++k) {
      st->bs.cfg[k] = pick[k];
      if (chosen) chosen[k] = kSlotBlocks[pick[k] % 3] * (pick[k] >= 3 ? -1 : 1);
      c->tot_ms[kid[k]] = 0.0;
      c->launches[kid[k]] = 0;
    }
  });
  if (ws) sph_workspace_destroy(ws);
  if (tape) sph_tape_destroy(tape);
  return rc;
}

int sph_ckpt_plan(int64_t epochs, int64_t nsnap, int64_t *out3) {
  return guarded([&] {
    if (epochs < 1 || nsnap < 0 || !out3) fail(SPH_E_ARG, "bad plan arguments");
    Schedule S;
    S.E = epochs;
    S.nsnap = nsnap;
    S.forward();
    S.backward();
    out3[0] = S.advanced + S.taped;   // forward epochs run in total
    out3[1] = S.taped;                // epochs run with a tape
    out3[2] = S.peak;                 // most snapshots alive at once (excluding the initial)
  });
}

int sph_ckpt_create(sph_ctx *ctx, const sph_params *p, int precision, int64_t N, int64_t nsteps,
                    int64_t nsnap, int64_t seg_epochs, double skin, int64_t interval,
                    int64_t maxnb, sph_ckpt **out) {
  return guarded([&] {
    if (!ctx || !out) fail(SPH_E_ARG, "NULL argument");
    check_params(p);
    need_precision(precision);
    if (nsteps < 1) fail(SPH_E_ARG, "nsteps must be >= 1");
    if (nsnap < 0) fail(SPH_E_ARG, "nsnap must be >= 0");
    if (interval < 1) fail(SPH_E_ARG, "interval must be >= 1");
    if (seg_epochs < 1) fail(SPH_E_ARG, "seg_epochs must be >= 1");
    auto *ck = new sph_ckpt;
    try {
      ck->ctx = ctx;
      ck->p = *p;
      ck->precision = precision;
      ck->N = N;
      ck->nsteps = nsteps;
      ck->interval = interval;
      ck->maxnb = maxnb;
      ck->skin = skin;
      ck->seg = seg_epochs * interval;
      ck->E = (nsteps + ck->seg - 1) / ck->seg;
      ck->Es = ck->E;
      ck->nsnap = nsnap < ck->E ? nsnap : ck->E;
      std::vector<double> zeros((size_t)N * 2, 0.0);
      int rc = sph_state_create(ctx, p, precision, N, zeros.data(), zeros.data(), skin, interval,
                                maxnb, SPH_LAYOUT_SLOT, &ck->work);
      if (rc != SPH_OK) throw Error(rc, g_err);
      const int64_t cap = ck->seg < nsteps ? ck->seg : nsteps;
      // the segment tape keeps (pterm, invrho) and, in the f32c form, the
      // exact masks of every step
      const int keep = SPH_TAPE_KEEP_PI | (precision == SPH_F32C ? SPH_TAPE_KEEP_LISTS : 0);
      rc = sph_tape_create_keep(ctx, N, cap, precision, keep, &ck->mini);
      if (rc != SPH_OK) throw Error(rc, g_err);
      ck->snaps.resize((size_t)ck->nsnap + 1);
      for (auto &q : ck->snaps) {
        q.X = dalloc<double2>((size_t)N);
        if (precision == SPH_F32C) q.V32 = dalloc<float2>((size_t)N);
        else q.V = dalloc<double2>((size_t)N);
        q.ids = dalloc<int>((size_t)N);
        if (precision == SPH_F32C) {
          q.scell = dalloc<int>((size_t)N);
          q.cstart = dalloc<int>((size_t)ck->work->g.ncell + 1);
          q.lane = dalloc<int>((size_t)N);
        }
      }
      ck->sched.E = ck->E;
      ck->sched.nsnap = ck->nsnap;
      ck_bind(ck);
    } catch (...) {
      ck_free(ck);
      delete ck;
      throw;
    }
    ctx_retain(ctx);
    *out = ck;
  });
}

int sph_ckpt_set_tape_steps(sph_ckpt *ck, int64_t tape_steps) {
  return guarded([&] {
    if (!ck) fail(SPH_E_ARG, "NULL argument");
    if (tape_steps < 1) fail(SPH_E_ARG, "tape_steps must be >= 1");
    // units the tape holds: the schedule's last unit and tail units after it
    int64_t units = tape_steps / ck->seg;
    units = units < 1 ? 1 : (units > ck->E ? ck->E : units);
    const int64_t tail = units - 1;
    const int64_t Es = ck->E - tail;
    // the tape holds any one unit, and the schedule's last unit with the tail
    const int64_t one = ck->seg < ck->nsteps ? ck->seg : ck->nsteps;
    const int64_t last = ck->nsteps - (Es - 1) * ck->seg;
    const int64_t cap = last > one ? last : one;
    sph_tape *tp = nullptr;
    const int keep = SPH_TAPE_KEEP_PI | (ck->precision == SPH_F32C ? SPH_TAPE_KEEP_LISTS : 0);
    (void)hipStreamSynchronize(ck->ctx->stream);
    int rc = sph_tape_create_keep(ck->ctx, ck->N, cap, ck->precision, keep, &tp);
    if (rc != SPH_OK) throw Error(rc, g_err);
    if (ck->mini) sph_tape_destroy(ck->mini);
    ck->mini = tp;
    ck->tail = tail;
    ck->Es = Es;
    ck->sched.E = Es;
    ck->sched.nsnap = ck->nsnap < Es ? ck->nsnap : Es;
    ck->forward_done = false;   // a forward under the old budget has no backward
  });
}

int sph_ckpt_destroy(sph_ckpt *ck) {
  return guarded([&] {
    if (!ck) return;
    (void)hipStreamSynchronize(ck->ctx->stream);
    ck_free(ck);
    sph_ctx *c = ck->ctx;
    delete ck;
    ctx_release(c);
  });
}

}  // extern "C"

namespace {

// forward over nsteps from (X0, V0) (id order; host memory, or device memory
// when `dev`); snapshots per the schedule
int ckpt_forward_impl(sph_ckpt *ck, const double *X0, const double *V0, const sph_field *theta,
                      bool dev) {
  return guarded([&] {
    if (!ck || !X0 || !V0 || !theta) fail(SPH_E_ARG, "NULL argument");
    if (theta->ngy != ck->p.ngy || theta->ngx != ck->p.ngx) fail(SPH_E_ARG, "theta shape mismatch");
    sph_state *st = ck->work;
    hipStream_t s = ck->ctx->stream;
    const size_t n = (size_t)ck->N;
    // reset the work state to (X0, V0) in particle-id order
    if (!ck->iota) {
      std::vector<int> io(n);
      for (size_t i = 0; i < n; ++i) io[i] = (int)i;
      ck->iota = dalloc<int>(n);
      HIPC(hipMemcpyAsync(ck->iota, io.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
      HIPC(hipStreamSynchronize(s));
    }
    HIPC(hipMemcpyAsync(st->ids, ck->iota, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    const hipMemcpyKind kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIPC(hipMemcpyAsync(st->X, X0, sizeof(double2) * n, kind, s));
    if (ck->precision == SPH_F32C) {
      const double2 *v0 = reinterpret_cast<const double2 *>(V0);
      if (!dev) {
        HIPC(hipMemcpyAsync(st->tmp2, V0, sizeof(double2) * n, kind, s));
        v0 = st->tmp2;
      }
      launch_d2_to_f2(s, v0, (int)n, st->V32);
    } else {
      HIPC(hipMemcpyAsync(st->V, V0, sizeof(double2) * n, kind, s));
    }
    if (!dev) HIPC(hipStreamSynchronize(s));  // the host buffers are released on return
    st->age = kAgeNever;
    st->pending_list = false;
    st->aux_state = 0;
    st->epoch += 1;
    copy_theta(st, theta);
    ck->theta = theta;
    ck->slot_of.assign((size_t)ck->E + 1, -1);
    ck->free_slots.clear();
    for (int k = (int)ck->sched.nsnap; k >= 1; --k) ck->free_slots.push_back(k);
    ck->at = 0;
    ck->on_tape = -1;
    ck_copy(ck, true, 0);
    ck->slot_of[0] = 0;
    ck->sched.advanced = ck->sched.taped = ck->sched.live = ck->sched.peak = 0;
    ck->sched.forward();
    ck->fwd_advanced = ck->sched.advanced;
    ck->fwd_taped = ck->sched.taped;
    ck->forward_done = true;
    HIPC(hipStreamSynchronize(s));
    prof_collect(ck->ctx);
  });
}

int ckpt_backward_impl(sph_ckpt *ck, sph_workspace *ws, const sph_field *theta,
                       const double *seedX, const double *seedV, bool dev) {
  return guarded([&] {
    if (!ck) fail(SPH_E_ARG, "NULL argument");
    check_bw_args(ws, theta, seedX, seedV);
    if (!ck->forward_done) fail(SPH_E_STATE, "sph_ckpt_backward needs a preceding sph_ckpt_forward");
    if (ws->N != ck->N || ws->precision != ck->precision)
      fail(SPH_E_ARG, "workspace does not match the checkpointer");
    if (theta != ck->theta) copy_theta(ck->work, theta);
    ck->theta = theta;
    ck->ws = ws;
    ensure_ws_list(ws, ck->maxnb);
    bw_seed(ws, seedX, seedV, dev);
    ck->sched.advanced = ck->sched.taped = 0;
    ck->sched.live = (int64_t)ck->sched.chain.size() - 1;
    ck->sched.backward();
    bw_finish(ws);
    ck->forward_done = false;
    HIPC(hipStreamSynchronize(ws->ctx->stream));
    prof_collect(ws->ctx);
  });
}

}  // namespace

extern "C" {

int sph_ckpt_forward(sph_ckpt *ck, const double *X0, const double *V0, const sph_field *theta) {
  return ckpt_forward_impl(ck, X0, V0, theta, false);
}

int sph_ckpt_forward_dev(sph_ckpt *ck, const double *X0, const double *V0,
                         const sph_field *theta) {
  return ckpt_forward_impl(ck, X0, V0, theta, true);
}

int sph_ckpt_backward(sph_ckpt *ck, sph_workspace *ws, const sph_field *theta, const double *seedX,
                      const double *seedV) {
  return ckpt_backward_impl(ck, ws, theta, seedX, seedV, false);
}

int sph_ckpt_backward_dev(sph_ckpt *ck, sph_workspace *ws, const sph_field *theta,
                          const double *seedX, const double *seedV) {
  return ckpt_backward_impl(ck, ws, theta, seedX, seedV, true);
}

int sph_ckpt_state(sph_ckpt *ck, sph_state **st) {
  return guarded([&] {
    if (!ck || !st) fail(SPH_E_ARG, "NULL argument");
    *st = ck->work;
  });
}

// {epochs, snapshots, forward epochs of the last forward, forward epochs of the
//  last backward (recomputation), epochs taped in the last backward}
int sph_ckpt_stats(const sph_ckpt *ck, int64_t *out5) {
  return guarded([&] {
    if (!ck || !out5) fail(SPH_E_ARG, "NULL argument");
    out5[0] = ck->E;
    out5[1] = ck->nsnap;
    out5[2] = ck->fwd_advanced + ck->fwd_taped + ck->tail;
    out5[3] = ck->sched.advanced + ck->sched.taped;
    out5[4] = ck->sched.taped;
  });
}

}  // extern "C"
