/*
 * sph_oracle_t.h -- type-generic body of the CPU oracle (TEST INFRASTRUCTURE).
 * Included by sph_oracle.c once with REAL=double/SFX=_f64 and once with
 * REAL=float/SFX=_f32.  Citations are to the reference's files.
 */
#define CAT2_(a, b) a##b
#define CAT_(a, b) CAT2_(a, b)
#define F_(name) CAT_(name, SFX)

typedef struct {
  REAL h, m, rho0, c, mu, gx, gy, dt, Lx, Ly, kw;
  int64_t ngx, ngy;
} F_(prm);

/* SPHParams{T}(...) converts every keyword to T (src/params.jl:14-28). */
static F_(prm) F_(prm_from)(const orc_params *q) {
  F_(prm) p;
  p.h = (REAL)q->h; p.m = (REAL)q->m; p.rho0 = (REAL)q->rho0; p.c = (REAL)q->c;
  p.mu = (REAL)q->mu; p.gx = (REAL)q->gx; p.gy = (REAL)q->gy; p.dt = (REAL)q->dt;
  p.Lx = (REAL)q->Lx; p.Ly = (REAL)q->Ly; p.kw = (REAL)q->kw;
  p.ngx = q->ngx; p.ngy = q->ngy;
  return p;
}

/* ---- src/kernels.jl ---------------------------------------------------- */

/* wnorm(h) = T(7) / (T(4) * T(pi) * h^2)   (src/kernels.jl:15) */
static inline REAL F_(wnorm_)(REAL h) {
  return (REAL)7 / (((REAL)4 * (REAL)ORC_PI) * (h * h));
}
/* w_kern (src/kernels.jl:17-26) */
static inline REAL F_(w_kern_)(REAL q, REAL A) {
  if (q < (REAL)2) {
    REAL u = (REAL)1 - q / (REAL)2;
    REAL u2 = u * u;
    return ((A * u2) * u2) * ((REAL)2 * q + (REAL)1);
  }
  return (REAL)0;
}
/* f_kern (src/kernels.jl:28-35): -5*A*u^3/h^2, with u^3 = u*u*u and h^2 = h*h */
static inline REAL F_(f_kern_)(REAL q, REAL A, REAL h) {
  if (q < (REAL)2) {
    REAL u = (REAL)1 - q / (REAL)2;
    return (((REAL)-5 * A) * ((u * u) * u)) / (h * h);
  }
  return (REAL)0;
}
/* g_kern (src/kernels.jl:37-44): 7.5*A*u^2/h^2 */
static inline REAL F_(g_kern_)(REAL q, REAL A, REAL h) {
  if (q < (REAL)2) {
    REAL u = (REAL)1 - q / (REAL)2;
    return (((REAL)7.5 * A) * (u * u)) / (h * h);
  }
  return (REAL)0;
}

/* ---- src/params.jl ----------------------------------------------------- */

static inline REAL F_(maxr)(REAL a, REAL b) { return (b > a) ? b : a; }

/* wall_accel (src/params.jl:36-40) */
static inline REAL F_(wall_accel)(REAL x, REAL L, REAL kw) {
  REAL lo = F_(maxr)((REAL)0, -x);
  REAL hi = F_(maxr)((REAL)0, x - L);
  return kw * (lo * lo - hi * hi);
}
/* wall_daccel (src/params.jl:42-46) */
static inline REAL F_(wall_daccel)(REAL x, REAL L, REAL kw) {
  REAL lo = F_(maxr)((REAL)0, -x);
  REAL hi = F_(maxr)((REAL)0, x - L);
  return kw * ((REAL)-2 * lo - (REAL)2 * hi);
}

/* Julia clamp(x, lo, hi) = ifelse(x > hi, hi, ifelse(x < lo, lo, x)) */
static inline REAL F_(clampr)(REAL x, REAL lo, REAL hi) {
  return (x > hi) ? hi : ((x < lo) ? lo : x);
}
static inline int64_t clampi64(int64_t x, int64_t lo, int64_t hi) {
  return (x > hi) ? hi : ((x < lo) ? lo : x);
}
/* unsafe_trunc(Int, v) on x86-64 yields typemin for NaN/out-of-range input;
 * every call site clamps the result to a small range afterwards, so any value
 * below the clamp floor reproduces it. */
static inline int64_t F_(trunc_i64)(REAL v) {
  if (!(v > (REAL)-1e15 && v < (REAL)1e15)) return INT64_MIN / 2;
  return (int64_t)v;
}
/* unsafe_trunc(Int32, v): x86-64 yields typemin(Int32) for NaN/out-of-range */
static inline int64_t F_(trunc_i32)(REAL v) {
  if (!(v > (REAL)-2147483648.0 && v < (REAL)2147483647.0)) return (int64_t)INT32_MIN;
  return (int64_t)(int32_t)v;
}

/* interp_cell (src/params.jl:53-68); returns 1-based i, j */
static inline void F_(interp_cell)(REAL x, REAL y, const F_(prm) *p, int64_t *ii, int64_t *jj,
                                   REAL *fx, REAL *fy, REAL *ax, REAL *ay) {
  REAL hx = p->Lx / (REAL)(p->ngx - 1);
  REAL hy = p->Ly / (REAL)(p->ngy - 1);
  REAL gx = x / hx;
  REAL gy = y / hy;
  *ax = ((gx > (REAL)0) & (gx < (REAL)(p->ngx - 1))) ? (REAL)1 : (REAL)0;
  *ay = ((gy > (REAL)0) & (gy < (REAL)(p->ngy - 1))) ? (REAL)1 : (REAL)0;
  REAL gxc = F_(clampr)(gx, (REAL)0, (REAL)(p->ngx - 1) - (REAL)1e-6);
  REAL gyc = F_(clampr)(gy, (REAL)0, (REAL)(p->ngy - 1) - (REAL)1e-6);
  int64_t i0 = clampi64(F_(trunc_i64)(FLOORR(gxc)), 0, p->ngx - 2);
  int64_t j0 = clampi64(F_(trunc_i64)(FLOORR(gyc)), 0, p->ngy - 2);
  *fx = gxc - (REAL)i0;
  *fy = gyc - (REAL)j0;
  *ii = i0 + 1;
  *jj = j0 + 1;
}

/* interp_alpha (src/params.jl:70-82); theta is (ngy,ngx) column-major */
static inline void F_(interp_alpha)(const REAL *theta, REAL x, REAL y, const F_(prm) *p,
                                    REAL *a, REAL *dadx, REAL *dady) {
  int64_t i, j;
  REAL fx, fy, ax, ay;
  F_(interp_cell)(x, y, p, &i, &j, &fx, &fy, &ax, &ay);
  const int64_t ny = p->ngy;
  REAL t00 = theta[(j - 1) + (i - 1) * ny];
  REAL t01 = theta[(j - 1) + i * ny];
  REAL t10 = theta[j + (i - 1) * ny];
  REAL t11 = theta[j + i * ny];
  REAL omx = (REAL)1 - fx, omy = (REAL)1 - fy;
  *a = (((omx * omy) * t00 + (fx * omy) * t01) + (omx * fy) * t10) + (fx * fy) * t11;
  REAL hx = p->Lx / (REAL)(p->ngx - 1);
  REAL hy = p->Ly / (REAL)(p->ngy - 1);
  *dadx = ((omy * (t01 - t00) + fy * (t11 - t10)) / hx) * ax;
  *dady = ((omx * (t10 - t00) + fx * (t11 - t01)) / hy) * ay;
}

/* ---- State / CellList / NeighborList (src/driver.jl:5-36, src/neighbors.jl) */

struct F_(orc_state) {
  F_(prm) p;
  int64_t N;
  REAL *X, *V, *a, *rho, *pterm, *invrho, *alpha;
  /* CellList (src/neighbors.jl:9-19) */
  int32_t *cellof, *counts, *cum, *starts, *cursor, *order;
  int64_t nx, ny, ncell;
  double cs;
  /* NeighborList (src/neighbors.jl:150-162) */
  int32_t *indices, *nbcount;
  int32_t flags[2];
  REAL *Xref;
  int64_t maxnb, sm, si, interval, age;
  double rc, dmax;
};

#define AGE_NEVER INT64_MAX

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

F_(orc_state) *F_(orc_state_create)(const orc_params *q, int64_t N, const REAL *X0,
                                    const REAL *V0, double skin, int64_t interval,
                                    int64_t maxnb, int layout) {
  F_(orc_state) *st = (F_(orc_state) *)xcalloc(1, sizeof(*st));
  st->p = F_(prm_from)(q);
  st->N = N;
  st->X = (REAL *)xcalloc(2 * N, sizeof(REAL));
  st->V = (REAL *)xcalloc(2 * N, sizeof(REAL));
  st->a = (REAL *)xcalloc(2 * N, sizeof(REAL));
  st->rho = (REAL *)xcalloc(N, sizeof(REAL));
  st->pterm = (REAL *)xcalloc(N, sizeof(REAL));
  st->invrho = (REAL *)xcalloc(N, sizeof(REAL));
  st->alpha = (REAL *)xcalloc(N, sizeof(REAL));
  if (X0) memcpy(st->X, X0, sizeof(REAL) * 2 * N);
  if (V0) memcpy(st->V, V0, sizeof(REAL) * 2 * N);
  /* NeighborList ctor (src/neighbors.jl:174-190):
   *   rc = Float64(support(p)) * (1 + skin), dmax = Float64(p.h) * skin */
  REAL support = (REAL)2 * st->p.h;
  st->rc = (double)support * (1.0 + skin);
  st->dmax = (double)st->p.h * skin;
  st->maxnb = maxnb;
  st->interval = interval;
  st->age = AGE_NEVER;
  if (layout == ORC_LAYOUT_SLOT) { st->sm = N; st->si = 1; }
  else { st->sm = 1; st->si = maxnb; }
  st->indices = (int32_t *)xcalloc(N * maxnb, sizeof(int32_t));
  st->nbcount = (int32_t *)xcalloc(N, sizeof(int32_t));
  st->Xref = (REAL *)xcalloc(2 * N, sizeof(REAL));
  /* CellList ctor with cs = nl.rc (src/driver.jl:34, src/neighbors.jl:28-36) */
  st->cs = st->rc;
  int64_t nx = (int64_t)ceil((double)st->p.Lx / st->cs) + 2;
  int64_t ny = (int64_t)ceil((double)st->p.Ly / st->cs) + 2;
  st->nx = nx > 1 ? nx : 1;
  st->ny = ny > 1 ? ny : 1;
  st->ncell = st->nx * st->ny;
  st->cellof = (int32_t *)xcalloc(N, sizeof(int32_t));
  st->order = (int32_t *)xcalloc(N, sizeof(int32_t));
  st->counts = (int32_t *)xcalloc(st->ncell, sizeof(int32_t));
  st->cum = (int32_t *)xcalloc(st->ncell, sizeof(int32_t));
  st->starts = (int32_t *)xcalloc(st->ncell, sizeof(int32_t));
  st->cursor = (int32_t *)xcalloc(st->ncell, sizeof(int32_t));
  return st;
}

void F_(orc_state_destroy)(F_(orc_state) *st) {
  if (!st) return;
  free(st->X); free(st->V); free(st->a); free(st->rho); free(st->pterm);
  free(st->invrho); free(st->alpha); free(st->cellof); free(st->counts); free(st->cum);
  free(st->starts); free(st->cursor); free(st->order); free(st->indices);
  free(st->nbcount); free(st->Xref); free(st);
}

int64_t F_(orc_state_n)(const F_(orc_state) *st) { return st->N; }

static REAL *F_(field_ptr)(const F_(orc_state) *st, int which, int64_t *len) {
  switch (which) {
    case ORC_X: *len = 2 * st->N; return st->X;
    case ORC_V: *len = 2 * st->N; return st->V;
    case ORC_A: *len = 2 * st->N; return st->a;
    case ORC_RHO: *len = st->N; return st->rho;
    case ORC_PTERM: *len = st->N; return st->pterm;
    case ORC_INVRHO: *len = st->N; return st->invrho;
    case ORC_ALPHA: *len = st->N; return st->alpha;
    case ORC_XREF: *len = 2 * st->N; return st->Xref;
  }
  *len = 0;
  return NULL;
}
void F_(orc_state_get)(const F_(orc_state) *st, int which, REAL *out) {
  int64_t n;
  REAL *src = F_(field_ptr)(st, which, &n);
  if (src) memcpy(out, src, sizeof(REAL) * n);
}
void F_(orc_state_set)(F_(orc_state) *st, int which, const REAL *in) {
  int64_t n;
  REAL *dst = F_(field_ptr)(st, which, &n);
  if (dst) memcpy(dst, in, sizeof(REAL) * n);
}
/* nx, ny, ncell, maxnb, sm, si, interval, age */
void F_(orc_state_info)(const F_(orc_state) *st, int64_t *o) {
  o[0] = st->nx; o[1] = st->ny; o[2] = st->ncell; o[3] = st->maxnb;
  o[4] = st->sm; o[5] = st->si; o[6] = st->interval; o[7] = st->age;
}
void F_(orc_cell_list_get)(const F_(orc_state) *st, int32_t *cellof, int32_t *counts,
                           int32_t *starts, int32_t *order) {
  if (cellof) memcpy(cellof, st->cellof, sizeof(int32_t) * st->N);
  if (counts) memcpy(counts, st->counts, sizeof(int32_t) * st->ncell);
  if (starts) memcpy(starts, st->starts, sizeof(int32_t) * st->ncell);
  if (order) memcpy(order, st->order, sizeof(int32_t) * st->N);
}
void F_(orc_neighbors_get)(const F_(orc_state) *st, int32_t *indices, int32_t *nbcount,
                           int32_t *flags) {
  if (indices) memcpy(indices, st->indices, sizeof(int32_t) * st->N * st->maxnb);
  if (nbcount) memcpy(nbcount, st->nbcount, sizeof(int32_t) * st->N);
  if (flags) { flags[0] = st->flags[0]; flags[1] = st->flags[1]; }
}
void F_(orc_flags_clear)(F_(orc_state) *st) { st->flags[0] = 0; st->flags[1] = 0; }

/* cell_index (src/neighbors.jl:38-42): floor(x / cs) -- a division, in T */
static inline int32_t F_(cell_index)(REAL x, REAL y, REAL cs, int64_t nx, int64_t ny) {
  int64_t cx = clampi64(F_(trunc_i64)(FLOORR(x / cs)) + 2, 1, nx);
  int64_t cy = clampi64(F_(trunc_i64)(FLOORR(y / cs)) + 2, 1, ny);
  return (int32_t)(cx + (cy - 1) * nx);
}

/* build! (src/neighbors.jl:97-114).  _cl_assign! and _cl_fill! run their
 * atomic increments sequentially here; the per-cell sort makes the final
 * `order` independent of the fill order, exactly as in the reference. */
void F_(orc_build)(F_(orc_state) *st) {
  const int64_t N = st->N, nc = st->ncell;
  const REAL cs = (REAL)st->cs;
  memset(st->counts, 0, sizeof(int32_t) * nc);                       /* fill!(counts, 0) */
  for (int64_t i = 0; i < N; ++i) {                                  /* _cl_assign! */
    int32_t c = F_(cell_index)(st->X[2 * i], st->X[2 * i + 1], cs, st->nx, st->ny);
    st->cellof[i] = c;
    st->counts[c - 1] += 1;
  }
  int32_t run = 0;                                                   /* cumsum! */
  for (int64_t c = 0; c < nc; ++c) { run += st->counts[c]; st->cum[c] = run; }
  for (int64_t c = 0; c < nc; ++c) {
    st->starts[c] = st->cum[c] - st->counts[c];                      /* starts .= cum .- counts */
    st->cursor[c] = st->starts[c];                                   /* cursor .= starts */
  }
  for (int64_t i = 0; i < N; ++i) {                                  /* _cl_fill! */
    int32_t c = st->cellof[i];
    int32_t pos = ++st->cursor[c - 1];        /* Atomix @atomic += returns the new value */
    st->order[pos - 1] = (int32_t)(i + 1);
  }
  /* _cl_sort_cells! (src/neighbors.jl:60-80): insertion sort per cell */
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nc; ++c) {
    int32_t s = st->starts[c], n = st->counts[c];
    int32_t *o = st->order + s;               /* o[k-1] == order[s+k] (1-based k) */
    for (int32_t k = 2; k <= n; ++k) {
      int32_t v = o[k - 1];
      int32_t j = k - 1;
      while (j >= 1 && o[j - 1] > v) { o[j] = o[j - 1]; j -= 1; }
      o[j] = v;
    }
  }
}

/* nl_index(m, i, sm, si) = m*sm + (i-1)*si + 1 (src/neighbors.jl:148); 0-based here */
static inline int64_t nl_slot0(int64_t m, int64_t i1, int64_t sm, int64_t si) {
  return m * sm + (i1 - 1) * si;
}

/* _nl_build! (src/neighbors.jl:192-245) */
static void F_(nl_build)(F_(orc_state) *st, int check) {
  const int64_t N = st->N;
  const REAL rc2 = (REAL)(st->rc * st->rc);
  const REAL r2min = (EPSR * st->p.h) * st->p.h;
  const REAL csT = (REAL)st->cs;
  const REAL invcs = (REAL)1 / csT;
  const REAL dmax2 = (REAL)(st->dmax * st->dmax);
  const int64_t nx = st->nx, ny = st->ny, mx = st->maxnb, sm = st->sm, si = st->si;
  int32_t f1 = 0, f2 = 0;
#pragma omp parallel for schedule(static) reduction(| : f1, f2)
  for (int64_t i0 = 0; i0 < N; ++i0) {
    const int64_t i1 = i0 + 1;
    REAL xi = st->X[2 * i0], yi = st->X[2 * i0 + 1];
    if (check) {
      REAL ddx = xi - st->Xref[2 * i0];
      REAL ddy = yi - st->Xref[2 * i0 + 1];
      if (ddx * ddx + ddy * ddy > dmax2) f2 = 1;
    }
    st->Xref[2 * i0] = xi;
    st->Xref[2 * i0 + 1] = yi;
    /* own cell by multiplication with invcs (not the division of cell_index) */
    int64_t cx = clampi64(F_(trunc_i32)(FLOORR(xi * invcs)) + 2, 1, nx);
    int64_t cy = clampi64(F_(trunc_i32)(FLOORR(yi * invcs)) + 2, 1, ny);
    int64_t m = 0;
    for (int64_t ddy = -1; ddy <= 1; ++ddy) {
      for (int64_t ddx = -1; ddx <= 1; ++ddx) {
        int64_t jx = cx + ddx, jy = cy + ddy;
        if (1 <= jx && jx <= nx && 1 <= jy && jy <= ny) {
          int64_t c = jx + (jy - 1) * nx;
          int32_t s = st->starts[c - 1], n = st->counts[c - 1];
          for (int32_t k = 1; k <= n; ++k) {
            int32_t j = st->order[s + k - 1];
            REAL dx = xi - st->X[2 * (j - 1)];
            REAL dy = yi - st->X[2 * (j - 1) + 1];
            REAL r2 = dx * dx + dy * dy;
            if (r2min < r2 && r2 < rc2) {
              if (m < mx) {
                st->indices[nl_slot0(m, i1, sm, si)] = j;
                m += 1;
              } else {
                f1 = 1;
              }
            }
          }
        }
      }
    }
    st->nbcount[i0] = (int32_t)m;
  }
  if (f1) st->flags[0] = 1;
  if (f2) st->flags[1] = 1;
}

/* build_neighbors! (src/neighbors.jl:252-263) */
void F_(orc_build_neighbors)(F_(orc_state) *st) {
  F_(orc_build)(st);
  int check = st->age != AGE_NEVER;
  F_(nl_build)(st, check);
  st->age = 0;
}

/* maybe_rebuild! (src/neighbors.jl:273-279) */
void F_(orc_maybe_rebuild)(F_(orc_state) *st) {
  if (st->age >= st->interval) F_(orc_build_neighbors)(st);
  st->age += 1;
}

/* ---- src/forward.jl ---------------------------------------------------- */

/* density_kernel! (src/forward.jl:25-65) */
void F_(orc_density)(F_(orc_state) *st) {
  const int64_t N = st->N, sm = st->sm, si = st->si;
  const REAL h = st->p.h, mass = st->p.m;
  const REAL A = F_(wnorm_)(h);
  const REAL invh = (REAL)1 / h;
  const REAL r2max = ((REAL)4 * h) * h;
#pragma omp parallel for schedule(static)
  for (int64_t i0 = 0; i0 < N; ++i0) {
    const int64_t i1 = i0 + 1;
    REAL xi = st->X[2 * i0], yi = st->X[2 * i0 + 1];
    REAL acc = 0;
    for (int64_t m = 0; m < st->nbcount[i0]; ++m) {
      int64_t j0 = st->indices[nl_slot0(m, i1, sm, si)] - 1;
      REAL dx = xi - st->X[2 * j0];
      REAL dy = yi - st->X[2 * j0 + 1];
      REAL r2 = dx * dx + dy * dy;
      if (r2 < r2max) {
        REAL q = SQRTR(r2) * invh;
        REAL u = (REAL)1 - (REAL)0.5 * q;
        REAL u2 = u * u;
        acc += mass * (((A * u2) * u2) * ((REAL)2 * q + (REAL)1));
      }
    }
    st->rho[i0] = acc + mass * A;
  }
}

/* eos_kernel! (src/forward.jl:68-75), c2 = p.c^2 */
void F_(orc_eos)(F_(orc_state) *st) {
  const REAL c2 = st->p.c * st->p.c, rho0 = st->p.rho0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < st->N; ++i) {
    REAL ri = st->rho[i];
    REAL inv = (REAL)1 / ri;
    st->invrho[i] = inv;
    st->pterm[i] = ((c2 * (ri - rho0)) * inv) * inv;
  }
}

/* accel_kernel! (src/forward.jl:77-132) */
void F_(orc_accel)(F_(orc_state) *st, const REAL *theta) {
  const F_(prm) *p = &st->p;
  const int64_t N = st->N, sm = st->sm, si = st->si;
  const REAL h = p->h, mass = p->m;
  const REAL A = F_(wnorm_)(h);
  const REAL invh = (REAL)1 / h;
  const REAL r2max = ((REAL)4 * h) * h;
  const REAL Fc = (((REAL)-5 * A) * invh) * invh;
  const REAL twomumass = ((REAL)2 * mass) * p->mu;
#pragma omp parallel for schedule(static)
  for (int64_t i0 = 0; i0 < N; ++i0) {
    const int64_t i1 = i0 + 1;
    REAL xi = st->X[2 * i0], yi = st->X[2 * i0 + 1];
    REAL vxi = st->V[2 * i0], vyi = st->V[2 * i0 + 1];
    REAL pti = st->pterm[i0], iri = st->invrho[i0];
    REAL ax = 0, ay = 0;
    for (int64_t m = 0; m < st->nbcount[i0]; ++m) {
      int64_t j0 = st->indices[nl_slot0(m, i1, sm, si)] - 1;
      REAL dx = xi - st->X[2 * j0];
      REAL dy = yi - st->X[2 * j0 + 1];
      REAL r2 = dx * dx + dy * dy;
      if (r2 < r2max) {
        REAL q = SQRTR(r2) * invh;
        REAL u = (REAL)1 - (REAL)0.5 * q;
        REAL F = ((Fc * u) * u) * u;
        REAL Pij = mass * (pti + st->pterm[j0]);
        ax -= (Pij * F) * dx;
        ay -= (Pij * F) * dy;
        REAL C = ((twomumass * iri) * st->invrho[j0]) * F;
        ax += C * (vxi - st->V[2 * j0]);
        ay += C * (vyi - st->V[2 * j0 + 1]);
      }
    }
    REAL al, d1, d2;
    F_(interp_alpha)(theta, xi, yi, p, &al, &d1, &d2);
    st->alpha[i0] = al;
    ax += p->gx + F_(wall_accel)(xi, p->Lx, p->kw);
    ay += p->gy + F_(wall_accel)(yi, p->Ly, p->kw);
    st->a[2 * i0] = ax;
    st->a[2 * i0 + 1] = ay;
  }
}

/* integrate_kernel! (src/forward.jl:137-147) */
void F_(orc_integrate)(F_(orc_state) *st) {
  const REAL dt = st->p.dt;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < st->N; ++i) {
    REAL D = (REAL)1 / ((REAL)1 + dt * st->alpha[i]);
    REAL v1 = (st->V[2 * i] + dt * st->a[2 * i]) * D;
    REAL v2 = (st->V[2 * i + 1] + dt * st->a[2 * i + 1]) * D;
    st->V[2 * i] = v1;
    st->V[2 * i + 1] = v2;
    st->X[2 * i] += dt * v1;
    st->X[2 * i + 1] += dt * v2;
  }
}

/* step! (src/forward.jl:161-173) */
void F_(orc_step)(F_(orc_state) *st, const REAL *theta) {
  F_(orc_maybe_rebuild)(st);
  F_(orc_density)(st);
  F_(orc_eos)(st);
  F_(orc_accel)(st, theta);
  F_(orc_integrate)(st);
}

/* ---- Tape (src/driver.jl:45-83) ---------------------------------------- */

struct F_(orc_tape) {
  REAL *X, *V;   /* (2, N, capacity) */
  int64_t N, cap, n;
};

F_(orc_tape) *F_(orc_tape_create)(int64_t N, int64_t capacity) {
  F_(orc_tape) *tp = (F_(orc_tape) *)xcalloc(1, sizeof(*tp));
  tp->N = N; tp->cap = capacity; tp->n = 0;
  tp->X = (REAL *)xcalloc(2 * N * capacity, sizeof(REAL));
  tp->V = (REAL *)xcalloc(2 * N * capacity, sizeof(REAL));
  return tp;
}
void F_(orc_tape_destroy)(F_(orc_tape) *tp) {
  if (!tp) return;
  free(tp->X); free(tp->V); free(tp);
}
void F_(orc_tape_reset)(F_(orc_tape) *tp) { tp->n = 0; }
int64_t F_(orc_tape_length)(const F_(orc_tape) *tp) { return tp->n; }
int64_t F_(orc_tape_capacity)(const F_(orc_tape) *tp) { return tp->cap; }
void F_(orc_tape_get)(const F_(orc_tape) *tp, int64_t k, REAL *X, REAL *V) {
  /* k is 1-based like tape.X[:, :, k] */
  if (X) memcpy(X, tp->X + 2 * tp->N * (k - 1), sizeof(REAL) * 2 * tp->N);
  if (V) memcpy(V, tp->V + 2 * tp->N * (k - 1), sizeof(REAL) * 2 * tp->N);
}

/* simulate! (src/driver.jl:92-115).  Returns 1 (ArgumentError) when the tape
 * cannot hold nsteps more entries.  Neighbour flags stay in the state for the
 * caller to report and clear (check_neighbor_flags). */
int F_(orc_simulate)(F_(orc_state) *st, const REAL *theta, int64_t nsteps, F_(orc_tape) *tape) {
  const int64_t N = st->N;
  if (tape) {
    if (tape->n + nsteps > tape->cap) return 1;
  }
  for (int64_t s = 0; s < nsteps; ++s) {
    if (tape) {
      tape->n += 1;
      memcpy(tape->X + 2 * N * (tape->n - 1), st->X, sizeof(REAL) * 2 * N);
      memcpy(tape->V + 2 * N * (tape->n - 1), st->V, sizeof(REAL) * 2 * N);
    }
    F_(orc_step)(st, theta);
  }
  return 0;
}

/* ---- AdjointWorkspace (src/driver.jl:117-133) -------------------------- */

struct F_(orc_ws) {
  int64_t N, ngx, ngy;
  REAL *gX, *gV, *gXs, *gVs, *abar, *grho, *galpha, *gtheta;
};

F_(orc_ws) *F_(orc_ws_create)(int64_t N, const orc_params *p) {
  F_(orc_ws) *ws = (F_(orc_ws) *)xcalloc(1, sizeof(*ws));
  ws->N = N; ws->ngx = p->ngx; ws->ngy = p->ngy;
  ws->gX = (REAL *)xcalloc(2 * N, sizeof(REAL));
  ws->gV = (REAL *)xcalloc(2 * N, sizeof(REAL));
  ws->gXs = (REAL *)xcalloc(2 * N, sizeof(REAL));
  ws->gVs = (REAL *)xcalloc(2 * N, sizeof(REAL));
  ws->abar = (REAL *)xcalloc(2 * N, sizeof(REAL));
  ws->grho = (REAL *)xcalloc(N, sizeof(REAL));
  ws->galpha = (REAL *)xcalloc(N, sizeof(REAL));
  ws->gtheta = (REAL *)xcalloc(p->ngx * p->ngy, sizeof(REAL));
  return ws;
}
void F_(orc_ws_destroy)(F_(orc_ws) *ws) {
  if (!ws) return;
  free(ws->gX); free(ws->gV); free(ws->gXs); free(ws->gVs); free(ws->abar);
  free(ws->grho); free(ws->galpha); free(ws->gtheta); free(ws);
}
void F_(orc_ws_get)(const F_(orc_ws) *ws, int which, REAL *out) {
  const int64_t N = ws->N;
  switch (which) {
    case ORC_GX: memcpy(out, ws->gX, sizeof(REAL) * 2 * N); break;
    case ORC_GV: memcpy(out, ws->gV, sizeof(REAL) * 2 * N); break;
    case ORC_GXS: memcpy(out, ws->gXs, sizeof(REAL) * 2 * N); break;
    case ORC_GVS: memcpy(out, ws->gVs, sizeof(REAL) * 2 * N); break;
    case ORC_ABAR: memcpy(out, ws->abar, sizeof(REAL) * 2 * N); break;
    case ORC_GRHO: memcpy(out, ws->grho, sizeof(REAL) * N); break;
    case ORC_GALPHA: memcpy(out, ws->galpha, sizeof(REAL) * N); break;
    case ORC_GTHETA: memcpy(out, ws->gtheta, sizeof(REAL) * ws->ngx * ws->ngy); break;
  }
}

/* ---- src/adjoint.jl ---------------------------------------------------- */

/* adj_pass1_kernel! (src/adjoint.jl:30-122) */
static void F_(adj_pass1)(F_(orc_ws) *ws, const F_(orc_state) *sc) {
  const F_(prm) *p = &sc->p;
  const int64_t N = sc->N, sm = sc->sm, si = sc->si;
  const REAL h = p->h, mass = p->m;
  const REAL A = F_(wnorm_)(h);
  const REAL c2 = p->c * p->c;
  const REAL invh = (REAL)1 / h;
  const REAL r2max = ((REAL)4 * h) * h;
  const REAL Fc = (((REAL)-5 * A) * invh) * invh;
  const REAL Gc = (((REAL)7.5 * A) * invh) * invh;
  const REAL twomumass = ((REAL)2 * mass) * p->mu;
  const REAL *X = sc->X, *V = sc->V, *abar = ws->abar, *pterm = sc->pterm, *invrho = sc->invrho;
#pragma omp parallel for schedule(static)
  for (int64_t i0 = 0; i0 < N; ++i0) {
    const int64_t i1 = i0 + 1;
    REAL xi = X[2 * i0], yi = X[2 * i0 + 1];
    REAL vxi = V[2 * i0], vyi = V[2 * i0 + 1];
    REAL abx = abar[2 * i0], aby = abar[2 * i0 + 1];
    REAL pti = pterm[i0], iri = invrho[i0];
    REAL gx = 0, gy = 0, gvx = 0, gvy = 0, grh = 0, sumS = 0;
    for (int64_t m = 0; m < sc->nbcount[i0]; ++m) {
      int64_t j0 = sc->indices[nl_slot0(m, i1, sm, si)] - 1;
      REAL dx = xi - X[2 * j0];
      REAL dy = yi - X[2 * j0 + 1];
      REAL r2 = dx * dx + dy * dy;
      if (r2 < r2max) {
        REAL r = SQRTR(r2);
        REAL q = r * invh;
        REAL u = (REAL)1 - (REAL)0.5 * q;
        REAL F = ((Fc * u) * u) * u;
        REAL G = (Gc * u) * u;
        REAL Pij = mass * (pti + pterm[j0]);
        REAL Cbase = (twomumass * iri) * invrho[j0];
        REAL C = Cbase * F;
        REAL dax = abx - abar[2 * j0];
        REAL day = aby - abar[2 * j0 + 1];
        REAL dvx = vxi - V[2 * j0];
        REAL dvy = vyi - V[2 * j0 + 1];
        REAL ad_d = -(dax * dx + day * dy);
        REAL ad_v = dax * dvx + day * dvy;
        gx -= (Pij * F) * dax;
        gy -= (Pij * F) * day;
        REAL Fbar = Pij * ad_d + Cbase * ad_v;
        REAL w = ((Fbar * G) * invh) / r;
        gx += w * dx;
        gy += w * dy;
        gvx += C * dax;
        gvy += C * day;
        grh -= (C * ad_v) * iri;
        sumS += F * ad_d;
      }
    }
    grh += ((((REAL)-2 * mass) * pti) * iri) * sumS;
    grh += (((c2 * mass) * iri) * iri) * sumS;
    gx += abx * F_(wall_daccel)(xi, p->Lx, p->kw);
    gy += aby * F_(wall_daccel)(yi, p->Ly, p->kw);
    ws->gXs[2 * i0] = gx;
    ws->gXs[2 * i0 + 1] = gy;
    ws->gVs[2 * i0] = gvx;
    ws->gVs[2 * i0 + 1] = gvy;
    ws->grho[i0] = grh;
  }
}

/* adj_pass2_kernel! (src/adjoint.jl:126-156) */
static void F_(adj_pass2)(F_(orc_ws) *ws, const F_(orc_state) *sc) {
  const int64_t N = sc->N, sm = sc->sm, si = sc->si;
  const REAL h = sc->p.h, mass = sc->p.m;
  const REAL A = F_(wnorm_)(h);
  const REAL invh = (REAL)1 / h;
  const REAL r2max = ((REAL)4 * h) * h;
  const REAL Fc = (((REAL)-5 * A) * invh) * invh;
  const REAL *X = sc->X, *grho = ws->grho;
#pragma omp parallel for schedule(static)
  for (int64_t i0 = 0; i0 < N; ++i0) {
    const int64_t i1 = i0 + 1;
    REAL xi = X[2 * i0], yi = X[2 * i0 + 1];
    REAL gri = grho[i0];
    REAL gx = 0, gy = 0;
    for (int64_t m = 0; m < sc->nbcount[i0]; ++m) {
      int64_t j0 = sc->indices[nl_slot0(m, i1, sm, si)] - 1;
      REAL dx = xi - X[2 * j0];
      REAL dy = yi - X[2 * j0 + 1];
      REAL r2 = dx * dx + dy * dy;
      if (r2 < r2max) {
        REAL q = SQRTR(r2) * invh;
        REAL u = (REAL)1 - (REAL)0.5 * q;
        REAL w = (mass * (((Fc * u) * u) * u)) * (gri + grho[j0]);
        gx += w * dx;
        gy += w * dy;
      }
    }
    ws->gXs[2 * i0] += gx;
    ws->gXs[2 * i0 + 1] += gy;
  }
}

/* adj_design_kernel! (src/adjoint.jl:166-185).  The four float atomics into
 * gtheta run in particle order here, one deterministic instance of the
 * reference's unordered atomic sum. */
static void F_(adj_design)(F_(orc_ws) *ws, const F_(orc_state) *sc, const REAL *theta) {
  const F_(prm) *p = &sc->p;
  const int64_t N = sc->N, ny = p->ngy;
  for (int64_t i0 = 0; i0 < N; ++i0) {
    REAL xi = sc->X[2 * i0], yi = sc->X[2 * i0 + 1];
    REAL ga = ws->galpha[i0];
    REAL al, dadx, dady;
    F_(interp_alpha)(theta, xi, yi, p, &al, &dadx, &dady);
    ws->gXs[2 * i0] += ga * dadx;
    ws->gXs[2 * i0 + 1] += ga * dady;
    int64_t ii, jj;
    REAL fx, fy, ax, ay;
    F_(interp_cell)(xi, yi, p, &ii, &jj, &fx, &fy, &ax, &ay);
    int64_t l00 = jj + (ii - 1) * ny;        /* 1-based linear index */
    ws->gtheta[l00 - 1] += (ga * ((REAL)1 - fx)) * ((REAL)1 - fy);
    ws->gtheta[l00 + ny - 1] += (ga * fx) * ((REAL)1 - fy);
    ws->gtheta[l00] += (ga * ((REAL)1 - fx)) * fy;
    ws->gtheta[l00 + ny] += (ga * fx) * fy;
  }
}

/* _adj_drag! (src/adjoint.jl:214-228) */
static void F_(adj_drag)(F_(orc_ws) *ws, const F_(orc_state) *sc) {
  const REAL dt = sc->p.dt;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < sc->N; ++i) {
    REAL D = (REAL)1 / ((REAL)1 + dt * sc->alpha[i]);
    REAL vx = (sc->V[2 * i] + dt * sc->a[2 * i]) * D;
    REAL vy = (sc->V[2 * i + 1] + dt * sc->a[2 * i + 1]) * D;
    REAL gx = ws->gV[2 * i], gy = ws->gV[2 * i + 1];
    ws->abar[2 * i] = (dt * D) * gx;
    ws->abar[2 * i + 1] = (dt * D) * gy;
    ws->galpha[i] = ((-dt) * D) * (vx * gx + vy * gy);
    ws->gV[2 * i] = D * gx;
    ws->gV[2 * i + 1] = D * gy;
  }
}

/* backward! (src/driver.jl:141-207).  A fresh scratch State is made for every
 * call (the `scratch = nothing` branch), so the reverse sweep rebuilds its
 * neighbour list on its own schedule starting at the last step. */
int F_(orc_backward)(F_(orc_ws) *ws, const F_(orc_tape) *tape, const REAL *theta,
                     const orc_params *q, const REAL *seedX, const REAL *seedV,
                     double skin, int64_t interval, int64_t maxnb, int layout,
                     int32_t *flags_out) {
  const int64_t N = ws->N;
  memcpy(ws->gX, seedX, sizeof(REAL) * 2 * N);
  memcpy(ws->gV, seedV, sizeof(REAL) * 2 * N);
  memset(ws->gtheta, 0, sizeof(REAL) * ws->ngx * ws->ngy);
  F_(orc_state) *sc = F_(orc_state_create)(q, N, NULL, NULL, skin, interval, maxnb, layout);
  const REAL dt = sc->p.dt;
  for (int64_t n = tape->n; n >= 1; --n) {
    memcpy(sc->X, tape->X + 2 * N * (n - 1), sizeof(REAL) * 2 * N);   /* _tape_load! */
    memcpy(sc->V, tape->V + 2 * N * (n - 1), sizeof(REAL) * 2 * N);
    F_(orc_maybe_rebuild)(sc);
    F_(orc_density)(sc);
    F_(orc_eos)(sc);
    F_(orc_accel)(sc, theta);
    for (int64_t i = 0; i < 2 * N; ++i) ws->gV[i] += dt * ws->gX[i];  /* _adj_integrate! */
    F_(adj_drag)(ws, sc);
    F_(adj_pass1)(ws, sc);
    F_(adj_design)(ws, sc, theta);
    F_(adj_pass2)(ws, sc);
    for (int64_t i = 0; i < 2 * N; ++i) ws->gX[i] += ws->gXs[i];      /* _axpy2! */
    for (int64_t i = 0; i < 2 * N; ++i) ws->gV[i] += ws->gVs[i];
  }
  if (flags_out) { flags_out[0] = sc->flags[0]; flags_out[1] = sc->flags[1]; }
  F_(orc_state_destroy)(sc);
  return 0;
}

/* target_objective (src/driver.jl:219-234): serial sum in T */
REAL F_(orc_target_objective)(const REAL *X, int64_t N, double xc, double yc, double sigma,
                              REAL *seed) {
  REAL J = 0;
  REAL s2 = (REAL)sigma * (REAL)sigma;
  for (int64_t i = 0; i < N; ++i) {
    REAL dx = X[2 * i] - (REAL)xc;
    REAL dy = X[2 * i + 1] - (REAL)yc;
    REAL e = EXPR(-(dx * dx + dy * dy) / ((REAL)2 * s2));
    J -= e;
    if (seed) {
      seed[2 * i] = (e * dx) / s2;
      seed[2 * i + 1] = (e * dy) / s2;
    }
  }
  return J;
}

/* exported scalar helpers for the kernel unit tests */
REAL F_(orc_w_kern)(REAL q, REAL A) { return F_(w_kern_)(q, A); }
REAL F_(orc_f_kern)(REAL q, REAL A, REAL h) { return F_(f_kern_)(q, A, h); }
REAL F_(orc_g_kern)(REAL q, REAL A, REAL h) { return F_(g_kern_)(q, A, h); }
REAL F_(orc_wnorm)(REAL h) { return F_(wnorm_)(h); }
void F_(orc_interp_alpha)(const REAL *theta, const orc_params *q, REAL x, REAL y, REAL *out3) {
  F_(prm) p = F_(prm_from)(q);
  F_(interp_alpha)(theta, x, y, &p, &out3[0], &out3[1], &out3[2]);
}

#undef F_
#undef CAT_
#undef CAT2_
