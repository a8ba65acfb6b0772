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
// Each rebuild builds a transposed neighbour list (entry m of slot s at
// nl[m*N + s]) of the 3x3 cell block around the particle's own cell, in the
// reference's order (ddy outer, ddx inner, ascending id inside a cell); an
// entry packs the neighbour slot with the code of its cell in the block,
// (j << 4) | code, from which the exact FP32 shift between the two cell
// frames follows.  Every sum has a fixed order: repeated runs are
// bit-identical.  Physics follows src/forward.jl:25-147 and
// src/adjoint.jl:30-228.
#include "sph_f32c.h"

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
};

// interp_cell (src/params.jl:53-68) from the FP64 position
__device__ __forceinline__ InterpF interp_cell(double x, double y, const PhysF32 &P) {
  InterpF r;
  double gx = x / P.hx;
  double gy = y / P.hy;
  r.ax = (gx > 0.0 && gx < P.gxmax) ? 1.f : 0.f;
  r.ay = (gy > 0.0 && gy < P.gymax) ? 1.f : 0.f;
  double gxc = (gx > P.gxclamp) ? P.gxclamp : ((gx < 0.0) ? 0.0 : gx);
  double gyc = (gy > P.gyclamp) ? P.gyclamp : ((gy < 0.0) ? 0.0 : gy);
  double fxf = floor(gxc), fyf = floor(gyc);
  int i0 = (fxf > -1e9 && fxf < 1e9) ? (int)fxf : -1000000000;
  int j0 = (fyf > -1e9 && fyf < 1e9) ? (int)fyf : -1000000000;
  i0 = i0 < 0 ? 0 : (i0 > P.ngx - 2 ? P.ngx - 2 : i0);
  j0 = j0 < 0 ? 0 : (j0 > P.ngy - 2 ? P.ngy - 2 : j0);
  r.fx = (float)(gxc - (double)i0);
  r.fy = (float)(gyc - (double)j0);
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
  int cy = cell / g.nx;
  int cx = cell - cy * g.nx;
  double ox = (double)(cx - 1) * (double)P.cs;
  double oy = (double)(cy - 1) * (double)P.cs;
  return make_float2((float)(x.x - ox), (float)(x.y - oy));
}

// shift that moves an xr of the cell with code (dy+1)*3 + (dx+1), where
// (dx, dy) is its offset from the sweeping particle's cell, into the frame of
// that particle (exact multiples of cs in FP32)
__device__ __forceinline__ float2 shift_of(int code, float cs) {
  const int row = (code * 11) >> 5;            // code / 3 for code < 9
  const int col = code - row * 3;
  return make_float2((float)(col - 1) * cs, (float)(row - 1) * cs);
}

struct Pair32 {
  float dx, dy, r2;
  int j;
};
// pair of the sweeping particle (own frame xi) with list entry e
__device__ __forceinline__ Pair32 pair32(float2 xi, const float2 *__restrict__ xr, int e,
                                         float cs) {
  const int j = e >> 4;
  const float2 sh = shift_of(e & 15, cs);
  const float2 xj = xr[j];
  Pair32 p;
  p.dx = (xi.x - sh.x) - xj.x;
  p.dy = (xi.y - sh.y) - xj.y;
  p.r2 = p.dx * p.dx + p.dy * p.dy;
  p.j = j;
  return p;
}

// neighbour list of the 3x3 block around the binning cell, cutoff rc; N < 2^27
__global__ void __launch_bounds__(kBlock)
k_nl_build32(const float2 *__restrict__ xr, const int *__restrict__ bcell,
             const int *__restrict__ cstart, Grid g, PhysF32 P, float rc2, int maxnb,
             int *__restrict__ nl, int *__restrict__ nbcnt, int *__restrict__ flags) {
  int s = xcd_block() * blockDim.x + threadIdx.x;
  if (s >= g.N) return;
  const float2 xi = xr[s];
  const int bc = bcell[s];
  const int by = bc / g.nx;
  const int bx = bc - by * g.nx;
  int m = 0;
  bool over = false;
#pragma unroll
  for (int dyi = -1; dyi <= 1; ++dyi) {
    const int jy = by + dyi;
    if (jy < 0 || jy >= g.ny) continue;
    const int base = jy * g.nx;
#pragma unroll
    for (int dxi = -1; dxi <= 1; ++dxi) {
      const int jx = bx + dxi;
      if (jx < 0 || jx >= g.nx) continue;
      const int code = (dyi + 1) * 3 + (dxi + 1);
      const float2 sh = shift_of(code, P.cs);
      const float xs = xi.x - sh.x, ys = xi.y - sh.y;
      const int e = cstart[base + jx + 1];
      for (int j = cstart[base + jx]; j < e; ++j) {
        const float2 xj = xr[j];
        const float dx = xs - xj.x, dy = ys - xj.y;
        const float r2 = dx * dx + dy * dy;
        if (r2 < rc2 && r2 > P.r2min) {
          if (m < maxnb) {
            nl[(size_t)m * g.N + s] = (j << 4) | code;
            ++m;
          } else {
            over = true;
          }
        }
      }
    }
  }
  nbcnt[s] = m;
  if (over) flags[0] = 1;
}

__device__ __forceinline__ float2 density_pair(float2 xi, const float2 *__restrict__ xr,
                                               const NbRow &row, const PhysF32 &P) {
  CSum acc;
  for (int m = 0; m < row.n; ++m) {
    const Pair32 pr = pair32(xi, xr, row.at(m), P.cs);
    if (pr.r2 < P.r2max) {
      float q = __builtin_amdgcn_sqrtf(pr.r2) * P.invh;
      float u = 1.f - 0.5f * q;
      float u2 = u * u;
      acc.add(P.mA * (u2 * u2 * (2.f * q + 1.f)));
    }
  }
  acc.add(P.mA);                  // self term
  // rho - rho0 from the compensated pair (exact subtraction, then the tail)
  float drho = (acc.s - P.rho0) + acc.c;
  return make_float2(P.rho0 + drho, drho);
}

__device__ __forceinline__ float2 eos(float2 rd, const PhysF32 &P) {
  float inv = 1.f / rd.x;
  return make_float2(P.c2 * rd.y * inv * inv, inv);    // (pterm, invrho)
}

__global__ void __launch_bounds__(kBlock)
k_density32(const float2 *__restrict__ xr, const int *__restrict__ nl,
            const int *__restrict__ nbcnt, Grid g, PhysF32 P, float2 *__restrict__ PI,
            double *__restrict__ rho_out) {
  int s = xcd_block() * blockDim.x + threadIdx.x;
  if (s >= g.N) return;
  float2 rd = density_pair(xr[s], xr, nb_row(nl, nbcnt, s, g.N), P);
  PI[s] = eos(rd, P);
  if (rho_out) rho_out[s] = (double)rd.x;
}

__global__ void __launch_bounds__(kBlock)
k_force32(const double2 *__restrict__ X, const float2 *__restrict__ xr,
          const float2 *__restrict__ V, const float2 *__restrict__ PI,
          const double *__restrict__ theta, const int *__restrict__ bcell,
          const int *__restrict__ nl, const int *__restrict__ nbcnt, Grid g, PhysF32 P,
          double2 *__restrict__ Xn, float2 *__restrict__ xrn, float2 *__restrict__ Vn,
          double2 *__restrict__ tapeX, float2 *__restrict__ tapeV, double2 *__restrict__ a_out,
          double *__restrict__ alpha_out) {
  int s = xcd_block() * blockDim.x + threadIdx.x;
  if (s >= g.N) return;
  const double2 x64 = X[s];
  const float2 xi = xr[s], vi = V[s], pii = PI[s];
  const NbRow row = nb_row(nl, nbcnt, s, g.N);
  CSum ax, ay;
  for (int m = 0; m < row.n; ++m) {
    const Pair32 pr = pair32(xi, xr, row.at(m), P.cs);
    if (pr.r2 < P.r2max) {
      float q = __builtin_amdgcn_sqrtf(pr.r2) * P.invh;
      float u = 1.f - 0.5f * q;
      float F = P.Fc * u * u * u;
      float2 pij = PI[pr.j];
      float PF = P.m * (pii.x + pij.x) * F;
      float C = P.twomumass * pii.y * pij.y * F;
      float2 vj = V[pr.j];
      ax.add(C * (vi.x - vj.x) - PF * pr.dx);
      ay.add(C * (vi.y - vj.y) - PF * pr.dy);
    }
  }
  InterpF ic = interp_cell(x64.x, x64.y, P);
  float al = interp_alpha(theta, ic, P, nullptr, nullptr);
  float axv = ax.value() + (P.gx + wall_accel((float)x64.x, P.Lx, P.kw));
  float ayv = ay.value() + (P.gy + wall_accel((float)x64.y, P.Ly, P.kw));
  if (a_out) {
    a_out[s] = make_double2(axv, ayv);
    alpha_out[s] = al;
    return;
  }
  float D = 1.f / (1.f + P.dt * al);
  float v1 = (vi.x + P.dt * axv) * D;
  float v2 = (vi.y + P.dt * ayv) * D;
  Vn[s] = make_float2(v1, v2);
  double2 xn = make_double2(x64.x + P.dt64 * (double)v1, x64.y + P.dt64 * (double)v2);
  Xn[s] = xn;
  xrn[s] = rel_coord(xn, bcell[s], g, P);
  if (tapeX) {
    tapeX[s] = x64;
    tapeV[s] = vi;
  }
}

__global__ void k_xrel(const double2 *__restrict__ X, const int *__restrict__ bcell, Grid g,
                       PhysF32 P, float2 *__restrict__ xr) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < g.N) xr[s] = rel_coord(X[s], bcell[s], g, P);
}

__global__ void k_permute32(const int *__restrict__ perm, const int *__restrict__ cellof,
                            const double2 *__restrict__ X, const float2 *__restrict__ V,
                            const double2 *__restrict__ Xref, int check, Grid g, PhysF32 P,
                            double dmax2, double2 *__restrict__ Xn, float2 *__restrict__ Vn,
                            float2 *__restrict__ xr, double2 *__restrict__ Xrefn,
                            int *__restrict__ bcell, int *__restrict__ flags) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int o = perm[t];
  double2 x = X[o];
  if (check) {
    double2 r = Xref[o];
    double ddx = x.x - r.x, ddy = x.y - r.y;
    if (ddx * ddx + ddy * ddy > dmax2) flags[1] = 1;
  }
  int c = cellof[o];
  Xn[t] = x;
  Vn[t] = V[o];
  Xrefn[t] = x;
  bcell[t] = c;
  xr[t] = rel_coord(x, c, g, P);
}

// adjoint R: density recompute, _adj_integrate!, _adj_drag! (src/adjoint.jl:193-228)
__global__ void __launch_bounds__(kBlock)
k_adj_r32(const double2 *__restrict__ X, const float2 *__restrict__ xr,
          const float2 *__restrict__ Vnext, const double *__restrict__ theta,
          const int *__restrict__ nl, const int *__restrict__ nbcnt, Grid g, PhysF32 P,
          float2 *__restrict__ PI, float2 *__restrict__ gX, float2 *__restrict__ gV,
          float2 *__restrict__ abar, float *__restrict__ galpha) {
  int s = xcd_block() * blockDim.x + threadIdx.x;
  if (s >= g.N) return;
  const double2 x64 = X[s];
  float2 rd = density_pair(xr[s], xr, nb_row(nl, nbcnt, s, g.N), P);
  PI[s] = eos(rd, P);
  InterpF ic = interp_cell(x64.x, x64.y, P);
  float al = interp_alpha(theta, ic, P, nullptr, nullptr);
  float2 gx = gX[s], gv = gV[s], vp = Vnext[s];
  float gvx = gv.x + P.dt * gx.x;
  float gvy = gv.y + P.dt * gx.y;
  float D = 1.f / (1.f + P.dt * al);
  float dtD = P.dt * D;
  abar[s] = make_float2(dtD * gvx, dtD * gvy);
  galpha[s] = -dtD * (vp.x * gvx + vp.y * gvy);
  gV[s] = make_float2(D * gvx, D * gvy);
}

// adjoint R without the density recompute (tapes that kept each step's
// pterm/invrho); also forms this step's relative coordinates xr
__global__ void __launch_bounds__(kBlock)
k_adj_rpw32(const double2 *__restrict__ X, const int *__restrict__ bcell,
            const float2 *__restrict__ Vnext, const double *__restrict__ theta, Grid g, PhysF32 P,
            float2 *__restrict__ xr, float2 *__restrict__ gX, float2 *__restrict__ gV,
            float2 *__restrict__ abar, float *__restrict__ galpha) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.N) return;
  const double2 x64 = X[s];
  xr[s] = rel_coord(x64, bcell[s], g, P);
  InterpF ic = interp_cell(x64.x, x64.y, P);
  float al = interp_alpha(theta, ic, P, nullptr, nullptr);
  float2 gx = gX[s], gv = gV[s], vp = Vnext[s];
  float gvx = gv.x + P.dt * gx.x;
  float gvy = gv.y + P.dt * gx.y;
  float D = 1.f / (1.f + P.dt * al);
  float dtD = P.dt * D;
  abar[s] = make_float2(dtD * gvx, dtD * gvy);
  galpha[s] = -dtD * (vp.x * gvx + vp.y * gvy);
  gV[s] = make_float2(D * gvx, D * gvy);
}

// adjoint P1: adj_pass1_kernel! + design term + gV += gVs (src/adjoint.jl:30-122,166-185)
__global__ void __launch_bounds__(kBlock)
k_adj_p132(const double2 *__restrict__ X, const float2 *__restrict__ xr,
           const float2 *__restrict__ V, const float2 *__restrict__ PI,
           const float2 *__restrict__ abar, const float *__restrict__ galpha,
           const double *__restrict__ theta, const int *__restrict__ nl,
           const int *__restrict__ nbcnt, Grid g, PhysF32 P, float2 *__restrict__ gXs,
           float2 *__restrict__ gV, float *__restrict__ grho) {
  int s = xcd_block() * blockDim.x + threadIdx.x;
  if (s >= g.N) return;
  const double2 x64 = X[s];
  const float2 xi = xr[s], vi = V[s], ab = abar[s], pii = PI[s];
  const float pti = pii.x, iri = pii.y;
  const NbRow row = nb_row(nl, nbcnt, s, g.N);
  CSum gx, gy, gvx, gvy, grh, sumS;
  for (int m = 0; m < row.n; ++m) {
    const Pair32 pr = pair32(xi, xr, row.at(m), P.cs);
    if (pr.r2 < P.r2max) {
      const int j = pr.j;
      const float dx = pr.dx, dy = pr.dy;
      float rinv = __builtin_amdgcn_rsqf(pr.r2);
      float q = (pr.r2 * rinv) * P.invh;
      float u = 1.f - 0.5f * q;
      float F = P.Fc * u * u * u;
      float G = P.Gc * u * u;
      float2 pij = PI[j];
      float Pij = P.m * (pti + pij.x);
      float Cbase = P.twomumass * iri * pij.y;
      float C = Cbase * F;
      float2 abj = abar[j];
      float2 vj = V[j];
      float dax = ab.x - abj.x;
      float day = ab.y - abj.y;
      float dvx = vi.x - vj.x;
      float dvy = vi.y - vj.y;
      float ad_d = -(dax * dx + day * dy);
      float ad_v = dax * dvx + day * dvy;
      float Fbar = Pij * ad_d + Cbase * ad_v;
      float wgt = Fbar * G * P.invh * rinv;
      float PF = Pij * F;
      gx.add(wgt * dx - PF * dax);
      gy.add(wgt * dy - PF * day);
      gvx.add(C * dax);
      gvy.add(C * day);
      grh.add(-(C * ad_v * iri));
      sumS.add(F * ad_d);
    }
  }
  float S = sumS.value();
  float grv = grh.value() + (-2.f * P.m * pti * iri) * S + (P.c2 * P.m * iri * iri) * S;
  float gxv = gx.value() + ab.x * wall_daccel((float)x64.x, P.Lx, P.kw);
  float gyv = gy.value() + ab.y * wall_daccel((float)x64.y, P.Ly, P.kw);
  float ga = galpha[s];
  InterpF ic = interp_cell(x64.x, x64.y, P);
  float dadx, dady;
  interp_alpha(theta, ic, P, &dadx, &dady);
  gxv += ga * dadx;
  gyv += ga * dady;
  gXs[s] = make_float2(gxv, gyv);
  grho[s] = grv;
  float2 gvo = gV[s];
  gV[s] = make_float2(gvo.x + gvx.value(), gvo.y + gvy.value());
}

// adjoint P2: adj_pass2_kernel! and gX += gXs (src/adjoint.jl:126-156,187-191)
__global__ void __launch_bounds__(kBlock)
k_adj_p232(const float2 *__restrict__ xr, const float *__restrict__ grho,
           const float2 *__restrict__ gXs, const int *__restrict__ nl,
           const int *__restrict__ nbcnt, Grid g, PhysF32 P, float2 *__restrict__ gX) {
  int s = xcd_block() * blockDim.x + threadIdx.x;
  if (s >= g.N) return;
  const float2 xi = xr[s];
  const float gri = grho[s];
  const NbRow row = nb_row(nl, nbcnt, s, g.N);
  CSum gx, gy;
  for (int m = 0; m < row.n; ++m) {
    const Pair32 pr = pair32(xi, xr, row.at(m), P.cs);
    if (pr.r2 < P.r2max) {
      float q = __builtin_amdgcn_sqrtf(pr.r2) * P.invh;
      float u = 1.f - 0.5f * q;
      float wgt = P.m * (P.Fc * u * u * u) * (gri + grho[pr.j]);
      gx.add(wgt * pr.dx);
      gy.add(wgt * pr.dy);
    }
  }
  float2 gs = gXs[s];
  float2 go = gX[s];
  gX[s] = make_float2(go.x + (gs.x + gx.value()), go.y + (gs.y + gy.value()));
}

}  // namespace

void launch_nl_build_f32c(hipStream_t s, const float2 *xr, const int *bcell, const int *cstart,
                          Grid g, PhysF32 P, float rc2, int maxnb, int *nl, int *nbcnt,
                          int *flags) {
  if (g.N > 0)
    k_nl_build32<<<nblocks(g.N), kBlock, 0, s>>>(xr, bcell, cstart, g, P, rc2, maxnb, nl, nbcnt,
                                                 flags);
}
void launch_density_f32c(hipStream_t s, const float2 *xr, const int *nl, const int *nbcnt, Grid g,
                         PhysF32 P, float2 *PI, double *rho) {
  if (g.N > 0) k_density32<<<nblocks(g.N), kBlock, 0, s>>>(xr, nl, nbcnt, g, P, PI, rho);
}
void launch_force_f32c(hipStream_t s, const double2 *X, const float2 *xr, const float2 *V,
                       const float2 *PI, const double *theta, const int *bcell, const int *nl,
                       const int *nbcnt, Grid g, PhysF32 P, double2 *Xn, float2 *xrn, float2 *Vn,
                       double2 *tapeX, float2 *tapeV, double2 *a_out, double *alpha_out) {
  if (g.N > 0)
    k_force32<<<nblocks(g.N), kBlock, 0, s>>>(X, xr, V, PI, theta, bcell, nl, nbcnt, g, P, Xn, xrn,
                                              Vn, tapeX, tapeV, a_out, alpha_out);
}
void launch_xrel_f32c(hipStream_t s, const double2 *X, const int *bcell, Grid g, PhysF32 P,
                      float2 *xr) {
  if (g.N > 0) k_xrel<<<nblocks(g.N), kBlock, 0, s>>>(X, bcell, g, P, xr);
}
void launch_permute_f32c(hipStream_t s, const int *perm, const int *cellof, const double2 *X,
                         const float2 *V, const double2 *Xref, int check, Grid g, PhysF32 P,
                         double dmax2, double2 *Xn, float2 *Vn, float2 *xr, double2 *Xrefn,
                         int *bcell, int *flags) {
  if (g.N > 0)
    k_permute32<<<nblocks(g.N), kBlock, 0, s>>>(perm, cellof, X, V, Xref, check, g, P, dmax2, Xn,
                                                Vn, xr, Xrefn, bcell, flags);
}
void launch_adj_r_f32c(hipStream_t s, const double2 *X, const float2 *xr, const float2 *Vnext,
                       const double *theta, const int *nl, const int *nbcnt, Grid g, PhysF32 P,
                       float2 *PI, float2 *gX, float2 *gV, float2 *abar, float *galpha) {
  if (g.N > 0)
    k_adj_r32<<<nblocks(g.N), kBlock, 0, s>>>(X, xr, Vnext, theta, nl, nbcnt, g, P, PI, gX, gV,
                                              abar, galpha);
}
void launch_adj_rpw_f32c(hipStream_t s, const double2 *X, const int *bcell, const float2 *Vnext,
                         const double *theta, Grid g, PhysF32 P, float2 *xr, float2 *gX,
                         float2 *gV, float2 *abar, float *galpha) {
  if (g.N > 0)
    k_adj_rpw32<<<nblocks(g.N), kBlock, 0, s>>>(X, bcell, Vnext, theta, g, P, xr, gX, gV, abar,
                                                galpha);
}
void launch_adj_p1_f32c(hipStream_t s, const double2 *X, const float2 *xr, const float2 *V,
                        const float2 *PI, const float2 *abar, const float *galpha,
                        const double *theta, const int *nl, const int *nbcnt, Grid g, PhysF32 P,
                        float2 *gXs, float2 *gV, float *grho) {
  if (g.N > 0)
    k_adj_p132<<<nblocks(g.N), kBlock, 0, s>>>(X, xr, V, PI, abar, galpha, theta, nl, nbcnt, g, P,
                                               gXs, gV, grho);
}
void launch_adj_p2_f32c(hipStream_t s, const float2 *xr, const float *grho, const float2 *gXs,
                        const int *nl, const int *nbcnt, Grid g, PhysF32 P, float2 *gX) {
  if (g.N > 0) k_adj_p232<<<nblocks(g.N), kBlock, 0, s>>>(xr, grho, gXs, nl, nbcnt, g, P, gX);
}

}  // namespace sph
