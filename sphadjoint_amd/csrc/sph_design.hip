// sph_design.hip -- the design-field chain of the reference's topology
// optimiser (scripts/03_optimize.jl:142-244) on the device, in FP64:
//
//   rd --[density filter, radius r, clamp to [0,1]]--> rdf --[alpha(rdf)]--> theta
//   gtheta --[* dalpha(rdf)]--> grdf --[filter transpose]--> grd
//
// plus the volume projection (project_volume, bisection on a uniform shift)
// and the line-search candidate project(rd - step * g / max|g|).
//
// Fields are (ngy, ngx) column-major: node (j, i) (0-based) at j + i*ngy.
// Each kernel evaluates the reference's expression in the reference's order,
// operation for operation (compiled with -ffp-contract=off), so that the
// filter, its transpose and the Borrvall-Petersson map give the reference's
// Float64 bits.  The filter weight of a node pair at squared grid distance k
// is wtab[k] = (r+1) - sqrt(k) and the normalisation of node (j, i) is snorm,
// both computed on the host with IEEE double arithmetic.
#pragma clang fp contract(off)
#include "sph_internal.h"

namespace sph {

namespace {

inline int nblocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// apply_filter (scripts/03_optimize.jl:166-183): acc = sum over the list of
// node (j, i) -- jj outer, ii inner -- of (w / s) * x[jj, ii]; clamp to [0, 1]
__global__ void k_filter(const double *__restrict__ x, const double *__restrict__ wtab,
                         const double *__restrict__ snorm, DesignGrid D, double *__restrict__ y) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= D.ngy * D.ngx) return;
  const int j = l % D.ngy, i = l / D.ngy;
  const double s = snorm[l];
  double acc = 0.0;
  for (int jj = max(0, j - D.r); jj <= min(D.ngy - 1, j + D.r); ++jj) {
    for (int ii = max(0, i - D.r); ii <= min(D.ngx - 1, i + D.r); ++ii) {
      const double w = wtab[(ii - i) * (ii - i) + (jj - j) * (jj - j)];
      if (w > 0.0) acc += (w / s) * x[jj + ii * D.ngy];
    }
  }
  y[l] = acc < 0.0 ? 0.0 : (acc > 1.0 ? 1.0 : acc);
}

// filter_adjoint (scripts/03_optimize.jl:186-194) as a gather: node (jj, ii)
// receives (w / s(j,i)) * g[j, i] from every node (j, i) whose list holds it,
// in the reference's scatter order (j outer, i inner)
__global__ void k_filter_adjoint(const double *__restrict__ g, const double *__restrict__ wtab,
                                 const double *__restrict__ snorm, DesignGrid D,
                                 double *__restrict__ out) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= D.ngy * D.ngx) return;
  const int jj = l % D.ngy, ii = l / D.ngy;
  double acc = 0.0;
  for (int j = max(0, jj - D.r); j <= min(D.ngy - 1, jj + D.r); ++j) {
    for (int i = max(0, ii - D.r); i <= min(D.ngx - 1, ii + D.r); ++i) {
      const double w = wtab[(ii - i) * (ii - i) + (jj - j) * (jj - j)];
      const int src = j + i * D.ngy;
      if (w > 0.0) acc += (w / snorm[src]) * g[src];
    }
  }
  out[l] = acc;
}

// alpha_of (scripts/03_optimize.jl:150): alpha_max * qp * (1 - rd) / (qp + rd)
__global__ void k_alpha(const double *__restrict__ rdf, DesignGrid D, double *__restrict__ theta) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= D.ngy * D.ngx) return;
  const double rd = rdf[l];
  theta[l] = D.alpha_max * D.qp * (1.0 - rd) / (D.qp + rd);
}

// grdf = gtheta .* dalpha_of.(rdf) (scripts/03_optimize.jl:151,226):
// dalpha_of(rd) = -alpha_max * qp * (qp + 1) / (qp + rd)^2
__global__ void k_chain(const double *__restrict__ gtheta, const double *__restrict__ rdf,
                        DesignGrid D, double *__restrict__ grdf) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= D.ngy * D.ngx) return;
  const double rd = rdf[l];
  const double q = D.qp + rd;
  grdf[l] = gtheta[l] * (-D.alpha_max * D.qp * (D.qp + 1.0) / (q * q));
}

// ---- one-workgroup reductions (fixed order: strided per thread, then a tree)
constexpr int kRedBlock = 1024;

__device__ __forceinline__ double block_sum(double v, double *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int w = kRedBlock / 2; w > 0; w >>= 1) {
    if (t < w) sh[t] = sh[t] + sh[t + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

// project_volume (scripts/03_optimize.jl:205-219), optionally applied to the
// line-search candidate rd - step * g / max|g| (scripts/03_optimize.jl:311):
//   x = clamp(in, 0, 1); if mean(x) >= volfrac return x;
//   60 bisection steps on mid in [0, 1]: mean(clamp(x + mid)) < volfrac ?
//   lo = mid : hi = mid; return clamp(x + (lo + hi) / 2).
// One workgroup; the sums run in a fixed order (each thread's nodes in
// ascending order, then a binary tree), so the result is the same on every
// run.  gmax (when g is given) receives max|g|.
__global__ void __launch_bounds__(kRedBlock)
k_project(const double *__restrict__ in, const double *__restrict__ g, double step, int n,
          double volfrac, double *__restrict__ out, double *__restrict__ gmax) {
  __shared__ double sh[kRedBlock];
  const int t = threadIdx.x;
  double gn = 0.0;
  if (g) {
    double m = 0.0;
    for (int l = t; l < n; l += kRedBlock) m = fmax(m, fabs(g[l]));
    sh[t] = m;
    __syncthreads();
    for (int w = kRedBlock / 2; w > 0; w >>= 1) {
      if (t < w) sh[t] = fmax(sh[t], sh[t + w]);
      __syncthreads();
    }
    gn = sh[0];
    __syncthreads();
    if (t == 0 && gmax) *gmax = gn;
  }
  // x = clamp(in - step * g / gn) (or clamp(in)), kept in `out`
  for (int l = t; l < n; l += kRedBlock) {
    const double v = g ? in[l] - step * g[l] / gn : in[l];
    out[l] = clamp01(v);
  }
  __syncthreads();
  double s = 0.0;
  for (int l = t; l < n; l += kRedBlock) s += out[l];
  const double dn = (double)n;
  if (block_sum(s, sh) / dn >= volfrac) return;
  double lo = 0.0, hi = 1.0;
  for (int it = 0; it < 60; ++it) {
    const double mid = (lo + hi) / 2.0;
    double ps = 0.0;
    for (int l = t; l < n; l += kRedBlock) ps += clamp01(out[l] + mid);
    if (block_sum(ps, sh) / dn < volfrac) lo = mid;
    else hi = mid;
  }
  const double shift = (lo + hi) / 2.0;
  __syncthreads();
  for (int l = t; l < n; l += kRedBlock) out[l] = clamp01(out[l] + shift);
}

}  // namespace

void launch_design_filter(hipStream_t s, const double *x, const double *wtab, const double *snorm,
                          const DesignGrid &D, double *y) {
  k_filter<<<nblocks((long long)D.ngy * D.ngx), kBlock, 0, s>>>(x, wtab, snorm, D, y);
}
void launch_design_filter_adjoint(hipStream_t s, const double *g, const double *wtab,
                                  const double *snorm, const DesignGrid &D, double *out) {
  k_filter_adjoint<<<nblocks((long long)D.ngy * D.ngx), kBlock, 0, s>>>(g, wtab, snorm, D, out);
}
void launch_design_alpha(hipStream_t s, const double *rdf, const DesignGrid &D, double *theta) {
  k_alpha<<<nblocks((long long)D.ngy * D.ngx), kBlock, 0, s>>>(rdf, D, theta);
}
void launch_design_chain(hipStream_t s, const double *gtheta, const double *rdf,
                         const DesignGrid &D, double *grdf) {
  k_chain<<<nblocks((long long)D.ngy * D.ngx), kBlock, 0, s>>>(gtheta, rdf, D, grdf);
}
void launch_design_project(hipStream_t s, const double *in, const double *g, double step, int n,
                           double volfrac, double *out, double *gmax) {
  k_project<<<1, kRedBlock, 0, s>>>(in, g, step, n, volfrac, out, gmax);
}

}  // namespace sph
