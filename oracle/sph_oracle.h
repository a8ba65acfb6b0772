/*
 * sph_oracle.h -- CPU restatement of SPHAdjoint's 2D WCSPH path (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, the smoke()
 * check and the cpu_baseline leg of bench.py load it.  It restates, function
 * by function, the reference's Julia/KernelAbstractions code as executed by its
 * CPU backend:
 *
 *   src/kernels.jl   wnorm / w_kern / f_kern / g_kern
 *   src/params.jl    SPHParams, wall_accel, wall_daccel, interp_cell, interp_alpha
 *   src/neighbors.jl CellList, cell_index, _cl_assign!, _cl_fill!, _cl_sort_cells!,
 *                    build!, NeighborList, _nl_build!, build_neighbors!,
 *                    maybe_rebuild!, check_neighbor_flags
 *   src/forward.jl   density_kernel!, eos_kernel!, accel_kernel!, integrate_kernel!, step!
 *   src/adjoint.jl   adj_pass1_kernel!, adj_pass2_kernel!, adj_design_kernel!,
 *                    _axpy2!, _adj_integrate!, _adj_drag!
 *   src/driver.jl    State, Tape, simulate!, AdjointWorkspace, backward!, target_objective
 *
 * Every floating-point expression keeps the reference's operation order and is
 * compiled with -ffp-contract=off (Julia does not contract a*b+c by default), so
 * the FP64 results follow the reference's CPU Float64 path operation for
 * operation.  Two types are built from one source: suffix _f64 (Float64) and
 * _f32 (Float32).
 *
 * Conventions follow the reference: X, V, a are (2,N) column-major, i.e.
 * interleaved x,y per particle; theta is (ngy,ngx) column-major; cell ids,
 * particle ids and `order` are 1-based; `starts` is an exclusive 0-based prefix.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SPHParams (src/params.jl:14-28); given in double and converted to T on use. */
typedef struct orc_params {
  double h, m, rho0, c, mu, gx, gy, dt, Lx, Ly, kw;
  int64_t ngx, ngy;
} orc_params;

enum { ORC_LAYOUT_SLOT = 0, ORC_LAYOUT_PARTICLE = 1 };

/* fields for orc_state_get/set */
enum {
  ORC_X = 0, ORC_V = 1, ORC_A = 2, ORC_RHO = 3, ORC_PTERM = 4, ORC_INVRHO = 5,
  ORC_ALPHA = 6, ORC_XREF = 7
};
/* fields for orc_ws_get */
enum {
  ORC_GX = 0, ORC_GV = 1, ORC_GXS = 2, ORC_GVS = 3, ORC_ABAR = 4, ORC_GRHO = 5,
  ORC_GALPHA = 6, ORC_GTHETA = 7
};

#define ORC_DECL(SFX, REAL)                                                          \
  typedef struct orc_state##SFX orc_state##SFX;                                      \
  typedef struct orc_tape##SFX orc_tape##SFX;                                        \
  typedef struct orc_ws##SFX orc_ws##SFX;                                            \
  orc_state##SFX *orc_state_create##SFX(const orc_params *p, int64_t N, const REAL *X0, \
                                       const REAL *V0, double skin, int64_t interval,  \
                                       int64_t maxnb, int layout);                    \
  void orc_state_destroy##SFX(orc_state##SFX *st);                                    \
  int64_t orc_state_n##SFX(const orc_state##SFX *st);                                 \
  void orc_state_get##SFX(const orc_state##SFX *st, int which, REAL *out);            \
  void orc_state_set##SFX(orc_state##SFX *st, int which, const REAL *in);             \
  void orc_state_info##SFX(const orc_state##SFX *st, int64_t *out8);                  \
  void orc_cell_list_get##SFX(const orc_state##SFX *st, int32_t *cellof, int32_t *counts, \
                              int32_t *starts, int32_t *order);                        \
  void orc_neighbors_get##SFX(const orc_state##SFX *st, int32_t *indices, int32_t *nbcount, \
                              int32_t *flags);                                         \
  void orc_flags_clear##SFX(orc_state##SFX *st);                                      \
  void orc_build##SFX(orc_state##SFX *st);                                            \
  void orc_build_neighbors##SFX(orc_state##SFX *st);                                  \
  void orc_maybe_rebuild##SFX(orc_state##SFX *st);                                    \
  void orc_density##SFX(orc_state##SFX *st);                                          \
  void orc_eos##SFX(orc_state##SFX *st);                                              \
  void orc_accel##SFX(orc_state##SFX *st, const REAL *theta);                         \
  void orc_integrate##SFX(orc_state##SFX *st);                                        \
  void orc_step##SFX(orc_state##SFX *st, const REAL *theta);                          \
  orc_tape##SFX *orc_tape_create##SFX(int64_t N, int64_t capacity);                   \
  void orc_tape_destroy##SFX(orc_tape##SFX *tp);                                      \
  void orc_tape_reset##SFX(orc_tape##SFX *tp);                                        \
  int64_t orc_tape_length##SFX(const orc_tape##SFX *tp);                              \
  int64_t orc_tape_capacity##SFX(const orc_tape##SFX *tp);                            \
  void orc_tape_get##SFX(const orc_tape##SFX *tp, int64_t k, REAL *X, REAL *V);       \
  int orc_simulate##SFX(orc_state##SFX *st, const REAL *theta, int64_t nsteps,        \
                        orc_tape##SFX *tape);                                          \
  orc_ws##SFX *orc_ws_create##SFX(int64_t N, const orc_params *p);                    \
  void orc_ws_destroy##SFX(orc_ws##SFX *ws);                                          \
  void orc_ws_get##SFX(const orc_ws##SFX *ws, int which, REAL *out);                  \
  int orc_backward##SFX(orc_ws##SFX *ws, const orc_tape##SFX *tape, const REAL *theta, \
                        const orc_params *p, const REAL *seedX, const REAL *seedV,     \
                        double skin, int64_t interval, int64_t maxnb, int layout,      \
                        int32_t *flags_out);                                           \
  REAL orc_target_objective##SFX(const REAL *X, int64_t N, double xc, double yc,      \
                                 double sigma, REAL *seed);                            \
  REAL orc_w_kern##SFX(REAL q, REAL A);                                               \
  REAL orc_f_kern##SFX(REAL q, REAL A, REAL h);                                       \
  REAL orc_g_kern##SFX(REAL q, REAL A, REAL h);                                       \
  REAL orc_wnorm##SFX(REAL h);                                                        \
  void orc_interp_alpha##SFX(const REAL *theta, const orc_params *p, REAL x, REAL y,  \
                             REAL *out3);

ORC_DECL(_f64, double)
ORC_DECL(_f32, float)

/* OpenMP thread count for the per-particle loops (0 = leave the default). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
