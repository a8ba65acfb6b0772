/*
 * sphadj.h -- C ABI of libsphadj.so, the MI355X (gfx950) implementation of
 * SPHAdjoint's differentiable 2D WCSPH step.
 *
 * The ABI mirrors the reference's public Julia API (src/SPHAdjoint.jl:39-40):
 * a Julia `ccall` wrapper (or Python ctypes) keeps the reference's names and
 * argument meaning and calls these entry points instead of launching
 * KernelAbstractions kernels.  Every entry point returns an sph_status; the
 * message of the last failure on the calling thread is sph_last_error().
 *
 * Conventions (the reference's):
 *   - X, V, a, seeds, gradients: (2, N) column-major doubles, i.e. interleaved
 *     x,y per particle, in particle-id order.
 *   - theta / gtheta: (ngy, ngx) column-major doubles.
 *   - cell ids, particle ids and `order` are 1-based; `starts` is the exclusive
 *     0-based prefix of `counts`.
 * Device memory is owned by the library behind opaque handles; host pointers
 * are borrowed for the duration of a call.  Calls are asynchronous on the
 * context's stream except where they return host data (get/objective/flags)
 * or where the reference synchronises (simulate, backward).
 */
#ifndef SPHADJ_H
#define SPHADJ_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sph_status {
  SPH_OK = 0,
  SPH_E_ARG = 1,       /* invalid argument (reference: ArgumentError on layout etc.) */
  SPH_E_CAPACITY = 2,  /* tape too small (reference: ArgumentError, src/driver.jl:96-98) */
  SPH_E_HIP = 3,       /* HIP runtime failure */
  SPH_E_STATE = 4,     /* call not valid in the current state */
  SPH_E_NOMEM = 5      /* device allocation failed */
} sph_status;

/* Arithmetic form of a state / tape / workspace. */
typedef enum sph_precision {
  SPH_F64 = 0,    /* plain FP64: the reference's Float64 arithmetic, operation for operation */
  SPH_F32C = 1    /* FP32 state and gathers with compensated sums; FP64 x += dt*v, J, gtheta */
} sph_precision;

typedef enum sph_layout { SPH_LAYOUT_SLOT = 0, SPH_LAYOUT_PARTICLE = 1 } sph_layout;

typedef enum sph_field_id {
  SPH_X = 0, SPH_V = 1, SPH_A = 2, SPH_RHO = 3, SPH_PTERM = 4, SPH_INVRHO = 5, SPH_ALPHA = 6
} sph_field_id;

typedef enum sph_grad_id { SPH_GX = 0, SPH_GV = 1, SPH_GTHETA = 2 } sph_grad_id;
/* what a tape keeps per step beyond X and V (sph_tape_create_keep) */
typedef enum sph_tape_keep { SPH_TAPE_KEEP_PI = 1, SPH_TAPE_KEEP_LISTS = 2 } sph_tape_keep;

/* SPHParams{T} (src/params.jl:14-28); gx defaults to 0 and gy to -9.81 there. */
typedef struct sph_params {
  double h, m, rho0, c, mu, gx, gy, dt, Lx, Ly, kw;
  int64_t ngx, ngy;
} sph_params;

typedef struct sph_ctx sph_ctx;
typedef struct sph_field sph_field;
typedef struct sph_state sph_state;
typedef struct sph_tape sph_tape;
typedef struct sph_workspace sph_workspace;

/* --- context: replaces the KernelAbstractions `backend` object (src/driver.jl:24) --- */
const char *sph_last_error(void);
int sph_ctx_create(int device, sph_ctx **out);
/* The same on a stream the caller owns (a hipStream_t passed as void*, e.g.
 * the HIP queue of a Julia AMDGPU task): every call of the context issues into
 * that stream, in the caller's order; the stream must belong to `device` and
 * outlive the context, which never destroys it.  The reference's kernels
 * launch on the backend's own queue (src/forward.jl:166-171). */
int sph_ctx_create_stream(int device, void *stream, sph_ctx **out);
int sph_ctx_destroy(sph_ctx *ctx);
/* free and total device memory of the context's device (hipMemGetInfo); no
 * counterpart in the reference */
int sph_mem_info(sph_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes);
/* KernelAbstractions.synchronize(backend) (src/driver.jl:110,204) */
int sph_sync(sph_ctx *ctx);

/* --- design field theta, a device array of shape (ngy, ngx) (src/params.jl:70-82) --- */
int sph_field_create(sph_ctx *ctx, int64_t ngy, int64_t ngx, const double *host, sph_field **out);
int sph_field_set(sph_field *f, const double *host);
int sph_field_get(const sph_field *f, double *host);
int sph_field_destroy(sph_field *f);

/* Device memory owned by the library, for the *_dev entry points' inputs and
 * outputs (the counterpart of a KernelAbstractions.allocate'd array).  No
 * counterpart entry point in the reference.  host == NULL zero-fills. */
typedef struct sph_buffer sph_buffer;
int sph_buffer_create(sph_ctx *ctx, int64_t nbytes, const void *host, sph_buffer **out);
int sph_buffer_destroy(sph_buffer *b);
int sph_buffer_ptr(const sph_buffer *b, void **ptr);   /* the device address */
int sph_buffer_read(const sph_buffer *b, void *host, int64_t nbytes);
int sph_buffer_write(sph_buffer *b, const void *host, int64_t nbytes);

/* --- State (src/driver.jl:5-36) with its CellList and NeighborList (src/neighbors.jl:28-36,174-190) --- */
int sph_state_create(sph_ctx *ctx, const sph_params *p, int precision, int64_t N,
                     const double *X0, const double *V0, double skin, int64_t interval,
                     int64_t maxnb, int layout, sph_state **out);
int sph_state_destroy(sph_state *st);
/* st.X / st.V / st.a / st.rho / st.pterm / st.invrho / st.alpha, in particle-id order */
int sph_state_get(sph_state *st, int field, double *host);
/* copyto!(st.X, ...) / copyto!(st.V, ...) */
int sph_state_set(sph_state *st, int field, const double *host);
/* out8 = {nx, ny, ncell, maxnb, sm, si, interval, age} */
int sph_state_info(const sph_state *st, int64_t *out8);

/* build!(cl, X, p, backend) (src/neighbors.jl:97-114) */
int sph_build(sph_state *st);
/* cl.cellof, cl.counts, cl.starts, cl.order of the last build (any pointer may be NULL) */
int sph_cell_list_get(sph_state *st, int32_t *cellof, int32_t *counts, int32_t *starts,
                      int32_t *order);
/* build_neighbors!(nl, cl, X, p, backend) (src/neighbors.jl:252-263) */
int sph_build_neighbors(sph_state *st);
/* nl.indices (N*maxnb, with the state's (sm, si) strides) and nl.counts of the last build */
int sph_neighbors_get(sph_state *st, int32_t *indices, int32_t *counts);
/* nl.flags: [0] row overflow, [1] displacement beyond skin*h (src/neighbors.jl:282-290) */
int sph_flags_get(sph_state *st, int32_t *flags2);
/* the same and [2]: a position that was not finite at a rebuild (NaN or
 * infinite coordinates bin into a clamped edge cell, as the reference's
 * unsafe_trunc does, src/neighbors.jl:38-42); no counterpart in the reference */
int sph_flags_get_all(sph_state *st, int32_t *flags3);
/* resets all three flags */
int sph_flags_clear(sph_state *st);

/* step!(st, theta, p, backend) (src/forward.jl:161-173) */
int sph_step(sph_state *st, const sph_field *theta);
/* simulate!(st, theta, p, backend, nsteps; tape) (src/driver.jl:92-115); tape may be NULL */
int sph_simulate(sph_state *st, const sph_field *theta, int64_t nsteps, sph_tape *tape);
/* simulate! without a tape, replaying a captured HIP graph of two rebuild
 * epochs (2*interval steps); results equal sph_simulate's bit for bit.  No
 * counterpart in the reference (whose steps are individual kernel launches,
 * src/forward.jl:161-173); for forward-only evaluations such as line-search
 * candidates.  *replays (may be NULL) = graph launches made. */
int sph_simulate_graph(sph_state *st, const sph_field *theta, int64_t nsteps, int64_t *replays);

/* --- Tape (src/driver.jl:45-83) --- */
int sph_tape_create(sph_ctx *ctx, int64_t N, int64_t capacity, int precision, sph_tape **out);
/* A tape that also keeps each step's (pterm, invrho) from the density pass
 * (keep & SPH_TAPE_KEEP_PI) and, in the f32c form, the pass's exact neighbour
 * masks (keep & SPH_TAPE_KEEP_LISTS), so that backward! skips the density
 * recompute of its R pass and the pair tests of P1.  No counterpart in the
 * reference, whose tape keeps X and V only (src/driver.jl:45-83).  Extra
 * device memory per particle-step: 16 B (FP64) or 8 B (f32c) for (pterm,
 * invrho), 48 B for the masks.  Gradients equal a plain tape's bit for bit. */
int sph_tape_create_keep(sph_ctx *ctx, int64_t N, int64_t capacity, int precision, int keep,
                         sph_tape **out);
int sph_tape_destroy(sph_tape *tp);
int sph_tape_reset(sph_tape *tp);                            /* reset!(tape) */
int sph_tape_length(const sph_tape *tp, int64_t *n);         /* length(tape) */
int sph_tape_capacity(const sph_tape *tp, int64_t *cap);     /* capacity(tape) */
/* tape.X[:, :, k], tape.V[:, :, k] (1-based k), particle-id order */
int sph_tape_get(sph_tape *tp, int64_t k, double *X, double *V);

/* --- AdjointWorkspace (src/driver.jl:117-133) and backward! (src/driver.jl:141-207) --- */
int sph_workspace_create(sph_ctx *ctx, int64_t N, const sph_params *p, int precision,
                         sph_workspace **out);
int sph_workspace_destroy(sph_workspace *ws);
/* ws.gX / ws.gV ((2,N), id order) and ws.gtheta ((ngy,ngx)) */
int sph_workspace_get(sph_workspace *ws, int which, double *host);
int sph_backward(sph_workspace *ws, sph_tape *tape, const sph_field *theta,
                 const double *seedX, const double *seedV, double skin, int64_t interval,
                 int64_t maxnb, int layout);

/* target_objective(st.X, xc, yc, sigma) (src/driver.jl:219-234): J in FP64 and,
 * when seed != NULL, dJ/dX_T as (2,N) in id order. */
int sph_target_objective(sph_state *st, double xc, double yc, double sigma, double *J,
                         double *seed);

/* Device-memory variants of the calls above: seedX, seedV and seed are device
 * pointers ((2,N) interleaved doubles, id order) on the context's device, e.g.
 * from sph_buffer_ptr.  The reference's backward! accepts device seeds the same
 * way (copyto!, src/driver.jl:144-146); callers order their own writes of
 * these buffers before the call.  J stays a host value. */
int sph_backward_dev(sph_workspace *ws, sph_tape *tape, const sph_field *theta,
                     const double *seedX, const double *seedV, double skin, int64_t interval,
                     int64_t maxnb, int layout);
int sph_target_objective_dev(sph_state *st, double xc, double yc, double sigma, double *J,
                             double *seed);
/* sph_target_objective_dev for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  It allocates nothing and waits for nothing:
 * the objective's partial sums go to a scratch array of the library's
 * (static device memory, one copy per GPU, held until the process ends; at
 * most 16384 partials, so N <= 4,194,304 -- larger states are refused with
 * SPH_E_ARG), and J (one double) and seed ((2,N) doubles, id order, may be
 * NULL) are device pointers the caller owns, written in stream order.  The
 * results equal sph_target_objective_dev's bit for bit.  Calls that overlap
 * on one GPU share its scratch array and race. */
int sph_target_objective_noalloc(sph_state *st, double xc, double yc, double sigma, double *J,
                                 double *seed);

/* --- batched scenes (the FP32-compensated form).  No counterpart in the
 * reference, which runs one scene per State; its line search evaluates the
 * design candidates one after another (scripts/03_optimize.jl:314-341).  A
 * batched state holds nscenes independent scenes side by side (each with its
 * own params, hence its own resolution, N and design grid); every kernel
 * launch covers all of them, each workgroup working inside one scene.
 * X0, V0: the scenes' particles in order, (2, sum N) interleaved.  The design
 * field of a batch is one sph_field holding the scenes' (ngy, ngx)
 * column-major fields one after the other (e.g. created with ngy = total
 * nodes, ngx = 1).  sph_step, sph_simulate, sph_tape_*, sph_backward and
 * sph_state_get/set work on batched states as on single ones; gradients and
 * gtheta come back concatenated in scene order. --- */
int sph_batch_create(sph_ctx *ctx, int64_t nscenes, const sph_params *params, const int64_t *N,
                     const double *X0, const double *V0, double skin, int64_t interval,
                     int64_t maxnb, sph_state **out);
int sph_batch_workspace_create(sph_ctx *ctx, const sph_state *st, sph_workspace **out);
/* *nscenes; when out != NULL, out[4k..4k+3] = {first particle, particle count,
 * first design node, design nodes} of scene k */
int sph_batch_info(const sph_state *st, int64_t *nscenes, int64_t *out);
/* target_objective per scene: J[nscenes] (host), seed (2, sum N) in id order
 * (host for sph_batch_objective, device for the _dev variant), or NULL */
int sph_batch_objective(sph_state *st, double xc, double yc, double sigma, double *J,
                        double *seed);
int sph_batch_objective_dev(sph_state *st, double xc, double yc, double sigma, double *J,
                            double *seed);

/* Autotuner of the f32c form (no counterpart in the reference): picks the
 * launch configuration of the neighbour-list build, density, force, adjoint
 * R, P1 and P2 kernels -- the workgroup size (64, 128 or 256) and whether the
 * pass stages its window's neighbour rows in LDS or reads them from global
 * memory -- by timing each of the six candidates over `steps` steps with a
 * tape that keeps (pterm, invrho) and the exact masks (the kind the
 * checkpointed and batched paths record) and their reverse sweep, and keeps
 * the choice with the state (tapes recorded afterwards inherit it).  At 256
 * threads P1 also runs the step's gtheta gather in its launch, so P1 is
 * compared on P1 + gather.  Advances st by 12*steps steps.  chosen (may be
 * NULL) receives the 6 configurations in that order: the workgroup size,
 * negated when the kernel reads without LDS staging.  Every configuration
 * gives the same results bit for bit. */
int sph_autotune(sph_state *st, const sph_field *theta, int64_t steps, int32_t *chosen);
/* the tuned configurations of a state (6 entries, as sph_autotune reports
 * them), and their transfer to another state of the same kind */
int sph_tuning_get(const sph_state *st, int32_t *sizes);
int sph_tuning_set(sph_state *st, const int32_t *sizes);

/* Candidate sweep of the f32c form's neighbour-list builds (no counterpart in
 * the reference, whose _nl_build! sweeps the 3x3 cells of width 2h(1+skin),
 * src/neighbors.jl:192-245; its README lists h-wide cells with a 5x5 sweep as
 * a possible change, README.md:429,795-796).  sweep 5: from the next rebuild
 * on, each cell's slots are ordered by quadrant (half cell in x, then in y),
 * then id, and the mask list is built from the 5x5 half cells around each
 * slot's quadrant -- about 52 candidates instead of 75, the same list (every
 * pair closer than the cell width lies in the half-cell block).  The pair
 * passes walk the same masks; sums within a cell follow the quadrant order, so
 * the f32c results change at rounding level.  sph_cell_list_get and
 * sph_neighbors_get still report the reference's order.  Checkpoint
 * snapshots keep the order of the layout they hold.  One-scene f32c states
 * with N < 2^29; sweep 3 (the default) restores the (cell, id) order at the
 * next rebuild. */
int sph_state_set_sweep(sph_state *st, int sweep);

/* --- binomial checkpointing of the time loop.  No counterpart in the
 * reference, whose Tape keeps every step (src/driver.jl:38-65) and whose README
 * recommends recompute checkpointing for long horizons.  The horizon is cut
 * into segments of `seg_epochs` neighbour-list epochs (seg_epochs*interval
 * steps); at most `nsnap` segment-start states are kept, at the positions of
 * the binomial (revolve) schedule, and each segment is recomputed onto a
 * one-segment tape right before its reverse sweep.  Segment starts are
 * neighbour-list rebuilds, so gradients equal the full-tape ones bit for bit.
 * Device memory: nsnap+1 snapshots (X, V, ids; the f32c form also keeps the
 * rebuilt layout: binning cells, cell starts, thread order, so that a
 * restored segment needs no cell-list build), one work state and a tape of
 * seg_epochs*interval steps. --- */
typedef struct sph_ckpt sph_ckpt;
int sph_ckpt_create(sph_ctx *ctx, const sph_params *p, int precision, int64_t N, int64_t nsteps,
                    int64_t nsnap, int64_t seg_epochs, double skin, int64_t interval,
                    int64_t maxnb, sph_ckpt **out);
int sph_ckpt_destroy(sph_ckpt *ck);
/* simulate! over nsteps from X0, V0 ((2,N), id order); the final state is sph_ckpt_state's */
int sph_ckpt_forward(sph_ckpt *ck, const double *X0, const double *V0, const sph_field *theta);
/* the same with X0, V0 in device memory */
int sph_ckpt_forward_dev(sph_ckpt *ck, const double *X0, const double *V0,
                         const sph_field *theta);
/* the checkpointer's state (borrowed; its X/V are the final state after sph_ckpt_forward) */
int sph_ckpt_state(sph_ckpt *ck, sph_state **st);
/* backward! from seeds at the final state; fills ws like sph_backward */
int sph_ckpt_backward(sph_ckpt *ck, sph_workspace *ws, const sph_field *theta,
                      const double *seedX, const double *seedV);
/* the same with seeds in device memory */
int sph_ckpt_backward_dev(sph_ckpt *ck, sph_workspace *ws, const sph_field *theta,
                          const double *seedX, const double *seedV);
/* Tape budget of a checkpointer (no counterpart in the reference): let its
 * segment tape hold up to tape_steps steps (whole segments, at least one).
 * The last segments of the horizon that fit, after the one the binomial
 * schedule tapes first, are then recorded in the forward and never
 * recomputed; the schedule covers the segments before them with the same
 * snapshot budget.  Gradients are the same bits as without the budget.
 * Device memory of the tape per particle-step: f32c 80 B (X, V, pterm/invrho,
 * exact masks), FP64 48 B, plus one layout record per epoch.  The default is
 * one segment.  Reallocates the tape; a forward run before it needs a new
 * forward. */
int sph_ckpt_set_tape_steps(sph_ckpt *ck, int64_t tape_steps);
/* {segments, snapshots, segments run by the last forward, segments run by the
 *  last backward (recomputation included), segments taped in the last backward} */
int sph_ckpt_stats(const sph_ckpt *ck, int64_t *out5);
/* the schedule alone, on the host: {segments run in total, segments taped,
 * most snapshots alive at once} for `epochs` segments and `nsnap` snapshots */
int sph_ckpt_plan(int64_t epochs, int64_t nsnap, int64_t *out3);

/* --- design-field chain of the topology optimiser, on the device in FP64.
 * The reference keeps it in its optimisation script (scripts/03_optimize.jl:142-244):
 *   rd -> apply_filter (linear weights (r+1) - dist, radius r grid points,
 *         normalised per node, clamped to [0,1]) -> rdf -> alpha_of(rdf) =
 *         alpha_max*qp*(1-rdf)/(qp+rdf) -> theta
 *   gtheta -> .* dalpha_of(rdf) -> filter_adjoint -> grd
 *   project_volume(x, volfrac) (bisection, scripts/03_optimize.jl:205-219) and
 *   the line-search candidate project_volume(rd - step*g/max|g|, volfrac)
 *   (scripts/03_optimize.jl:311).
 * Fields are sph_field handles of (ngy, ngx) column-major doubles. */
typedef struct sph_design sph_design;
/* build_filter (scripts/03_optimize.jl:156-164) for an (ngy, ngx) grid */
int sph_design_create(sph_ctx *ctx, int64_t ngy, int64_t ngx, int64_t radius, double alpha_max,
                      double qp, sph_design **out);
int sph_design_destroy(sph_design *d);
/* rdf = apply_filter(rd); theta[offset + node] = alpha_of(rdf[node]).  rdf is
 * (ngy, ngx); theta holds at least offset + ngy*ngx nodes (offset > 0 writes
 * one scene's field of a batched design field). */
int sph_design_theta(sph_design *d, const sph_field *rd, sph_field *rdf, sph_field *theta,
                     int64_t offset);
/* grd = filter_adjoint(gtheta .* dalpha_of(rdf)) with gtheta the nodes
 * [offset, offset + ngy*ngx) of the workspace's dJ/dtheta (after sph_backward) */
int sph_design_grad(sph_design *d, const sph_field *rdf, const sph_workspace *ws, int64_t offset,
                    sph_field *grd);
/* out = project_volume(x, volfrac); out and x are distinct fields */
int sph_design_project(sph_design *d, const sph_field *x, double volfrac, sph_field *out);
/* out = project_volume(rd - step * g / max|g|, volfrac); *gmax (host, may be
 * NULL) receives max|g|; out is distinct from rd and g */
int sph_design_candidate(sph_design *d, const sph_field *rd, const sph_field *g, double step,
                         double volfrac, sph_field *out, double *gmax);

/* --- measurement: per-kernel GPU time with events on the context's stream.
 * No counterpart in the reference (its stage timing lives in
 * scripts/06_metal_profile.jl:158-290); used by bench.py for the roofline. */
/* on = 0: off; 1: every kernel; 2 + k: kernel id k only (two events per
 * launch of that kernel, none elsewhere).  Also resets the counters. */
int sph_profile_enable(sph_ctx *ctx, int on);
int sph_profile_get(sph_ctx *ctx, int kernel, int64_t *launches, double *total_ms);
const char *sph_profile_name(int kernel);          /* NULL past the last kernel id */

#ifdef __cplusplus
}
#endif
#endif
