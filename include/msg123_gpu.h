/* msg123_gpu.h -- C ABI of the MI355X Newton-Krylov step for the 3D
 * mixed-form Richards equation (libmsg123_gpu.so).
 *
 * The reference solver has no FFI: its hot path is a set of Fortran module
 * procedures sharing module state.  Each entry point below replaces one of
 * them; a Fortran shim binds them with iso_c_binding and keeps the
 * reference's public names (see the comment on each function).
 *
 * Conventions: plain pointers and sizes, host memory on the boundary (the
 * library copies to and from device memory), int return codes (0 = ok,
 * negative = error), one context per domain, not thread-safe per context.
 * Arrays are 0-based C arrays holding the reference's 1-based Fortran
 * arrays element for element; CSR column ids stay 1-based as in
 * crs_index(1)%offrow.
 */
#ifndef MSG123_GPU_H
#define MSG123_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSG123_OK 0
#define MSG123_EINVAL (-1)     /* bad argument or unsupported mesh */
#define MSG123_EDEVICE (-2)    /* HIP runtime error */
#define MSG123_ENAN (-3)       /* NaN detected in the Newton update */
#define MSG123_ENOMEM (-4)     /* device allocation failed */
#define MSG123_EIO (-5)        /* a file could not be opened, written or read */
#define MSG123_ECONV (-6)      /* msg123_run_box: the failure budget of the call is spent */

/* Static arrays of the reference's level-1 system.  Replaces the module
 * state built by set_connect / make_cell_info / assign_* and copied into
 * allocate_solution by allocate_matvec (Allocate_solution.f90:95-197).
 * The mesh must be a full structured box of nx*ny*nz active cells in the
 * reference's natural order (Set_cell.f90:880-908) with the 7-point CSR
 * pattern in neighbour order up, N, W, E, S, down. */
typedef struct msg123_mesh {
  int32_t nx, ny, nz;
  int32_t ncalc;                 /* nx*ny*nz */
  int32_t ncals;                 /* nx*ny, the top layer */
  int32_t nconn;                 /* crs_index(1)%offind(ncalc) */
  const int32_t *offind;         /* [ncalc+1] crs_index(1)%offind(0:ncalc) */
  const int32_t *offrow;         /* [nconn]   crs_index(1)%offrow, 1-based */
  const int32_t *dir_conn;       /* [nconn]   dir_conn, 1..6 */
  const double *hydf_conn;       /* [nconn]   hydf_conn */
  const double *abyd_conn;       /* [nconn]   abyd_conn */
  const double *cell_vol;        /* [ncalc]   make_cell::cell_vol */
  const double *cell_top;        /* [ncalc]   make_cell::cell_top */
  const float *reta, *retn, *resi, *poro, *ss; /* [ncalc] assign_calc::read_* */
  const double *hydf_surf, *abyd_surf;         /* [ncals] set_condition */
  const double *surf_top, *surf_bott;          /* [ncals] prep_calculation / assign_calc */
  double len_scal;               /* read_input::len_scal */
  double newper;                 /* initial_module::newper */
} msg123_mesh;

/* Solver settings (initial_module: maxout_iter, maxinn_iter, criteria). */
#define MSG123_PRECON_DILU 0     /* reference D-ILU(0), FP64, bit-exact sweeps */
#define MSG123_PRECON_LINE32 1   /* FP32 vertical-line (tridiagonal) blocks, FP64 BiCGStab */
/* With MSG123_PRECON_LINE32, the Newton systems msg123_calc_numsol and
 * msg123_run_box assemble keep their seven diagonals in FP32 (the products
 * and the solve's vectors stay FP64; the difference-quotient Jacobian holds
 * about eight digits); systems assembled by msg123_make_matvec or loaded by
 * msg123_load_system stay FP64. */
/* The reference's precon_type = 1 path.  Its serial V-cycle returns the
 * D-ILU result (the post-smoother starts from the rhs again,
 * Linear_solution.f90:688-718,820), so the Krylov numerics are D-ILU's; the
 * path also leaves the D-ILU diagonal and the last preconditioner input in
 * the level-1 system (:517-529), which the line search's slope
 * (Nonlinear_solution.f90:500) and the convergence log (:236-237) read.
 * Both side effects are reproduced. */
#define MSG123_PRECON_AMG 2
/* Smoothed-aggregation AMG on the device as an FP32 preconditioner inside
 * the FP64 BiCGStab: the hierarchy is aggregated once on the host (the
 * reference's strength filter and heap aggregation, Make_amg_matrix.f90:
 * 67-222), its prolongators and Galerkin products R A P are recomputed on
 * the device from every new Jacobian, and one V-cycle with damped-Jacobi
 * smoothing (msg123_amg_opts) is applied per preconditioner call.  No
 * reference counterpart: the reference's serial V-cycle discards its coarse
 * correction (Linear_solution.f90:688-718,820). */
#define MSG123_PRECON_AMG32 3
/* FP32 multicolour ILU(0) inside the FP64 BiCGStab: red-black colouring of
 * the 7-point grid, pivots d_r = a_rr and d_b = a_bb - sum_k a_bk a_kb / a_kk
 * computed in FP64, factors stored and applied in FP32, each sweep one pass
 * per colour.  It replaces precon_dilu / solve_ilu (Linear_solution.f90:
 * 742-855), whose natural-order sweeps are serial, with the same ILU(0) on a
 * parallel ordering; the reference has no counterpart. msg123_precon_dilu
 * returns the FP64 pivots. */
#define MSG123_PRECON_MCILU32 4
typedef struct msg123_params {
  int32_t maxout_iter;           /* Newton iteration cap (default 20) */
  int32_t maxinn_iter;           /* BiCGStab iteration cap (default 10) */
  double criteria;               /* max-norm head change tolerance, metres */
  int32_t precon;                /* MSG123_PRECON_* */
  /* 0: the reference's Jacobian, which adds -K A/d kr to every top-cell
   *    diagonal even where the seepage face is inactive
   *    (Make_linearsystem.f90:476-478);
   * 1: that term only where the face is active (h >= surf_bott), i.e. the
   *    derivative of the residual.  The residual, and so the converged
   *    heads, are unchanged; Newton converges in fewer iterations. */
  int32_t consistent_jacobian;
} msg123_params;

/* Options of MSG123_PRECON_AMG32 (msg123_set_amg_opts). */
typedef struct msg123_amg_opts {
  int32_t max_levels;            /* levels including the box (amg_nlevel); default 12 */
  float theta;                   /* strength threshold (amg_theta); default 0.05 */
  float omega;                   /* prolongator smoothing weight (jac_omega); default 0.67 */
  float smooth_omega;            /* damped-Jacobi weight of the smoother; default 0.67 */
  int32_t npre, npost;           /* sweeps before / after the coarse correction; 2, 2 */
  int32_t coarse_max;            /* stop coarsening at or below this size; default 64 */
  int32_t coarse_sweeps;         /* Jacobi sweeps on a coarsest level too large for the
                                    dense solve (> 80 unknowns); default 20 */
  /* 0: the reference's aggregates exactly (its orphan fix-up stops after the
   *    first node, Make_amg_matrix.f90:173-186, leaving nodes outside every
   *    aggregate);
   * 1: every such node with an aggregated strong neighbour joins the first
   *    one's aggregate (default; the V-cycle then contracts). */
  int32_t complete_orphans;
  /* 1 (default): build and smooth on the M-matrix projection of the
   *    Jacobian (couplings clipped at zero, diagonal -max(|a_ii|, sum of the
   *    clipped couplings)): the V-cycle then contracts even where a wetting
   *    front gives the Jacobian couplings or a diagonal of the wrong sign,
   *    and BiCGStab on the true Jacobian absorbs the few rows that differ;
   * 0: use the Jacobian as it is. */
  int32_t project;
} msg123_amg_opts;

/* Per-call statistics of msg123_calc_numsol. */
typedef struct msg123_newton_stats {
  int32_t out_iter;              /* Newton iterations (prep_calculation::out_iter) */
  int32_t conv_flag;             /* 1 converged (prep_calculation::conv_flag) */
  int32_t in_iter_last;          /* BiCGStab iterations of the last solve (linear_solution::in_iter) */
  int32_t in_iter_total;         /* summed over the Newton iterations */
  int32_t back_iter_total;       /* backtracking function evaluations */
  int32_t func_evals;            /* residual evaluations */
  double max_change;             /* last max |dh| * len_scal, metres */
  double l2_new;                 /* last ||F||^2 */
} msg123_newton_stats;

typedef struct msg123_ctx msg123_ctx;

/* Create a device context for one domain.  Replaces allocate_solvar
 * (Allocate_solution.f90:56-93) for the level-1 system. */
int msg123_ctx_create(const msg123_mesh *mesh, const msg123_params *params,
                      msg123_ctx **out);
void msg123_ctx_destroy(msg123_ctx *ctx);

/* Load the time-level state read by calc_numsol: head_new, head_old,
 * srat_old [ncalc]; surf_head, surf_rati, calc_rech (recharge volume per top
 * cell) [ncals]; delt.  Replaces the allocate_solution / calc_boundary /
 * prep_calculation module variables (Time_module.f90:34-112 writes them). */
int msg123_set_state(msg123_ctx *ctx, const double *head_new, const double *head_old,
                     const double *srat_old, const double *surf_head,
                     const double *surf_rati, const double *calc_rech, double delt);
/* Read back head_new, srat_new, rel_perm [ncalc] and surf_head, surf_rati
 * [ncals]; any pointer may be NULL. */
int msg123_get_state(msg123_ctx *ctx, double *head_new, double *srat_new, double *rel_perm,
                     double *surf_head, double *surf_rati);

/* Saturation and relative permeability: calc_srat_rperm
 * (Calc_parameter.f90:15-77).  pres, srat, rperm, sstor [ncalc]; sstor may
 * be NULL. */
int msg123_calc_srat_rperm(msg123_ctx *ctx, double pertur, const double *pres, double *srat,
                           double *rperm, double *sstor);

/* Residual F(h): calc_func (Calc_function.f90:24-101).  Also updates the
 * context's srat_new, rel_perm and surf_rati, as the reference does. */
int msg123_calc_func(msg123_ctx *ctx, const double *h, double *funcv);

/* J.v without the Jacobian: calc_vecjacf (Calc_function.f90:515-621) at x
 * along v, jv = -(F(x + eps v) - F(x))/eps with Brown and Saad's eps;
 * x, v, jv [ncalc].  Leaves the context's srat_new, rel_perm and surf_rati
 * of F(x + eps v), as the reference's calc_func side effects do. */
int msg123_calc_vecjacf(msg123_ctx *ctx, const double *x, const double *v, double *jv);

/* The reference's output and restart files, stream-unformatted (raw
 * little-endian values without record markers, open_new_wbin,
 * Utility_module.f90:174-175).
 * msg123_write_head: one record of the head file, write_out_headf
 * (Write_output.f90:336-362): a float32 time (write_header_bin,
 * Write_module.f90:209-225) and nx*ny*nz float32 values
 * real(head_new*len_scal, SP), x fastest (write_3dbin, :146-183).  append = 0
 * replaces the file, 1 adds a record.
 * msg123_write_restart: write_out_restf (Write_output.f90:415-432), the
 * FP64 time and head_new*len_scal in FP64; the file is replaced.
 * msg123_read_restart: the restart branch of assign_initv
 * (Assign_calculation.f90:416-435): heads[ncalc] = value * len_scal_inv
 * with len_scal_inv = 1/len_scal in FP32 (Read_input.f90:1108), *time =
 * real(header, SP); host only, no context.  Load the heads with
 * msg123_set_state. */
int msg123_write_head(msg123_ctx *ctx, const char *path, int32_t append, float time);
int msg123_write_restart(msg123_ctx *ctx, const char *path, float time);
int msg123_read_restart(const char *path, int32_t ncalc, double len_scal, double *heads,
                        float *time);

/* rhs = -F(head_new) and the FD Jacobian: make_matvec
 * (Make_linearsystem.f90:24-68).  The system stays on the device for
 * msg123_solve_linalg; rhs, dmat [ncalc] and lumat [nconn] (CSR order) are
 * copied out when not NULL. */
int msg123_make_matvec(msg123_ctx *ctx, double *rhs, double *dmat, double *lumat);

/* Solve J dh = rhs for the system of the last make_matvec (or of
 * msg123_load_system): solve_linalg (Linear_solution.f90:26-58) with
 * bnorm = init_norm and the convergence tolerance errtol (check_insol,
 * Check_simulation.f90:16-35).  dh [ncalc] is the initial guess on entry
 * (the reference passes zero) and the solution on exit. */
int msg123_solve_linalg(msg123_ctx *ctx, double init_norm, double errtol, double *dh,
                        double *last_norm, int32_t *in_iter, int32_t *conv_flag);

/* Load a system given in the reference's CSR layout (dmat, lumat, rhs). */
int msg123_load_system(msg123_ctx *ctx, const double *rhs, const double *dmat,
                       const double *lumat);

/* Preconditioner pieces, for parity testing: D-ILU diagonal
 * (precon_dilu, Linear_solution.f90:742-782) of the loaded system into d,
 * and z = M^-1 r (solve_ilu, :803-855) with that diagonal.  With
 * MSG123_PRECON_LINE32 the same two calls set up and apply the FP32
 * vertical-line preconditioner instead (d is left untouched; nz <= 640). */
int msg123_precon_dilu(msg123_ctx *ctx, double *d);
int msg123_solve_ilu(msg123_ctx *ctx, const double *r, double *z);
/* y = A x for the loaded system (calc_matvec, :917-951). */
int msg123_matvec(msg123_ctx *ctx, const double *x, double *y);

/* One implicit step attempt: calc_numsol (Nonlinear_solution.f90:33-294)
 * from the loaded state.  On return head_new (and srat_new, surf_head) in
 * the context hold the result; stats may be NULL. */
int msg123_calc_numsol(msg123_ctx *ctx, msg123_newton_stats *stats);

/* ---- Standalone box driver (no reference set-up needed) ------------- */

/* Inputs of a uniform box in SI units (grid type 7, closed box, constant
 * top recharge, init_type 7 initial head). */
typedef struct msg123_box_spec {
  int32_t nx, ny, nz;
  float dx, dy, dz;
  const float *alpha, *vg_n, *theta_r, *theta_s, *ks, *ss; /* [nx*ny*nz] */
  float wt_depth;                /* metres below the top surface */
  float recharge_mmh;            /* mm/h */
} msg123_box_spec;

typedef struct msg123_box_mesh {
  msg123_mesh mesh;              /* arrays owned by this struct */
  double *rech_area;             /* [ncals] set_condition::rech_area */
  float *read_rech;              /* [ncals] assign_boundary::read_rech (scaled) */
  double *h0;                    /* [ncalc] initial head (scaled) */
} msg123_box_mesh;

/* Build the reference's static arrays for a box (host only). */
int msg123_mesh_box(const msg123_box_spec *spec, msg123_box_mesh *out);
void msg123_box_mesh_free(msg123_box_mesh *b);

/* Time-step control of the reference (Time_module.f90:34-273,770-858,
 * 1522-1543) around calc_numsol, for a box with constant or
 * piecewise-constant recharge:
 * run until end_time or nstep_max accepted steps.  The time-loop state is
 * kept in the context: a call continues where the previous one stopped
 * unless reset = 1.  Per accepted step,
 * heads_out[k*ncalc..] (may be NULL), times_out[k], dts_out[k],
 * newton_out[k], inner_out[k] receive the results.  Returns the number of
 * accepted steps, or a negative error. */
typedef struct msg123_run_spec {
  float init_step, max_step, end_time; /* seconds */
  int32_t nstep_max;
  int32_t reset;                 /* 1: start from t = 0 */
  /* Optional prescribed schedule: accepted step k ends at time
   * sched_times[k] (float32 seconds, as the reference's current_t); a step
   * that fails is halved and sub-stepped until that time is reached.  With
   * n_sched = 0 the reference's heuristic (apply_heuri) chooses dt. */
  const float *sched_times;
  int32_t n_sched;
  /* Newton initial head of each step attempt:
   * MSG123_WARM_PERSIST  the previous head (the reference, Time_module.f90);
   * MSG123_WARM_LINEAR   h_n + (dt_n/dt_{n-1}) (h_n - h_{n-1});
   * MSG123_WARM_LEARNED  the linear extrapolation corrected by the network
   *                      attached with msg123_set_warm_net in the cells
   *                      unsaturated at h_n, taken only when its residual
   *                      norm is below that of persistence (the reference's
   *                      start, used otherwise) and only while the network's
   *                      steps cost fewer Newton iterations: it is first
   *                      tried on two steps whose network-started solution
   *                      is discarded and which are solved again from
   *                      persistence (msg123_warm_stats.shadows), then used
   *                      only if it cost less than 0.8 of persistence's
   *                      Newton iterations on those steps (a trial is cut
   *                      off, unconverged, at 0.8 of the Newton count of
   *                      the last step accepted from persistence (at least
   *                      one), and the network is then declined without a
   *                      second trial); an attempt from
   *                      it that fails is repeated with the same dt from
   *                      persistence, so the network never changes the
   *                      step sequence;
   * MSG123_WARM_SELECT   persistence or the linear extrapolation, whichever
   *                      has the smaller residual norm (the choice without
   *                      the network). */
  int32_t warm_start;
  /* Failure budget (no reference counterpart: the reference halves a failed
   * step until dt < DSMAL and then stops the program, Time_module.f90:72-74).
   * max_fail: consecutive failed step attempts after which the call returns
   * MSG123_ECONV (0: 64); max_newton: Newton iterations spent on failed
   * attempts per call after which it returns MSG123_ECONV (0: no cap).  A
   * call that returns MSG123_ECONV leaves the time loop at the failed
   * attempt; the next call halves dt and continues.  dt < DSMAL also returns
   * MSG123_ECONV. */
  int32_t max_fail;
  int32_t max_newton;
  /* Probes (observation cells): with n_probe > 0, after every accepted step
   * k < nstep_max the heads of cells probe_idx[0..n_probe) (0-based) are
   * gathered on the device and written to probe_out[k*n_probe ..]. */
  const int32_t *probe_idx;
  int32_t n_probe;
  double *probe_out;
  /* Time-varying recharge: the reference's recharge file with a time unit
   * (Open_file.f90:744-800), whose records are read at their end times by
   * set_nextet / set_nextvar (Time_module.f90:274-330,802-858) and applied
   * by update_tstep (:81-99).  Record k is uniform over the top, rech_mmh[k]
   * mm/h, in effect until rech_etime[k] seconds (float32, increasing; the
   * first one is clipped to end_time, as the reference clips it).  After the
   * last record the recharge is zero, as set_nextet resets it
   * (reset_value, :341-350,1502-1520; the reference then also aborts on its
   * log message, :343, so its decks cover end_time).  Steps are cut to end
   * at the record ends and are not grown on the step after one, as in
   * set_delt and calc_nextst.  With a prescribed schedule a step takes the
   * record in effect at its start.  n_rech = 0: the box's constant
   * recharge. */
  const float *rech_etime;
  const float *rech_mmh;
  int32_t n_rech;
} msg123_run_spec;
#define MSG123_WARM_PERSIST 0
#define MSG123_WARM_LINEAR 1
#define MSG123_WARM_LEARNED 2
#define MSG123_WARM_SELECT 3
int msg123_run_box(msg123_ctx *ctx, const msg123_box_mesh *box, const msg123_run_spec *run,
                   double *heads_out, double *times_out, double *dts_out,
                   int32_t *newton_out, int32_t *inner_out);

/* Statistics of the warm starts of the current run (msg123_run_box calls
 * since the last reset). */
typedef struct msg123_warm_stats {
  int32_t attempts;              /* step attempts with a learned candidate */
  int32_t learned_taken;         /* ... where its residual norm was below persistence's */
  double sum_ratio;              /* sum of ||F(learned)|| / ||F(persistence)|| */
  int32_t retries;               /* attempts from the learned start that failed and were
                                    repeated, same dt, from persistence */
  int32_t shadows;               /* trial steps solved from the learned start, their result
                                    discarded and the step solved again from persistence */
  int32_t shadow_newton;         /* Newton iterations of those trial solves */
  int32_t net_on;                /* the gate's decision: -1 trials still running,
                                    0 persistence chosen, 1 the network used */
} msg123_warm_stats;
int msg123_warm_stats_get(msg123_ctx *ctx, msg123_warm_stats *out);

/* Kernel timing support: total duration (ms) and launch count of the named
 * kernel group, measured with HIP events on the context's stream.
 * enable = 1 times every group, 2 only the SpMV group "spmv" (two events per
 * launch: the other groups' events would add to a timed wall clock), 0 none. */
int msg123_set_timing(msg123_ctx *ctx, int32_t enable);
int msg123_get_timing(msg123_ctx *ctx, const char *name, double *total_ms, int64_t *launches);

/* Append the reference's convergence log (conv_<type>_<name>.txt,
 * Nonlinear_solution.f90:59-63,250-252: a header per time-step attempt and
 * one line per Newton iteration) to `path`; NULL or "" closes it. */
int msg123_set_conv_log(msg123_ctx *ctx, const char *path);

/* Newton iterations of the last msg123_calc_numsol, one record per line of
 * the reference's convergence log (Nonlinear_solution.f90:236-252): counts,
 * the signed largest head change (metres) and, at its cell, the diagonal,
 * right-hand side and head in the reference's units (x L^2, L^3, L).  Up to
 * cap records are copied; *count receives the number available and
 * *end_reason why the loop ended (MSG123_END_*; the reference's messages
 * 12-14 at :265-285).  Lets a caller write the log through its own unit. */
#define MSG123_END_CONVERGED 0
#define MSG123_END_BACKTRACK 1   /* "Stop due to maximum value or change in backtracking" */
#define MSG123_END_MAXITER 2     /* "Stop due to maximum number of nonlinear iteration" */
#define MSG123_END_DIVERGED 3    /* "Didn't converge due to maximum value or change" */
#define MSG123_END_NAN 4         /* NaN in the Newton update */
typedef struct msg123_iter_rec {
  int32_t out_iter, in_iter, back_iter;
  int32_t cell_i, cell_j, cell_k;  /* 1-based cell of the largest change */
  double max_change, dmat, rhs, head;
} msg123_iter_rec;
int msg123_iter_log(msg123_ctx *ctx, msg123_iter_rec *recs, int32_t cap, int32_t *count,
                    int32_t *end_reason);

/* Newton options without a reference counterpart (the reference's values
 * are fixed).  eta_first / eta_max: Eisenstat-Walker forcing (set_eater,
 * Nonlinear_solution.f90:119-123,414-439), the linear tolerance of the first
 * Newton iteration (the reference's 0.5) and the cap of the later ones (the
 * reference's 0.9); a tighter cap trades cheap Krylov iterations for
 * expensive Newton iterations.  algebraic_slope = 1: the line search's
 * slope rhs . (-J dh) from the linear solve's residual r (J dh = rhs - r)
 * instead of calc_vecjacf's difference quotient (Calc_function.f90:515-621),
 * one residual evaluation less per Newton iteration (not with
 * MSG123_PRECON_AMG, whose slope is the reference's quirk).  beta_refine =
 * 0: the line search keeps the step its sufficient-decrease (alpha)
 * condition accepted, without run_backtr's beta-condition refinement
 * (Nonlinear_solution.f90:586-628), whose probes cost one residual
 * evaluation each.  The Newton convergence test, and so the answer, is
 * unchanged by all four; the reference's values are 0.5, 0.9, 0, 1. */
int msg123_set_newton(msg123_ctx *ctx, double eta_first, double eta_max,
                      int32_t algebraic_slope, int32_t beta_refine);

/* Who drives a Newton iteration of msg123_calc_numsol (no reference
 * counterpart; the arithmetic is the same either way):
 * MSG123_CONTROL_HOST    (default) the host reads every norm, dot product
 *                        and test (a wait per BiCGStab half-step, per
 *                        update, per line-search evaluation);
 * MSG123_CONTROL_DEVICE  the Krylov and Newton scalars stay in
 *                        device memory, updated by the reduction folds; the
 *                        kernels of an iteration (assembly, the solve, the
 *                        update, the line search's full step, the residual
 *                        at the new head, the convergence tests and the
 *                        forcing term) are enqueued without waiting, the
 *                        Krylov kernels skipping once the solve has ended,
 *                        and the host reads one record per iteration.  A
 *                        step the line search must shorten, or a step capped
 *                        by maxstep, finishes on the host.  Applies to the
 *                        D-ILU, line and reference-AMG preconditioners; the
 *                        device AMG keeps host control.  The Krylov
 *                        iterations are enqueued in chunks as long as the
 *                        previous solve's, and an iteration's tail after
 *                        each chunk: kernels past the end of the solve
 *                        still occupy the queue, so where the iteration
 *                        count varies it costs more than host control.
 * MSG123_CONTROL_GRAPH   as MSG123_CONTROL_DEVICE, with each chunk's kernels
 *                        (the head of the iteration, its Krylov iterations and
 *                        its tail) captured once in a HIP graph and replayed:
 *                        one graph launch and one synchronisation per chunk.
 *                        A graph is kept per launch shape, law, slope mode,
 *                        bottom, preconditioner, chunk and head buffer, and
 *                        captured anew when one of them differs.  Applies to
 *                        the line and multicolour ILU(0) preconditioners and
 *                        the D-ILU launched per tile diagonal, without kernel
 *                        timing; otherwise the iteration runs as under
 *                        MSG123_CONTROL_DEVICE (or the host's, as there). */
#define MSG123_CONTROL_HOST 0
#define MSG123_CONTROL_DEVICE 1
#define MSG123_CONTROL_GRAPH 2
int msg123_set_control(msg123_ctx *ctx, int32_t mode);

/* How the D-ILU(0) factorisation and sweeps (precon_dilu / solve_ilu,
 * Linear_solution.f90:742-855) are launched; no reference counterpart (the
 * numbers are the same, bit for bit, either way):
 * MSG123_ILU_DIAGONALS  (default) one launch per tile diagonal I+J+K (38
 *                       per sweep at 256x256x64);
 * MSG123_ILU_WAVE       one launch per sweep: blocks take the 16x16x8 tiles
 *                       in dependency order from a ticket counter and wait
 *                       for their three predecessor tiles on per-tile flags
 *                       (bounded waits; a wait that gives up makes the call
 *                       return MSG123_EDEVICE).  Each hand-off between tiles
 *                       costs an agent-scope release and acquire (L2
 *                       write-back and invalidation across the dies), which
 *                       makes it slower than the per-diagonal launches at
 *                       256x256x64. */
#define MSG123_ILU_WAVE 0
#define MSG123_ILU_DIAGONALS 1
int msg123_set_ilu_launch(msg123_ctx *ctx, int32_t mode);

/* Options of the device AMG (MSG123_PRECON_AMG32); the next preconditioner
 * set-up rebuilds the hierarchy with them.  opts = NULL restores the
 * defaults. */
int msg123_set_amg_opts(msg123_ctx *ctx, const msg123_amg_opts *opts);
/* Drop the device hierarchy so that the next set-up aggregates again. */
int msg123_amg_refresh(msg123_ctx *ctx);
/* Levels of the device hierarchy (0 before the first set-up) and, per
 * level up to cap, unknowns and off-diagonal entries. */
int msg123_amg_device_info(msg123_ctx *ctx, int32_t *nlevel, int32_t *n, int32_t *nnz,
                           int32_t *np, int32_t cap);
/* Hierarchy set-ups since the context was created, and linear solves that
 * fell back from the AMG to the FP32 line preconditioner (D-ILU for columns
 * over 640 cells) after a BiCGStab breakdown or 200 iterations without
 * convergence. */
int msg123_amg_stats(msg123_ctx *ctx, int32_t *setups, int32_t *fallbacks);
/* Copy level `level` (1 = the box, as the reference numbers them; >= 2
 * here) of the device hierarchy in the layout of msg123_amg_level: the
 * matrix (offind [n+1], offrow [nnz] 1-based, dmat [n], lumat [nnz]) and the
 * prolongator from level-1 (pindex [nfine+1], poffrow, pval [np]).  Any
 * pointer may be NULL. */
int msg123_amg_device_level(msg123_ctx *ctx, int32_t level, int32_t *offind, int32_t *offrow,
                            double *dmat, double *lumat, int32_t *pindex, int32_t *poffrow,
                            double *pval);

/* Form of the van Genuchten-Mualem evaluation (calc_srat_rperm,
 * Calc_parameter.f90:15-77) in every kernel of the context:
 * MSG123_LAW_REFERENCE the reference's four pow per unsaturated cell;
 * MSG123_LAW_LOGEXP    the same law with five transcendentals (beta = x^n,
 *                      Se = exp(-m log1p beta), 1 - (1 - Se^(1/m))^m =
 *                      -expm1(-m log1p(1/beta))), which also avoids the
 *                      reference form's cancellation in dry cells.  No
 *                      reference counterpart. */
#define MSG123_LAW_REFERENCE 0
#define MSG123_LAW_LOGEXP 1
int msg123_set_law(msg123_ctx *ctx, int32_t law);

/* Free-drainage bottom (unit head gradient below the lowest layer).  No
 * reference counterpart: the reference's bottom face is closed, and its
 * other boundary fluxes (Calc_function.f90:277-513) are not this.  cond has
 * nx*ny values in column order (x fastest): Ks_z of the bottom cell times
 * its bottom-face area, in the mesh's scaled units (len_scal^3/s at
 * kr = 1).  The bottom cell then loses cond * kr each second: the residual,
 * the Jacobian (FD derivative of kr, as on the top face) and the mass
 * balance's connection term carry it.  NULL restores the closed bottom.
 * MSG123_EINVAL for a negative or non-finite value; the array is copied
 * before return. */
int msg123_set_bottom_drain(msg123_ctx *ctx, const double *cond);

/* River, lake, well and sea terms of the residual, the Jacobian and the mass
 * balance: func_riveterm, func_laketerm, func_wellterm, func_sealterm
 * (Calc_function.f90:277-302,393-513), set_rivebound, set_lakebound,
 * set_seabound (Make_linearsystem.f90:287-408,499-547) and their calc_mass
 * terms (Calc_function.f90:103-174).  The arrays are the ones the
 * reference's set-up and time module leave in its module variables, with
 * 1-based cell indices:
 *   rivers  rive2cals(nrive) (top cells), rive_head, rive_bott, abyd_rive
 *           (calc_boundary, set_boundary); lakes likewise;
 *   wells   well_conn(nwell_conn) and calc_well(ncalc) (set_condition,
 *           assign_boundary): each listed cell gets calc_well of that cell;
 *   sea     seal2calc, seal2seal, hydf_seal, abyd_seal (allocate_solution)
 *           and the float32 read_seal(nseal_levels) (assign_boundary).
 * A top cell listed twice takes its last river (lake) entry, as the
 * reference's assignment does; sea connections of one cell are summed in
 * order.  The terms hold until the next call (the reference changes rivers,
 * lakes and calc_well between steps: call again before each step that
 * should see new values); NULL or all counts zero removes them.  Rivers and
 * lakes are evaluated inside the fused stencil kernels; well and sea cells
 * by a pass over their list after them.  Contexts with terms run the Newton
 * iteration under host control (msg123_set_control).  MSG123_EINVAL for an
 * index out of range or a river/lake on a cell outside the top layer. */
typedef struct msg123_bterms {
  int32_t nrive;
  const int32_t *rive2cals;
  const double *rive_head, *rive_bott, *abyd_rive;
  int32_t nlake;
  const int32_t *lake2cals;
  const double *lake_head, *lake_bott, *abyd_lake;
  int32_t nwell_conn;
  const int32_t *well_conn;
  const double *calc_well;       /* [ncalc] */
  int32_t nseal;
  const int32_t *seal2calc, *seal2seal;
  const double *hydf_seal, *abyd_seal;
  int32_t nseal_levels;
  const float *read_seal;
} msg123_bterms;
int msg123_set_bterms(msg123_ctx *ctx, const msg123_bterms *bt);
/* The boundary terms' part of calc_mass at the current state: well, river,
 * lake and sea volumes of the step (Calc_function.f90:103-174, times delt,
 * scaled units, summed over the domain).  msg123_calc_mass's residual
 * (out[4]) includes them; zeros without terms. */
int msg123_calc_mass_bterms(msg123_ctx *ctx, double out[4]);

/* FP64 mass balance: calc_mass (Calc_function.f90:103-174) of the current
 * step (head_new against head_old, times delt), summed over the domain as
 * calc_out_mass does for the whole-domain output (Calc_output.f90:54-166):
 * out[0] storage change (the new and the old head's Ss switch, as calc_mass
 * passes them), out[1] connection flow (zero up to rounding: the flows
 * cancel pairwise; with a free-drainage bottom minus the outflow through
 * it), out[2] recharge, out[3] surface exchange (against
 * surf_old), and beside them out[4] the sum of the Newton residual of
 * calc_func at head_new (current-iterate Ss switch for both levels) and
 * out[5] the sum of |recharge|, all times delt, in the reference's scaled
 * units.  Reads the context's state and writes none of it. */
int msg123_calc_mass(msg123_ctx *ctx, double out[6]);
/* msg123_calc_mass for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same result as
 * msg123_calc_mass (same kernels, launch shape and fixed-order fold), but the
 * call allocates nothing: the per-block partial sums live in a file-scope
 * device array of 3 MiB per GPU (the library's `scratch`, sized for
 * MSG123_NOALLOC_MAX_CELLS cells; larger contexts get MSG123_EINVAL), the
 * constitutive law of head_new goes to work arrays the context allocated at
 * creation.  All calls on one GPU share that GPU's scratch: calls that
 * overlap on one GPU (two host threads, two contexts) race and must be
 * serialised by the caller.  Overwrites the context's saturation and
 * relative-permeability work arrays (not head_new, srat_new, rel_perm). */
#define MSG123_NOALLOC_MAX_CELLS 16777216
int msg123_calc_mass_noalloc(msg123_ctx *ctx, double out[6]);

/* Cumulative mass balance of the accepted steps of msg123_run_box, enabled
 * (and reset) by msg123_set_mass_balance.  The relative error
 * |residual| / recharge_abs is the conservation measure of the Newton
 * solution; total is the reference's mass-balance total. */
typedef struct msg123_mass_balance {
  double storage, connection, recharge, surface, total;
  double residual, recharge_abs;
  int32_t steps;
  double well, river, lake, sea;  /* msg123_set_bterms terms */
} msg123_mass_balance;
int msg123_set_mass_balance(msg123_ctx *ctx, int32_t enable);
int msg123_mass_balance_get(msg123_ctx *ctx, msg123_mass_balance *out);

/* Launch-shape autotuner (no reference counterpart).  The residual and the
 * Jacobian assembly run as fused kernels (the constitutive law evaluated on
 * chip, tiles of 16 x 16 columns marching down chunks of z levels); their
 * launch shape -- the occupancy the compiler is held to (1: unconstrained,
 * 4 or 5 waves per SIMD) and the block count the z chunking aims at (512 ..
 * 8192) -- is chosen by timing every candidate `reps` times on the context's
 * current state (calc_func and make_matvec at head_new, side effects
 * included; delt must be set) with HIP events.  The choice stays with the
 * context; before a tuning every context uses 1 / 1024.  *report (may be
 * NULL) receives the choices and their average durations beside the
 * default's. */
typedef struct msg123_tune_report {
  int32_t residual_waves, residual_blocks;
  int32_t assembly_waves, assembly_blocks;
  double residual_us, residual_default_us;
  double assembly_us, assembly_default_us;
  int32_t candidates;
} msg123_tune_report;
int msg123_autotune(msg123_ctx *ctx, int32_t reps, msg123_tune_report *report);

/* Device stream the context launches on (hipStream_t as void*). */
void *msg123_stream(msg123_ctx *ctx);

/* ---- Learned warm start (no reference counterpart) ------------------ */

/* A 3D convolution network, four 3x3x3 layers (4 -> 16 -> 16 -> 16 -> 1
 * channels, ReLU, zero padding), bf16 weights and activations on the matrix
 * cores with fp32 accumulation, fp32 master parameters and Adam state.  It
 * predicts the correction to the linear extrapolation of the head from the
 * features (extrapolation increment / its RMS, pressure head, recharge
 * depth of the step, lnKs).  Buffers are sized for one grid; all array
 * arguments are device pointers (features [N][4] bf16 bits, outputs and
 * targets [N] float, heads [N] double), work runs on `stream` (NULL: the
 * network's own) and every call returns after it finished. */
#define MSG123_NET_NPARAM 16033
typedef struct msg123_net msg123_net;
int msg123_net_create(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, void *stream,
                      msg123_net **out);
void msg123_net_destroy(msg123_net *net);
/* copy the fp32 master parameters out (set = 0) or in (set = 1); layout per
 * layer: W[tap][cout][cin] (tap = (dz+1)*9 + (dy+1)*3 + (dx+1)), then b[cout] */
int msg123_net_params(msg123_net *net, float *host, int32_t set);
int msg123_net_forward(msg123_net *net, const uint16_t *x, float *y);
/* mean squared error against t and its gradient (host, MSG123_NET_NPARAM) */
int msg123_net_grad(msg123_net *net, const uint16_t *x, const float *t, float *grad_host,
                    double *loss);
/* one Adam step (beta 0.9 / 0.999) on the loss of one sample */
int msg123_net_train_step(msg123_net *net, const uint16_t *x, const float *t, float lr,
                          double *loss);
/* features of a step from heads h_n, h_nm1 (scaled), cell tops (scaled),
 * lnKs, r = dt_n/dt_{n-1}, len_scal, recharge depth of the step (m); writes
 * x, h_lin (may be NULL) and the increment scale (device double) */
int msg123_net_features(msg123_net *net, const double *h_n, const double *h_nm1,
                        const double *cell_top, const float *lnks, double r, double len_scal,
                        double rech_depth, uint16_t *x, double *h_lin, double *scale);
/* training target (h_next - h_lin) * len_scal / scale */
int msg123_net_target(msg123_net *net, const double *h_next, const double *h_lin,
                      const double *scale, double len_scal, float *t);
/* Attach a network (sized for this context's grid) for MSG123_WARM_LEARNED
 * runs: lnks [ncalc] (host) and the recharge rate (m/s) of the top cells. */
int msg123_set_warm_net(msg123_ctx *ctx, msg123_net *net, const float *lnks, double rech_ms);

/* ---- Smoothed-aggregation AMG hierarchy (host) ---------------------- */

/* One level of the hierarchy.  Level 1 is the input system.  CSR arrays
 * follow the reference: offind[0..n] holds row ends (offind[0] = 0),
 * column ids in offrow / poffrow / roffrow are 1-based.  From level 2 on,
 * P (rows = the finer level's nfine unknowns, columns = this level's
 * aggregates) is pindex/poffrow/pval and R = P^T is rindex/roffrow/rval. */
typedef struct msg123_amg_level {
  int32_t n, nnz, nfine;
  int32_t *offind, *offrow;
  double *dmat, *lumat;
  int32_t *pindex, *poffrow, *rindex, *roffrow;
  double *pval, *rval;
} msg123_amg_level;

typedef struct msg123_amg {
  int32_t nlevel;
  msg123_amg_level *levels;      /* [nlevel], owned */
} msg123_amg;

/* make_amgmat (Make_amg_matrix.f90:20-65) on the system (dmat, lumat) with
 * the mesh's CSR pattern: strength filter, heap aggregation, Jacobi-smoothed
 * prolongator, R = P^T and the Galerkin product, level by level up to
 * amg_nlevel or an empty level (set_amg namelist, Read_input.f90:373;
 * amg_theta and jac_omega are float32 there).  Serial, element for element
 * the reference's single-thread result.  Free with msg123_amg_free. */
int msg123_amg_setup(const int32_t *offind, const int32_t *offrow, const double *dmat,
                     const double *lumat, int32_t n, int32_t amg_nlevel, float amg_theta,
                     float jac_omega, msg123_amg *out);
void msg123_amg_free(msg123_amg *amg);

#ifdef __cplusplus
}
#endif
#endif
