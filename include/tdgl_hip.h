/*
 * tdgl_hip.h -- C ABI of libtdglhip.so, the MI355X (gfx950) implementation of
 * tdgl3d's trapezoidal time step (omedeiro/nanowire_tdgl).
 *
 * The reference exposes no FFI: its seam is the Python call chain
 *   solve (src/tdgl3d/solvers/runner.py:40-163)
 *     -> trapezoidal (src/tdgl3d/solvers/integrators.py:92-206)
 *       -> newton_gcr_trap (src/tdgl3d/solvers/newton.py:76-120)
 *         -> tgcr_matrix_free_trap (src/tdgl3d/solvers/tgcr.py:96-156)
 *           -> eval_f (src/tdgl3d/physics/rhs.py:147-239)
 * Each entry point below names the reference function it replaces.
 *
 * Conventions
 *   - Every call returns an int status: TDGL_OK (0) or a negative TDGL_E_*.
 *     tdgl_last_error() describes the last failure on a context.
 *   - Pointers named *_dev are device pointers owned by the caller (for
 *     example torch tensors); they must stay valid until the work queued on
 *     the context's stream has finished.  Host pointers are read during the
 *     call only.
 *   - All work is queued on the context's stream (tdgl_set_stream); calls
 *     that return host values synchronise that stream.
 *   - A context is not thread-safe; use one context per stream.
 *   - "State" buffers use the library's internal layout: interior nodes only,
 *     i fastest, then j, then k; psi as interleaved (re, im) pairs followed
 *     by real phi_x, phi_y and (3-D only) phi_z arrays:
 *         [psi re/im: 2n][phi_x: n][phi_y: n][phi_z: n]   (n = n_interior)
 *     tdgl_permute_in/out convert from/to the reference's complex128 state
 *     vector [psi; phi_x; phi_y; phi_z] in its interior order (i slowest,
 *     k fastest, src/tdgl3d/mesh/indices.py:271-278).
 */
#ifndef TDGL_HIP_H
#define TDGL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDGL_OK 0
#define TDGL_E_INVALID (-1)    /* bad argument or unsupported configuration */
#define TDGL_E_HIP (-2)        /* a HIP runtime call failed */
#define TDGL_E_NOMEM (-3)      /* device allocation failed */
#define TDGL_E_COMPLEX_PHI (-4) /* a link variable has a non-zero imaginary part */
#define TDGL_E_NONFINITE (-5)  /* a norm came back NaN or Inf */

#define TDGL_PREC_F64 0        /* everything in FP64 (plain method) */
#define TDGL_PREC_MIXED 1      /* FP32 stencil/JVP/Krylov vectors, FP64 Newton and Gram scalars */

typedef struct tdgl_ctx tdgl_ctx;
typedef struct tdgl_batch tdgl_batch;

/* Grid and physics parameters; mirrors SimulationParameters
 * (src/tdgl3d/core/parameters.py:9-94).  Nx, Ny, Nz count cells. */
typedef struct tdgl_grid {
  int32_t Nx, Ny, Nz;
  int32_t periodic_x, periodic_y, periodic_z;
  double hx, hy, hz;
  double kappa;
} tdgl_grid;

/* Newton / TGCR options; mirrors the keyword arguments of
 * trapezoidal (integrators.py:100-111), newton_gcr_trap (newton.py:82-87)
 * and tgcr_matrix_free_trap (tgcr.py:100-103).  tgcr_max_iter <= 0 selects
 * the reference default max(N, round(0.2 N)) with N = n_state. */
typedef struct tdgl_solver_opts {
  double newton_tol_f;
  double newton_tol_dx;
  int32_t newton_max_iter;
  int32_t precision; /* TDGL_PREC_F64 or TDGL_PREC_MIXED */
  double tol_gcr;
  double eps_mf;
  int64_t tgcr_max_iter;
  /* TDGL_PREC_MIXED only: the TGCR tolerance used is max(tol_gcr,
   * mixed_tol_floor) (<= 0 selects 1e-6).  FP32 Krylov vectors cannot resolve
   * relative residuals much below 1e-7; the FP64 Newton residual still
   * decides convergence (inexact Newton). */
  double mixed_tol_floor;
  /* TDGL_PREC_MIXED only (no counterpart in the reference): right
   * preconditioner of TGCR, a Chebyshev polynomial M = p(A) of A = I - dt/2 J
   * on the interval [1 - dt/2, 1 + dt/2 ||J||_inf] (||J||_inf bounded from the
   * stencil coefficients with |psi| <= 2).  0: none; > 0: polynomial degree
   * (minus one) = this many stencil sweeps per TGCR iteration; < 0: the
   * degree that contracts the spectrum to within 0.25 of 1 (in a batch, each
   * film's own).  TGCR's stop test
   * stays on the true residual of (I - dt/2 J) dx = -g; the Newton residual
   * is unchanged.  Off when dt/2 >= 0.9 (the interval would reach 0). */
  int32_t mixed_precond;
  /* TDGL_PREC_F64 steps of tdgl_trap_step only (no counterpart in the
   * reference; tdgl_newton_trap and tdgl_tgcr_trap stay the reference's): the
   * same Chebyshev polynomial as a right preconditioner of the FP64 TGCR, with the
   * reference's finite-difference JVP (eps from ||r||) as the operator in its
   * recurrence; TGCR keeps every direction p_k = M r_k it adds, so the update
   * and the stop test stay those of the preconditioned system.  0 (the
   * default): the reference's unpreconditioned TGCR; < 0: the automatic
   * degree; > 0: that degree.  Everything stays FP64; Newton converges to the
   * same discrete solution. */
  int32_t f64_precond;
  /* Predictor (no counterpart in the reference, which starts Newton from
   * X + dt f(X), integrators.py:156-158): with x_prev non-NULL, a film whose
   * dt_prev entry is > 0 starts from the linear extrapolation
   * X + (dt / dt_prev) (X - x_prev) of its last accepted step instead.
   * x_prev: the state before that step (device, internal layout, the same
   * length and offsets as X_dev); dt_prev: host array of the step's sizes, one
   * entry per film (one for tdgl_trap_step); an entry <= 0 keeps the
   * reference predictor for that film.  gamma = X + dt/2 f(X), the equations
   * and the convergence test are unchanged: only Newton's starting point moves.
   * Both NULL (the default) reproduce the reference. */
  const double* x_prev;
  const double* dt_prev;
} tdgl_solver_opts;

/* Per-step counters (the reference keeps none; they make iteration-count
 * parity testable). */
typedef struct tdgl_step_stats {
  int32_t converged;        /* Newton converged (newton.py:118) */
  int32_t newton_iters;     /* value newton_gcr_trap returns as `iterations` */
  int32_t tgcr_failed;      /* the last TGCR returned the empty array (tgcr.py:154-155) */
  int32_t tgcr_iters_last;  /* Krylov iterations of the last TGCR call */
  int32_t tgcr_iters_total; /* Krylov iterations summed over the step */
  int32_t rhs_evals;        /* stencil passes (RHS, predictor, Newton residual) */
  int32_t jvp_evals;        /* fused finite-difference JVP passes */
  int32_t reserved;
  double g_inf;             /* last ||g||_inf (newton.py:112) */
  double dx_inf;            /* last ||dx||_inf (newton.py:113) */
  double bytes_moved;       /* algorithmic HBM bytes of the step */
  double tgcr_r0;           /* ||r0||_2 of the last TGCR call (tgcr.py:116) */
  double tgcr_rk;           /* its final ||r||_2 */
} tdgl_step_stats;

/* Context: grid geometry + device buffers.  Replaces construct_indices'
 * role for the RHS (src/tdgl3d/mesh/indices.py:239-450): index maps become
 * index arithmetic on the internal layout. */
int tdgl_ctx_create(const tdgl_grid* grid, tdgl_ctx** out);
int tdgl_ctx_destroy(tdgl_ctx* ctx);
int tdgl_set_stream(tdgl_ctx* ctx, void* hip_stream);
/* Tuning (no counterpart in the reference): z planes per FP64 stencil work
 * item (tiles are 64 x 4 x kz nodes), including the mixed method's FP64
 * residual stencil; the mixed sweeps size their z pieces to their work
 * (tdgl_batch_set_sweep_split). */
int tdgl_set_stencil_kz(tdgl_ctx* ctx, int kz);
int tdgl_get_stencil_kz(const tdgl_ctx* ctx);
/* Tuning (no counterpart in the reference): tile shape of the mixed-precision
 * stencil sweeps, nodes per plane: 0 = 64 x 4 (default), 1 = 32 x 8,
 * 2 = 16 x 16, 3 = each film a tile of its own, W x R nodes with W from 16
 * to 64 and W R <= 256, the one with the fewest lanes per plane (idle ones
 * included, a little more per lane for narrow tiles).  Results are the same
 * bits for every shape. */
int tdgl_set_sweep_tile(tdgl_ctx* ctx, int shape);
int tdgl_get_sweep_tile(const tdgl_ctx* ctx);
const char* tdgl_last_error(const tdgl_ctx* ctx);
/* Number of doubles in one internal state buffer ((2 + n_phi) * n_interior). */
int64_t tdgl_state_len(const tdgl_ctx* ctx);
int64_t tdgl_n_interior(const tdgl_ctx* ctx);

/* Per-node material and hole data (host arrays, copied to the device).
 *   link_bits_full : dim_x bytes in the reference's full-grid numbering
 *                    m = i + (Nx+1) j + (Nx+1)(Ny+1) k; bit 0/1/2 = the x/y/z
 *                    link at m is a hole-boundary link (GridIndices.hole_*_bc_mask,
 *                    indices.py:126-134, zeroed by rhs.py:132-137 and rhs.py:229-234).
 *                    NULL = no holes.
 *   kappa_interior : n doubles, MaterialMap.kappa at the interior nodes in the
 *                    reference interior order (sparse_operators.py:19-28); NULL
 *                    = uniform grid->kappa (no material).
 *   sc_interior    : n doubles, MaterialMap.interior_sc_mask (material.py:35-41,
 *                    read by sparse_operators.py:221-224); NULL = no material.
 * Replaces MaterialMap / Device.add_hole (core/material.py:26-112,
 * core/device.py:80-153) as seen by eval_f. */
int tdgl_set_material(tdgl_ctx* ctx, const uint8_t* link_bits_full,
                      const double* kappa_interior, const double* sc_interior);

/* Applied field of BoundaryVectors (physics/rhs.py:32-40) built by
 * build_boundary_field_vectors (physics/applied_field.py:59-91): uniform values
 * on the face-inner nodes. */
int tdgl_set_field(tdgl_ctx* ctx, double bx, double by, double bz);
/* General BoundaryVectors: three host arrays of dim_x doubles (full-grid
 * numbering); only their face-inner entries are read (rhs.py:95-127). */
int tdgl_set_field_vectors(tdgl_ctx* ctx, const double* bx_full, const double* by_full,
                           const double* bz_full);

/* Layout conversion between the reference state vector (complex128, n_state,
 * StateVector layout core/state.py:11-30) and the internal layout.
 * tdgl_permute_in fails with TDGL_E_COMPLEX_PHI if any phi entry has a
 * non-zero imaginary part (the build stores phi as real). */
int tdgl_permute_in(tdgl_ctx* ctx, const void* ref_state_dev, double* state_dev);
int tdgl_permute_out(tdgl_ctx* ctx, const double* state_dev, void* ref_state_dev);

/* f = eval_f(x) (physics/rhs.py:147-239), FP64, internal layout. */
int tdgl_rhs(tdgl_ctx* ctx, const double* x_dev, double* f_dev);

/* Ap = p - dt/2 * (f(x + eps p) - f0) * (1/eps)  (tgcr.py:129-133), FP64,
 * with f0 = f(x) supplied by the caller. */
int tdgl_jvp_trap_fd(tdgl_ctx* ctx, const double* x_dev, const double* f0_dev,
                     const double* p_dev, double eps, double dt, double* ap_dev);

/* Ap = p - dt/2 * J(x) p with the analytic Jacobian of eval_f, FP32 vectors in
 * the internal layout.  The TGCR of TDGL_PREC_MIXED uses this in place of the
 * finite difference of tgcr.py:129-133 (its eps -> 0 limit). */
int tdgl_jvp_trap32(tdgl_ctx* ctx, const float* x32_dev, const float* p32_dev, double dt,
                    float* ap32_dev);

/* One application of the mixed TGCR's Chebyshev preconditioner (no
 * counterpart in the reference; see tdgl_solver_opts.mixed_precond), FP32,
 * internal layout, at the linearisation point x32 with A = I - dt/2 J(x32):
 *   w32 = A M v = (I - P(A)) v,  z32 = M v,
 * P(t) = T_d((theta - t)/delta) / T_d(sigma) on the interval [a, b] the
 * library picks (theta = (a + b)/2, delta = (b - a)/2, sigma = theta/delta).
 * degree: > 0 the polynomial degree d, < 0 the automatic choice.
 * info (host double[4]) receives a, b, d, and 1.0 when d = 1 (z32 is then
 * not written: z = v / theta). */
int tdgl_precond_cheb32(tdgl_ctx* ctx, const float* x32_dev, const float* v32_dev, double dt,
                        int degree, float* w32_dev, float* z32_dev, double* info);

/* Adjoint of eval_f (no counterpart in the reference; the hand-written
 * transpose of the Jacobian of rhs.py:147-239 at x, at the current field and
 * material), FP64:
 *   gx_dev  = J(x)^T w                                       (internal layout)
 *   dfield  = d<w, F(x)>/d(Bx, By, Bz) of the uniform applied field (host
 *             double[3]; NULL to skip)
 *   dkappa  = d<w, F(x)>/d kappa with every node's kappa tied (host; NULL to skip)
 *   dsc_dev = d<w, F(x)>/d sc_q, sc the superconducting mask of interior
 *             node q (internal order; NULL to skip)
 * <a, b> is the real inner product of the internal layout (psi as re, im). */
int tdgl_vjp_f(tdgl_ctx* ctx, const double* x_dev, const double* w_dev, double* gx_dev,
               double* dfield, double* dkappa, double* dsc_dev);
/* Adjoint of the B_z observable (core/solution.py:124-158 -> bfield.py:58-68)
 * at the current field: gx_dev = (dB_z/dx)^T gbz (gbz in the reference interior
 * order, n values), dfield = d<gbz, B_z>/d(Bx, By, Bz) (host double[3] or NULL). */
int tdgl_vjp_bz(tdgl_ctx* ctx, const double* gbz_dev, double* gx_dev, double* dfield);
/* Adjoint step solve (no counterpart in the reference; the linear system of
 * the trapezoidal step's adjoint): out_dev = mu with (I - hdt J(x1)^T) mu = rhs,
 * by TGCR (the solver of tgcr.py:96-156, with the transposed operator) to
 * ||residual|| <= tol ||rhs||, at most max_iter iterations (<= 0: the
 * reference's default cap); *iters (may be NULL) = iterations taken.
 * precond: the Chebyshev right preconditioner's degree (-1 automatic, as the
 * mixed method's; 0 none).  Uses the current field and material.
 * TDGL_E_NONFINITE when TGCR stops short. */
int tdgl_adjoint_solve(tdgl_ctx* ctx, const double* x1_dev, const double* rhs_dev, double hdt,
                       double tol, int64_t max_iter, int precond, double* out_dev, int* iters);

/* Reductions over internal-layout vectors (the reference's np.vdot(...).real,
 * np.linalg.norm(., 2) and np.linalg.norm(., inf): tgcr.py:116,129,136,140,149,152;
 * newton.py:100,112-113).  Results are returned to the host. */
int tdgl_dot(tdgl_ctx* ctx, const double* a_dev, const double* b_dev, double* out);
int tdgl_norm_inf(tdgl_ctx* ctx, const double* a_dev, double* out);

/* One trapezoidal step attempt at the current field: predictor, Newton,
 * TGCR (integrators.py:151-172 + newton.py:76-120 + tgcr.py:96-156).
 * On convergence X_dev is overwritten with the new state and stats->converged
 * is 1; otherwise X_dev is left unchanged (the caller divides dt by 10 as
 * integrators.py:185-197 does).  opts->precision selects the plain FP64
 * method or the mixed-precision method (the film as a batch of one, see
 * tdgl_batch_trap_step). */
int tdgl_trap_step(tdgl_ctx* ctx, double* X_dev, double dt, const tdgl_solver_opts* opts,
                   tdgl_step_stats* stats);

/* newton_gcr_trap (src/tdgl3d/solvers/newton.py:76-120) on its own, FP64, at
 * the current field: Newton on g(x) = x - dt/2 f(x) - gamma from x_dev (in:
 * the start x0; out: the last iterate -- the converged one, the one whose
 * TGCR failed, or the last of newton_max_iter iterations, as the reference
 * returns it).  stats->converged / newton_iters are the reference's
 * (converged, iterations).  opts->precision must be TDGL_PREC_F64. */
int tdgl_newton_trap(tdgl_ctx* ctx, double* x_dev, const double* gamma_dev, double dt,
                     const tdgl_solver_opts* opts, tdgl_step_stats* stats);

/* tgcr_matrix_free_trap (src/tdgl3d/solvers/tgcr.py:96-156) on its own, FP64,
 * at the current field: dx_dev with (I - dt/2 J(x_lin)) dx = b, J p by the
 * finite difference of eval_f at x_lin (opts->eps_mf), tolerance
 * opts->tol_gcr, at most opts->tgcr_max_iter iterations (<= 0: the
 * reference's default).  *converged = 0 is the reference's empty-array return
 * (dx_dev is then zero); stats->tgcr_iters_last = its iteration count.
 * opts->precision must be TDGL_PREC_F64. */
int tdgl_tgcr_trap(tdgl_ctx* ctx, const double* x_lin_dev, const double* b_dev, double dt,
                   const tdgl_solver_opts* opts, double* dx_dev, int* converged,
                   tdgl_step_stats* stats);

/* One Forward-Euler step X += dt f(X) (integrators.py:78-83). */
int tdgl_euler_step(tdgl_ctx* ctx, double* X_dev, double dt);

/* tdgl_euler_step without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  X + dt f(X) is staged in a file-scope device
 * array of tdgl_noalloc_max_len() doubles per GPU (32 MiB; its address is
 * resolved when the first context on that GPU is created) instead of a work
 * buffer the library allocates.  TDGL_E_INVALID when the state is longer.
 * Every call on one GPU shares that array: calls that overlap on a GPU race. */
int tdgl_euler_step_noalloc(tdgl_ctx* ctx, double* X_dev, double dt);
int64_t tdgl_noalloc_max_len(void);

/* Check-path observables on an internal-layout state, written in the
 * reference interior order (3 x n doubles each):
 *   B  = curl(phi) on the expanded state with BCs at the current field
 *        (core/solution.py:124-158 -> physics/bfield.py:12-78);
 *   Js = Im[conj(psi) e^{-i phi} psi(+a)] on the expanded state without BCs
 *        (core/solution.py:180-217 -> physics/current_density.py:14-72).
 * Either output pointer may be NULL. */
int tdgl_observables(tdgl_ctx* ctx, const double* x_dev, double* b_dev, double* js_dev);

/* ---- Batches of films (mixed precision; no counterpart in the reference,
 * whose trapezoidal() advances one film per call) ----
 * A batch advances several films -- each a context with its own grid,
 * material and field -- by one trapezoidal step, with tdgl3d's per-film
 * control flow (integrators.py:151-172, newton.py:76-120, tgcr.py:96-156)
 * run in lockstep: every kernel launch covers all films still iterating.
 * The contexts must outlive the batch; field and material changes made on
 * them are picked up at the next step.  All films are 3-D, or all 2-D.
 * A batched state is one device buffer of tdgl_batch_total_len doubles; film
 * f occupies tdgl_state_len(film f) doubles from tdgl_batch_offset(f) in the
 * internal layout; the padding between films is never read. */
int tdgl_batch_create(tdgl_ctx* const* films, int nfilm, tdgl_batch** out);
int tdgl_batch_destroy(tdgl_batch* batch);
int tdgl_batch_set_stream(tdgl_batch* batch, void* hip_stream);
/* The batch's own (non-blocking) stream, the one its kernels run on: work that
 * a caller orders on it (as the caller's stream of tdgl_batch_set_stream)
 * needs no event between the two.  No counterpart in the reference: it lets
 * sub-batches stepped concurrently each keep to one stream. */
int tdgl_batch_work_stream(const tdgl_batch* batch, void** hip_stream);
/* Move the batch's work to a new stream that has a hardware queue of its own
 * (made with a CU mask of every CU; HIP otherwise shares its
 * GPU_MAX_HW_QUEUES queues among a process's streams, and batches stepped at
 * once on one queue run one after the other).  The new stream is blocking
 * (ordered after work on the legacy stream).  No counterpart in the reference. */
int tdgl_batch_own_queue(tdgl_batch* batch);
const char* tdgl_batch_last_error(const tdgl_batch* batch);
int64_t tdgl_batch_total_len(const tdgl_batch* batch);
int64_t tdgl_batch_offset(const tdgl_batch* batch, int film);
/* Tuning: z planes per FP64 stencil work item of film `film` (the batch's
 * predictor and Newton residual stencils). */
int tdgl_batch_set_kz(tdgl_batch* batch, int film, int kz);
int tdgl_batch_get_kz(const tdgl_batch* batch, int film);
/* Tuning: tile shape of the batch's FP32 sweeps (as tdgl_set_sweep_tile; the
 * autotuner's other knob). */
int tdgl_batch_set_sweep_tile(tdgl_batch* batch, int shape);
int tdgl_batch_get_sweep_tile(const tdgl_batch* batch);
/* Tuning: how each FP32 sweep's work is split into blocks.  Every tile column
 * of the films taking part in a sweep is cut along z into near-equal pieces of
 * at most P planes, one block per piece, P = (the sweep's tile-planes / the
 * blocks the device holds at once) clamped to [min_planes, max_planes]: deep
 * pieces when every film sweeps, shallow ones when a sweep covers a few films.
 * Defaults 8 and 2; results are the same bits for every split. */
int tdgl_batch_set_sweep_split(tdgl_batch* batch, int max_planes, int min_planes);
int tdgl_batch_get_sweep_split(const tdgl_batch* batch, int* max_planes, int* min_planes);
/* Average milliseconds of `reps` launches of the batch's FP32 JVP stencil over
 * its work buffers, as currently laid out (what the autotuner measures). */
int tdgl_batch_time_stencil(tdgl_batch* batch, int reps, double* ms);
/* TGCR iterations after the first replay one captured hipGraph of the
 * iteration's kernels (default on); off: the same kernels launched eagerly.
 * Results are identical either way. */
int tdgl_batch_set_graphs(tdgl_batch* batch, int enable);
/* One step attempt for every film f with active[f] != 0 (active may be NULL:
 * all), with its own dt[f] (host array).  stats[f] as tdgl_trap_step's;
 * opts->precision must be TDGL_PREC_MIXED. */
int tdgl_batch_trap_step(tdgl_batch* batch, double* X_dev, const double* dt, const int* active,
                         const tdgl_solver_opts* opts, tdgl_step_stats* stats);
int tdgl_batch_profile_enable(tdgl_batch* batch, int enable);
int tdgl_batch_profile_read(tdgl_batch* batch, int kernel_class, double* ms_total,
                            int64_t* launches, double* bytes_total);

/* Instrumentation (no counterpart in the reference): when enabled, the driver
 * brackets every kernel launch of the trapezoidal step with HIP events on the
 * context's stream and accumulates, per kernel class, the summed duration,
 * the launch count and the algorithmic HBM bytes (each vector pass counted
 * once).  tdgl_profile_read synchronises the stream. */
#define TDGL_K_PRED 0      /* predictor + gamma stencil */
#define TDGL_K_NEWTON 1    /* Newton residual stencil (f, r = -g, norms) */
#define TDGL_K_JVP 2       /* fused JVP stencil (FP64 finite difference; mixed: analytic FP32) */
#define TDGL_K_MGS 3       /* MGS pass: Ap -= beta Ap_j, next beta (mixed: CGS dot products) */
#define TDGL_K_MGS_LAST 4  /* last MGS update + ||Ap|| (mixed: CGS update Ap -= Q c) */
#define TDGL_K_FINAL_P 5   /* p_k assembly + normalisation + alpha (mixed: Q_k, alpha, r update) */
#define TDGL_K_UPDATE_R 6  /* r -= alpha Ap + ||r|| (mixed: stop tests) */
#define TDGL_K_ACCUM 7     /* dx = sum alpha p; x += dx; ||dx||_inf (mixed: H y = alpha, x += R y) */
#define TDGL_K_CGS2 8      /* mixed only: the conditional second CGS pass */
#define TDGL_K_CHEB 9      /* mixed only: Chebyshev preconditioner sweeps (stencil) */
#define TDGL_K_COUNT 10
int tdgl_profile_enable(tdgl_ctx* ctx, int enable);
int tdgl_profile_read(tdgl_ctx* ctx, int kernel_class, double* ms_total, int64_t* launches,
                      double* bytes_total);

#ifdef __cplusplus
}
#endif

#endif /* TDGL_HIP_H */
