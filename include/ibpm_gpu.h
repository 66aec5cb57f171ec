/*
 * ibpm_gpu.h -- C ABI of the gfx950 (MI355X) projection-step library.
 *
 * Plain pointers and sizes only.  Every call returns 0 on success and a
 * negative value on failure (never aborts); ibpm_last_error() then describes
 * the failure (per host thread).  Calls are ordered on the context's single
 * HIP stream; a context is not thread-safe, but batches on different contexts
 * may be driven from different host threads at once (their work then shares
 * the GPU, one stream each).  "d_" pointers are device pointers (allocate
 * them with ibpm_malloc or any HIP allocator on the context's device).
 *
 * Layouts are the reference's (so a host Scalar/Flux/BoundaryVector buffer
 * copies 1:1):
 *   Scalar  [lev][i=1..nx-1][j=1..ny-1]           reference src/Scalar.cc:81-89
 *   Flux    [lev][X (nx+1)*ny | Y nx*(ny+1)]       reference src/Flux.cc:42-50
 *   BoundaryVector [X_0..X_{n-1}, Y_0..Y_{n-1}]    reference src/BoundaryVector.h:65-69
 *   BC ring 2(nx+ny): left(0..ny), top(1..nx), right(ny-1..0), bottom(nx-1..1)
 *                                                  reference src/BC.h:22-79
 */
#ifndef IBPM_GPU_H
#define IBPM_GPU_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ibpm_ctx ibpm_ctx;
typedef struct ibpm_batch ibpm_batch;

/* ---- errors, devices, memory -------------------------------------------- */
const char* ibpm_last_error(void);
int ibpm_device_count(void);
int ibpm_ctx_create(int device, ibpm_ctx** out);
int ibpm_ctx_destroy(ibpm_ctx* ctx);
int ibpm_ctx_synchronize(ibpm_ctx* ctx);
int ibpm_malloc(ibpm_ctx* ctx, size_t bytes, void** d_ptr);
int ibpm_free(ibpm_ctx* ctx, void* d_ptr);
int ibpm_memcpy_h2d(ibpm_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int ibpm_memcpy_d2h(ibpm_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);

/* ---- elliptic solves (EllipticSolver2d / EllipticSolver) ------------------ */
/* Unnormalised 2-D DST-I (RODFT00 x RODFT00) of an n1 x n2 row-major array.
   Replaces EllipticSolver2d::sinTransform, reference src/EllipticSolver2d.cc:57-71. */
int ibpm_dst2d(ibpm_ctx* ctx, int n1, int n2, const double* d_in, double* d_out);

/* Single-level solve with boundary values: kind 0 Poisson (L u = f),
   1 Helmholtz ((1 + alpha L) u = f); d_bc may be NULL (zero boundary).
   Replaces EllipticSolver2d::solve(f, [bc,] u), src/EllipticSolver2d.cc:87-104. */
int ibpm_elliptic2d_solve(ibpm_ctx* ctx, int kind, int nx, int ny, double dx, double alpha,
                          const double* d_f, const double* d_bc, double* d_u);

/* Multi-domain solve over ngrid levels (coarsify, coarsest with zero BC,
   finer levels with BCs from the coarser solution).  d_f and d_u may alias.
   Replaces EllipticSolver::solve, src/EllipticSolver.cc:47-73. */
int ibpm_elliptic_solve(ibpm_ctx* ctx, int kind, int nx, int ny, int ngrid, double dx,
                        double alpha, const double* d_f, double* d_u);

/* The multi-domain solve of ibpm_elliptic_solve without any allocation.  No
   counterpart in the reference: it is for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc).  Its own buffers (right-hand-side copy, coarsify and
   transpose buffers, tables, job descriptors) live in a fixed 32 MiB device
   array that every GPU the library runs on holds until the process ends; a
   grid needs 16 (nx-1)(ny-1) ngrid bytes plus about 16 (nx-1)(ny-1) for the
   transposes, so e.g. 512 x 512 with 5 levels fits and 1024 x 1024 with 6 does
   not: sizes beyond the array, or more than 16 levels, return -2 with a
   message.  The caller passes d_f and d_u (which may alias).  Calls on one GPU
   share its array, so calls that overlap on one GPU (other threads, other
   contexts of the same device) race; calls are synchronous. */
int ibpm_elliptic_solve_noalloc(ibpm_ctx* ctx, int kind, int nx, int ny, int ngrid, double dx,
                                double alpha, const double* d_f, double* d_u);

/* ibpm_elliptic_solve with a choice of transform precision: 0 FP64 (the
   reference's arithmetic; ibpm_elliptic_solve), 1 FP32 transforms and
   transposed intermediates with the right-hand side, boundary terms,
   eigenvalue scaling and output in FP64.  No counterpart in the reference
   (whose solve is FP64 FFTW, src/EllipticSolver2d.cc:57-104); for measuring
   the reduced-precision path's error per solve.  2: the split-BF16 GEMM-DST
   on MFMA (square grids of N = nx - 1 <= 511 only). */
int ibpm_elliptic_solve_precision(ibpm_ctx* ctx, int kind, int nx, int ny, int ngrid, double dx,
                                  double alpha, const double* d_f, double* d_u, int precision);

/* ---- grid operators (VectorOperations) ----------------------------------- */
/* Curl(Flux -> Scalar), src/VectorOperations.cc:33-54 */
int ibpm_curl_flux(ibpm_ctx* ctx, int nx, int ny, int ngrid, double dx, const double* d_q, double* d_w);
/* Curl(Scalar -> Flux) + uniform flow (u0, v0) * dx_lev, src/VectorOperations.cc:63-123,
   src/Flux.cc:79-99 (pass u0 = v0 = 0 for the bare curl) */
int ibpm_curl_scalar(ibpm_ctx* ctx, int nx, int ny, int ngrid, double dx, double u0, double v0,
                     const double* d_w, double* d_q);
/* NonlinearIBSolver::N = Curl(CrossProduct(q, omega)), src/IBSolver.cc:205-209 */
int ibpm_nonlinear(ibpm_ctx* ctx, int nx, int ny, int ngrid, double dx, const double* d_q,
                   const double* d_w, double* d_n);
/* CrossProduct(Flux q, Scalar omega) -> Flux (omega v, -omega u), src/VectorOperations.cc:438-457 */
int ibpm_cross_product(ibpm_ctx* ctx, int nx, int ny, int ngrid, double dx, const double* d_q,
                       const double* d_w, double* d_out);
/* CrossProduct(Flux q1, Flux q2) -> Scalar u1 v2 - u2 v1, coarsified over the
   overlap regions, src/VectorOperations.cc:465-486 */
int ibpm_cross_flux_flux(ibpm_ctx* ctx, int nx, int ny, int ngrid, double dx, const double* d_q1,
                         const double* d_q2, double* d_out);
/* Laplacian = -Curl(Curl(omega)), src/VectorOperations.cc:133-140 */
int ibpm_laplacian(ibpm_ctx* ctx, int nx, int ny, int ngrid, double dx, const double* d_w, double* d_out);
/* Right-hand side of one substep, IBSolver::advanceSubstep (src/IBSolver.cc:167-183) with
   N = NonlinearIBSolver::N (:205-209):
     N = Curl(CrossProduct(q, omega)),
     a = omega + dt (0.5 nu (an + bn) Laplacian(omega) + an N + bn N_prev).
   d_nprev NULL means no previous N yet (the reference then uses N itself).
   Writes d_a and d_n (d_n may alias d_nprev). */
int ibpm_rhs(ibpm_ctx* ctx, int nx, int ny, int ngrid, double dx, double dt, double nu, double an, double bn,
             const double* d_q, const double* d_w, const double* d_nprev, double* d_a, double* d_n);
/* NavierStokesModel::computeFlux (src/NavierStokesModel.cc:114-124): psi solves
   L psi = -coarsify(omega) over all levels (:131-139), q = Curl(psi) + q0 with the
   uniform base flow (u0, v0) * dx_lev (src/Flux.cc:79-99). */
int ibpm_flux_from_vorticity(ibpm_ctx* ctx, int nx, int ny, int ngrid, double dx, double u0, double v0,
                             const double* d_w, double* d_q);

/* ---- Regularizer (finest level; src/Regularizer.cc:31-128) ----------------- */
typedef struct ibpm_reg ibpm_reg;
/* Regularizer(grid, geometry) + update() for npts points h_xy = [x_0.., y_0..]. */
int ibpm_reg_create(ibpm_ctx* ctx, int nx, int ny, int ngrid, double length, double xoffset, double yoffset,
                    int npts, const double* h_xy, ibpm_reg** out);
/* Regularizer::update after the points moved (same count). */
int ibpm_reg_update(ibpm_reg* reg, const double* h_xy);
int ibpm_reg_destroy(ibpm_reg* reg);
/* Regularizer::toFlux: d_q (all levels) = 0 except level 0 = E^T f * dx (Regularizer.cc:92-109). */
int ibpm_reg_to_flux(ibpm_reg* reg, const double* d_f, double* d_q);
/* Regularizer::toBoundary: d_f = E q / dx from level 0 of d_q (Regularizer.cc:111-128). */
int ibpm_reg_to_boundary(ibpm_reg* reg, const double* d_q, double* d_f);

/* ---- geometry ------------------------------------------------------------- */
/* Number of boundary points of a .geom text (Geometry::getNumPoints after
   Geometry::load, src/Geometry.cc:143-191), or -1 if it does not parse. */
int ibpm_geom_num_points(const char* geom_text);
/* Geometry::getNumPoints, getNumBodies, isStationary and the centre of the
   first body (Geometry.cc:111-113, 72-77; only "motion fixed" and no motion
   are stationary, Motion.h:28, FixedPosition.h:39).  Any output may be null.
   Returns 0, or -1 if the text does not parse. */
int ibpm_geom_info(const char* geom_text, int* npoints, int* nbodies, int* stationary, double* xc0, double* yc0);
/* Geometry::moveBodies(t) at each of times[0..ntimes) in order, then
   getPoints / getVelocities (src/Geometry.cc:115-128, RigidBody.cc:586-609,
   the Motion classes FixedPosition.h ... MotionFilePeriodic.h).  Motions that
   integrate over successive calls (FixedVelocity, Eldredge2,
   EldredgeCombined2) keep their state across the sequence as the reference's
   do.  Call c writes [x.., y..] to xy + 2 n c and [u.., v..] to uv + 2 n c;
   cap is the length of each array.  Host memory, no GPU work.  Returns n, the
   number of points, or -1. */
int ibpm_geom_points(const char* geom_text, int ntimes, const double* times, double* xy, double* uv, long cap);

/* ---- whole projection-step solver (IBSolver / ProjectionSolver / ...) ----- */
typedef struct {
    int nx, ny, ngrid;          /* Grid(nx, ny, ngrid, length, xoffset, yoffset), src/Grid.h:34-41 */
    double length, xoffset, yoffset;
    double reynolds;            /* NavierStokesModel(grid, geom, Re, q0) */
    double dt;
    int scheme;                 /* 0 Euler, 1 AB2, 2 RK3, 3 RK3b (src/Scheme.h:14-56) */
    double alpha_deg;           /* base-flow angle in degrees (ibpm.cc -alpha) */
    double cg_tol;              /* CG absolute residual tolerance (IBSolver setTol) */
    int force_cg;               /* 1: CG even for stationary bodies */
    const char* geom;           /* .geom file text (src/Geometry.cc:143-191) */
    int ubf;                    /* 1: unsteady base flow (ibpm.cc -ubf, :166-178): the motion of the
                                   only body moves to the base flow and the flow is solved in the
                                   body frame; more than one body is an error */
    int model;                  /* ibpm.cc -model (:185-235): 0 nonlinear, 1 linear, 2 adjoint,
                                   3 linear periodic, 4 SFD.  The linear models (1-3) have no
                                   potential-flow part (NavierStokesModel(grid, geom, Re)) and need
                                   base states (ibpm_batch_set_base_state) before init */
    int period;                 /* linear periodic: number of base states (phase = step % period) */
    double sfd_chi, sfd_delta;  /* SFD gain and filter width (ibpm.cc -chi -Delta) */
} ibpm_flow_params;

/* ---- NavierStokesModel / ProjectionSolver operators of one flow -------------
   An operator-level interface over the stand-alone entry points above (one
   flow, synchronous calls); the batch below is the fast path for stepping.
   Device arrays: Scalar (ns doubles), Flux, BoundaryVector (2 Nb). */
typedef struct ibpm_model ibpm_model;
/* NavierStokesModel(grid, geometry, Re, BaseFlow(grid, 1, alpha)) + init():
   p gives the grid, Reynolds number, base-flow angle (alpha_deg) and .geom
   text (other fields unused); the regulariser is built for the points at
   time t0.  Reference src/NavierStokesModel.h:40-56, NavierStokesModel.cc:19-57. */
int ibpm_model_create(ibpm_ctx* ctx, const ibpm_flow_params* p, double t0, ibpm_model** out);
int ibpm_model_destroy(ibpm_model* m);
/* NavierStokesModel::getNumPoints */
int ibpm_model_num_points(ibpm_model* m);
/* The uniform base flow q0 of magnitude mag at angle alpha (radians):
   BaseFlow(grid, mag, alpha) (BaseFlow.h, UniformFlow Flux.cc:79-99);
   ibpm_model_create sets magnitude 1 at p->alpha_deg, mag 0 is the
   model without a potential-flow part (NavierStokesModel.h:48-52). */
int ibpm_model_set_base_flow(ibpm_model* m, double mag, double alpha);
/* NavierStokesModel::updateOperators(time) (NavierStokesModel.cc:93-99): the
   bodies move to `time` (the positions of one Geometry::moveBodies call at
   that time) and the regulariser is updated. */
int ibpm_model_update_operators(ibpm_model* m, double time);
/* NavierStokesModel::B: omega = Curl(E^T f)  (NavierStokesModel.cc:101-105) */
int ibpm_model_B(ibpm_model* m, const double* d_f, double* d_omega);
/* NavierStokesModel::C: f = E Curl(psi), L psi = -coarsify(omega)  (:107-119, 133-140) */
int ibpm_model_C(ibpm_model* m, const double* d_omega, double* d_f);
/* NavierStokesModel::computeFlux: q = Curl(psi) + q0  (:121-125) */
int ibpm_model_compute_flux(ibpm_model* m, const double* d_omega, double* d_q);
/* NavierStokesModel::getConstraints: h_b = body velocities - E q0 (host
   BoundaryVector, :86-91, 142-146) */
int ibpm_model_constraints(ibpm_model* m, double* h_b);
/* ProjectionSolver::M: y = C A^-1 (beta B f), A = 1 - beta/(2 Re) L
   (ProjectionSolver.cc:85-101, the Helmholtz alpha of :33) */
int ibpm_model_M(ibpm_model* m, double beta, const double* d_f, double* d_y);
/* CholeskySolver::computeMatrixM (CholeskySolver.cc:63-79): the n x n
   (n = 2 Nb) matrix M, row-major, one GPU M application per column. */
int ibpm_cholesky_build(ibpm_model* m, double beta, double* h_M);
/* CholeskySolver::computeFactorization (CholeskySolver.cc:89-109), FP64: the
   factor L of M from M's upper part; h_lower n x n row-major with the strict
   lower triangle (zeros above), h_diag its diagonal.  -2 if M is not
   positive definite. */
int ibpm_cholesky_factor(int n, const double* h_M, double* h_lower, double* h_diag);
/* CholeskySolver::Minv (CholeskySolver.cc:200-224), FP64 forward and back
   substitution with the reference's summation order; h_x may alias h_b. */
int ibpm_cholesky_solve(int n, const double* h_lower, const double* h_diag, const double* h_b, double* h_x);
/* CholeskySolver::save / load (CholeskySolver.cc:113-197): the factor in the
   reference's .cholesky text layout (size, alpha*beta, diagonal, '#', strict
   lower triangle by rows).  load returns -2 when the file is missing or its
   size or alpha*beta differ (the reference's "false"). */
int ibpm_cholesky_save(const char* path, int n, double alpha_beta, const double* h_lower, const double* h_diag);
int ibpm_cholesky_load(const char* path, int n, double alpha_beta, double* h_lower, double* h_diag);
/* ConjugateGradientSolver::Minv (ConjugateGradientSolver.cc:35-65): CG on M
   from h_x0 (NULL: zero) with FP64 vectors and residual, each M application
   on the GPU; stops when r.r <= tol^2 (absolute) or after max_it iterations
   (max_it <= 0: no cap, as the reference).  Writes the iteration count and
   whether the test was met. */
int ibpm_cg_solve(ibpm_model* m, double beta, double tol, int max_it, const double* h_b, const double* h_x0,
                  double* h_x, int* it_count, int* converged);

/* A batch of independent flows advanced together by the same kernels. */
int ibpm_batch_create(ibpm_ctx* ctx, int nflows, const ibpm_flow_params* params, ibpm_batch** out);
int ibpm_batch_destroy(ibpm_batch* b);
/* ibpm.cc:283-305: move bodies, regulariser update, Cholesky build (stationary)
   or CG set-up (moving), refreshState. */
int ibpm_batch_init(ibpm_batch* b);
/* IBSolver::advance, src/IBSolver.cc:142-150, nsteps times. */
int ibpm_batch_advance(ibpm_batch* b, int nsteps);
/* Record each step's net force on the device (State::computeNetForce). */
int ibpm_batch_num_points(ibpm_batch* b, int flow);
int ibpm_batch_get_omega(ibpm_batch* b, int flow, double* h_out);
int ibpm_batch_set_omega(ibpm_batch* b, int flow, const double* h_in);   /* + refreshState */
int ibpm_batch_get_q(ibpm_batch* b, int flow, double* h_out);
int ibpm_batch_get_f(ibpm_batch* b, int flow, double* h_out);
int ibpm_batch_get_points(ibpm_batch* b, int flow, double* h_xy);
/* Cd and Cl after the last step, ibpm.cc:346-355 (2 x the rotated net force). */
int ibpm_batch_get_forces(ibpm_batch* b, int flow, double* cd, double* cl);
/* Force history: (Cd, Cl) of every step since init, 2 doubles per step. */
int ibpm_batch_force_history(ibpm_batch* b, int flow, double* h_out, int max_steps);
/* Base-flow angle (radians) and magnitude that rotate the current force into
   drag and lift: the free-stream angle, or for an unsteady base flow
   BaseFlow::getAlpha / getMag after setAlphaMag at the current time
   (ibpm.cc:347-351, BaseFlow.cc:111-125).  Returns 1 for an unsteady base
   flow, 0 otherwise, -1 on error. */
int ibpm_batch_base_flow(ibpm_batch* b, int flow, double* alpha, double* mag);
/* CG iterations of each substep of the last step (-1 when Cholesky is used). */
int ibpm_batch_cg_iters(ibpm_batch* b, int flow, int* h_iters, int max_substeps);
double ibpm_batch_time(ibpm_batch* b);
/* Base state k (vorticity h_omega, flux h_q of the flow's grid) of a linear
   model: LinearizedIBSolver / AdjointIBSolver x0 (k = 0), the k-th phase of
   LinearizedPeriodicIBSolver (k < period); reference IBSolver.h:92-187,
   ibpm.cc:190-219.  Before ibpm_batch_init. */
int ibpm_batch_set_base_state(ibpm_batch* b, int flow, int k, const double* h_omega, const double* h_q);
/* SFD: the filtered vorticity (SFDSolver::_xhat.omega) of a flow, read or
   written (loadFilteredState / saveFilteredState, IBSolver.cc:281-311). */
int ibpm_batch_get_filtered(ibpm_batch* b, int flow, double* h_omega);
int ibpm_batch_set_filtered(ibpm_batch* b, int flow, const double* h_omega);
/* Capture one timestep as a hipGraph and replay it (stationary bodies). */
int ibpm_batch_use_graph(ibpm_batch* b, int enable);
/* Cholesky factor cache of a stationary-body flow, named before ibpm_batch_init:
   init reads each substep's factor from <basename>_01.cholesky, _02, ... when the
   file matches (size and alpha*beta) and otherwise builds the factor and writes
   the file -- the reference's load-else-init-and-save (src/ibpm.cc:291-296,
   IBSolver::load/save src/IBSolver.cc:77-108, file format CholeskySolver::load/save
   src/CholeskySolver.cc:113-197).  basename NULL or "" turns the cache off. */
int ibpm_batch_cholesky_cache(ibpm_batch* b, int flow, const char* basename);
/* Number of substep factors init read from cache files (0 = built). */
int ibpm_batch_cholesky_loaded(ibpm_batch* b, int flow);

/* ---- restart files ---------------------------------------------------------
   The reference's binary restart layout (State::save / State::load,
   src/State.cc:65-223).  ibpm_state_load with null q/omega/f reads only the
   header, so buffers can be sized first. */
int ibpm_state_save(const char* path, int nx, int ny, int ngrid, double dx, double x0, double y0, int npts,
                    const double* q, const double* omega, const double* f, int timestep, double time);
int ibpm_state_load(const char* path, int* nx, int* ny, int* ngrid, double* dx, double* x0, double* y0,
                    int* npts, double* q, double* omega, double* f, int* timestep, double* time);
/* Resume: sets a flow's vorticity and force, the batch clock (time, step),
   then refreshState (ibpm.cc:250-305 loads an initial condition this way). */
int ibpm_batch_set_state(ibpm_batch* b, int flow, const double* h_omega, const double* h_f, double time,
                         int timestep);

/* ---- learned CG warm start (moving bodies) ---------------------------------
   No counterpart in the reference: a per-boundary-point network (13 -> 32 tanh
   -> 2, FP64) predicts the CG start from the force history, anchored on the
   force of the same RK3 substep one step earlier (the reference starts from
   the previous substep, src/ProjectionSolver.cc:72, whose beta differs); CG
   then runs to the same absolute residual tolerance, so only iteration
   counts change.
   mode: 0 zero start, 1 previous force (reference), 2 learned. */
int ibpm_batch_set_warm_start(ibpm_batch* b, int mode);
/* Record (features, converged increment) samples of every substep during
   ibpm_batch_advance into a device data set of `capacity` points. */
int ibpm_batch_collect(ibpm_batch* b, int enable, long capacity);
long ibpm_batch_num_samples(ibpm_batch* b);
/* Full-batch Adam (iters steps, learning rate lr) on the collected samples,
   with the gradient and update computed by HIP kernels; *loss = final loss. */
int ibpm_batch_train(ibpm_batch* b, int iters, double lr, double* loss);
/* Copies the network parameters out (h_get) and/or in (h_set); returns the count. */
int ibpm_batch_nn_params(ibpm_batch* b, double* h_get, const double* h_set);
/* Mean-squared-error loss and its gradient over n host samples (X: n x 13,
   Y: n x 2), computed by the training kernels (gradient checks). */
int ibpm_nn_loss_grad(ibpm_ctx* ctx, const double* h_params, const double* h_x, const double* h_y, long n,
                      double* h_grad, double* loss);

/* ---- preconditioned CG force solve (moving bodies) -------------------------
   No counterpart in the reference, which runs plain CG
   (src/ConjugateGradientSolver.cc:35-65).  mode 0: plain CG.  mode 1: CG
   preconditioned by
     H = (I - S G Y^T) P (I - Y G S^T) + S G S^T,  G = (S^T Y)^-1,
   where P is the FP64 Cholesky inverse of M at an earlier body position
   (built like the stationary-body factor and rebuilt once the extra
   iterations since the last build exceed the cost of a build, one M
   application per unknown), scaled by beta to each RK substep, and (S, Y=M S)
   are the first `pairs` (0..8) search directions of the previous solve of the
   same substep: a low-rank update learned from the solver's own Krylov data.
   Every solve ends on the unpreconditioned test r.r <= tol^2 of plain CG, and
   plain CG finishes any solve in which the preconditioner breaks down, so the
   answer is the plain solver's to within its tolerance. */
int ibpm_batch_set_precond(ibpm_batch* b, int mode, int pairs);
/* How P is built (no counterpart in the reference).  levels 0: from M itself.
   levels 1: from M's finest-level form, E L0^-1 A0^-1 (beta B) with L0^-1 the
   level-0 Poisson solve under zero boundary values, symmetric positive
   definite like M and a fraction of its cost per column (needs the local
   Helmholtz path, ibpm_batch_set_local_helmholtz; without it P comes from M).
   levels 2: from the kernel table, M = beta E G E^T with the level-0
   flux-to-flux operator G inside M read off its response to one unit flux of
   each component at the level's centre (a translation-invariant lattice
   kernel, computed once per grid), at the current points: no M applications.
   levels 3: as 2, and before every solve one Newton-Schulz step moves P
   toward the inverse of the table's M at that substep's points, P <- P (2 I
   - M~ P), two batched FP64 GEMMs, so that P follows the moving bodies; the
   rebuild rule then counts only iterations more than two above a substep's
   first solve after a build.
   P only steers the iterations: the solves still end on M's own residual.
   rebuild_cost: a build's cost in M applications per unknown, which sets the
   rebuild rule's break-even (1 for levels 0).  Marks P for a rebuild. */
int ibpm_batch_set_precond_build(ibpm_batch* b, int levels, double rebuild_cost);
/* Inspection of the preconditioner (no counterpart in the reference): for a
   moving flow, h_mtable (2 nb x 2 nb, row-major, may be null) receives the
   kernel-table M of substep `substep` at the flow's current points, and h_p
   (may be null) the dense P in use scaled to that substep (an error if none
   is built).  Synchronises the batch's stream. */
int ibpm_batch_precond_matrices(ibpm_batch* b, int flow, int substep, double* h_mtable, double* h_p);
/* Force-solve counters since init: [0] iterations summed over flows and
   substeps, [1] solves (flow-substeps), [2] batched M applications in solves,
   [3] batched M applications spent building preconditioners, [4] builds,
   [5] solves finished by plain CG after a breakdown, [6] 1 when some M was not
   positive definite and preconditioning stopped, [7] host microseconds in the
   per-substep kinematics and regulariser updates (GPU idle), [8] host
   microseconds in solver polls (stream synchronisation and state read-back),
   [9] batched M applications that took the local Helmholtz path,
   [10] wall microseconds in preconditioner builds, [11] of those, host
   microseconds in the factorisations and inverse assemblies, [12] refreshes
   of P (build 3).
   Returns the count written. */
int ibpm_batch_solver_stats(ibpm_batch* b, long* stats, int nstats);
/* No counterpart in the reference (its M application and its step solve
   A^-1 = (1 + alpha Lap)^-1 on every level with the DST,
   ProjectionSolver.cc:57-101, EllipticSolver.cc:40-70).  on = 1 (the
   default): in moving-body batches, A^-1 of a smeared force beta B f (inside
   each M application of the force solve, and in the step's
   omega = omega* - A^-1 c) is computed by Jacobi sweeps over a box around the
   body on each level, wherever alpha / dx^2 is small enough that the solution
   decays below 1e-18 of its maximum before the box reaches a level's boundary
   or the ring of the finer level (the substep falls back to the DST solve
   otherwise); the result equals the DST solve's to rounding.  The L^-1 solve
   inside M stays a DST solve, run over the parts of each level whose psi is
   read.  on = 0: the DST solves always. */
int ibpm_batch_set_local_helmholtz(ibpm_batch* b, int on);

/* Transform precision of the batch's elliptic solves (no counterpart in the
   reference, whose solves are FP64): 0 (the default) FP64 everywhere; 1 the
   step's own solves (A^-1 a, L^-1 omega*, A^-1 c, the refresh's L^-1) run FP32
   transforms and transposed intermediates, with FP64 right-hand sides, boundary
   terms, eigenvalue scaling and fields, so omega, psi and q stay FP64; 2 also
   the solves inside the force solve's M applications (CG's vectors, its
   residual test and the Cholesky factor stay FP64).  Each reduced-precision
   stage perturbs the FP64 answer by about the FP32 rounding of one solve. */
int ibpm_batch_set_precision(ibpm_batch* b, int mode);
/* Modes 3 and 4 of ibpm_batch_set_precision take, for the step's solves on
   square levels of N <= 511, the dense sine-matrix GEMM-DST on MFMA
   (split-BF16 products, FP32 accumulation): mode 4 wherever it applies, mode 3
   where the library's autotuner, timing both on the batch's own jobs at the
   call, finds it faster than the FP32 radix passes.  This reports per N the
   choice (use_gemm 1) and the timed microseconds per solve of the group's
   jobs (0 when not timed); returns the number of lengths written. */
int ibpm_batch_precision_report(ibpm_batch* b, int* n, int* use_gemm, double* t_radix_us, double* t_gemm_us,
                                int max);

/* Measurement: advances nsteps with a HIP event pair around every kernel
   launch on the batch stream and returns per kernel class the launch count,
   total milliseconds and algorithmic HBM bytes.  Returns the number of
   classes; ibpm_kernel_class_name names them. */
int ibpm_batch_kernel_stats(ibpm_batch* b, int nsteps, int* counts, double* ms, double* bytes, int nclass);
const char* ibpm_kernel_class_name(int cls);

#ifdef __cplusplus
}
#endif
#endif
