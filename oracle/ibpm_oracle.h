/*
 * ibpm_oracle.h -- CPU FP64 restatement of the ibpm projection-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (ibpm_amd/, libibpm_gpu.so) never links or calls it.
 *
 * Layouts follow the reference exactly (Array.h: last index contiguous):
 *   Scalar  : [lev][i=1..nx-1][j=1..ny-1]            (Scalar.cc:81-89)
 *   Flux    : [lev][ X block (nx+1)*ny | Y block nx*(ny+1) ]
 *             X(i,j) -> i*ny + j,  Y(i,j) -> (nx+1)*ny + i*(ny+1) + j
 *                                                     (Flux.cc:42-50, Flux.h:176-185)
 *   BC ring : 2*(nx+ny) doubles, left/top/right/bottom accessors of BC.h:22-79
 *   BoundaryVector : [X_0..X_{nb-1}, Y_0..Y_{nb-1}]   (BoundaryVector.h:65-69)
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference's root).  Pinning of each function is listed in the header of
 * tests/test_oracle.py.
 */
#ifndef IBPM_ORACLE_H
#define IBPM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- transforms --------------------------------------------------------- */
/* Unnormalised 2-D DST-I (FFTW RODFT00 x RODFT00), FFT based.
   in/out: n1 x n2 row-major (may alias).  EllipticSolver2d.cc:57-71 */
void og_dst2d(int n1, int n2, const double* in, double* out);
/* Same transform by direct summation, O(n1 n2 (n1+n2)); pins og_dst2d. */
void og_dst2d_direct(int n1, int n2, const double* in, double* out);

/* ---- multi-domain scalar helpers ---------------------------------------- */
void og_coarsify(int nx, int ny, int ngrid, double* s);                 /* Scalar.cc:46-63 */
void og_getbc(int nx, int ny, int ngrid, const double* s, int lev, double* bc); /* Scalar.cc:91-127 */

/* ---- elliptic solves ----------------------------------------------------- */
/* kind 0 = Poisson (L u = f), 1 = Helmholtz ((1 + alpha L) u = f).
   Single level; bc may be NULL (zero boundary).  EllipticSolver2d.cc:87-185 */
void og_elliptic2d_solve(int kind, int nx, int ny, double dx, double alpha,
                         const double* f, const double* bc, double* u);
/* Multi-domain solve over all levels; f and u may alias.  EllipticSolver.cc:47-73 */
void og_elliptic_solve(int kind, int nx, int ny, int ngrid, double dx, double alpha,
                       const double* f, double* u);

/* ---- grid operators (VectorOperations.cc) -------------------------------- */
void og_curl_flux(int nx, int ny, int ngrid, double dx, const double* q, double* w);   /* :33-54 */
void og_curl_scalar(int nx, int ny, int ngrid, const double* w, double* q);           /* :63-123 */
void og_laplacian(int nx, int ny, int ngrid, double dx, const double* w, double* out);/* :133-140 */
void og_laplacian2d_bc(int nx, int ny, double dx, const double* f, const double* bc, double* g); /* :176-214 */
void og_flux_to_xvel(int nx, int ny, int ngrid, double dx, const double* q, double* u); /* :488-577 */
void og_flux_to_yvel(int nx, int ny, int ngrid, double dx, const double* q, double* v); /* :579-665 */
void og_xvel_to_flux(int nx, int ny, int ngrid, double dx, const double* u, double* q); /* :669-791 */
void og_yvel_to_flux(int nx, int ny, int ngrid, double dx, const double* v, double* q); /* :793-915 */
void og_cross_product(int nx, int ny, int ngrid, double dx, const double* q, const double* w,
                      double* out);                                                    /* :438-457 */
void og_cross_flux_flux(int nx, int ny, int ngrid, double dx, const double* q1, const double* q2,
                        double* out);   /* VectorOperations.cc:465-486 */
void og_nonlinear(int nx, int ny, int ngrid, double dx, const double* q, const double* w,
                  double* n_out);                                  /* IBSolver.cc:205-209 */
void og_linear_N(int nx, int ny, int ngrid, double dx, const double* q0, const double* w0, const double* q,
                 const double* w, double* n_out);                  /* IBSolver.cc:211-216 */
void og_adjoint_N(int nx, int ny, int ngrid, double dx, const double* q0, const double* w0, const double* q,
                  double* n_out);                                  /* IBSolver.cc:218-222 */

/* Multi-domain inner products, VectorOperations.cc:223-405 */
double og_ip_scalar(int nx, int ny, int ngrid, double dx, const double* f, const double* g);
double og_ip_flux(int nx, int ny, int ngrid, const double* p, const double* q);

/* ---- regulariser (Regularizer.cc) ---------------------------------------- */
/* Builds the association list in reference order (dir -> point -> flux index).
   Returns the number of associations; arrays must hold 2*18*npts entries. */
int og_reg_build(int nx, int ny, double dx, double xoff, double yoff, int npts,
                 const double* px, const double* py, int* bidx, int* fidx, double* w); /* :54-90 */
void og_reg_toflux(int nx, int ny, int ngrid, double dx, int nassoc, const int* bidx,
                   const int* fidx, const double* w, const double* f, double* q);        /* :92-109 */
void og_reg_toboundary(int npts, double dx, int nassoc, const int* bidx, const int* fidx,
                       const double* w, const double* q, double* f);                     /* :111-128 */

/* Points [x..,y..] and velocities [u..,v..] of the bodies in a .geom text after
   Geometry::moveBodies(t) (Geometry.cc:115, RigidBody.cc:586-609). */
int og_geom_points(const char* geom, double t, double* xy, double* uv, int cap);
int og_geom_sequence(const char* geom, int ntimes, const double* times, double* xy, double* uv, int cap);

/* ---- whole-solver simulation -------------------------------------------- */
typedef struct og_sim og_sim;
/* scheme: 0 Euler, 1 AB2, 2 RK3, 3 RK3b (Scheme.h:14-56).  alpha_deg: base-flow
   angle in degrees (ibpm.cc:163-166).  geom: .geom file text (Geometry.cc:143). */
og_sim* og_sim_create(int nx, int ny, int ngrid, double length, double xoff, double yoff,
                      double Re, double dt, int scheme, double alpha_deg,
                      const char* geom_text, double cg_tol, int force_cg);
void og_sim_destroy(og_sim* s);
const char* og_sim_error(void);
int  og_sim_set_ubf(og_sim* s);
/* Model of ibpm.cc:185-235: 0 nonlinear, 1 linear, 2 adjoint, 3 linear
   periodic (period base states), 4 SFD (gain chi, filter width delta); call
   before og_sim_init. */
int  og_sim_set_model(og_sim* s, int model, int period, double chi, double delta);
/* Base state k (vorticity, flux) of the linear models. */
int  og_sim_set_base_state(og_sim* s, int k, const double* w, const double* q);
/* SFD filtered vorticity: set (marked saved, as loadFilteredState) and read. */
int  og_sim_set_filtered(og_sim* s, const double* w);
int  og_sim_get_filtered(const og_sim* s, double* w);
int  og_baseflow_sequence(const char* geom, int nx, int ny, int ngrid, double length, double xoff, double yoff,
                          double alpha_deg, int ntimes, const double* times, double* q_out, double* am_out);
int  og_sim_num_points(const og_sim* s);
int  og_sim_uses_cg(const og_sim* s);
/* ibpm.cc:283-305: move bodies to t, regulariser update, build solvers, refresh */
int  og_sim_init(og_sim* s);
/* og_sim_init with identity Cholesky factors (CPU-baseline timing only). */
int  og_sim_init_timing(og_sim* s);
/* One full timestep (IBSolver.cc:142-150); returns 0 on success. */
int  og_sim_advance(og_sim* s);
/* Cd, Cl as printed by ibpm.cc:346-355 (2 x rotated net force). */
void og_sim_forces(const og_sim* s, double* cd, double* cl);
void og_sim_get_omega(const og_sim* s, double* out);
void og_sim_set_omega(og_sim* s, const double* in);
void og_sim_get_q(const og_sim* s, double* out);
void og_sim_get_f(const og_sim* s, double* out);
void og_sim_get_points(const og_sim* s, double* xy);
double og_sim_time(const og_sim* s);
/* CG iterations used in each substep of the last advance (-1 for Cholesky). */
int  og_sim_last_cg_iters(const og_sim* s, int* iters_per_substep);
/* Projection pieces for identity tests (ProjectionSolver.cc:57-101). */
void og_sim_projection_solve(og_sim* s, int substep, const double* a, const double* b,
                             double* omega, double* f);
void og_sim_M(og_sim* s, int substep, const double* f, double* y);
void og_sim_B(og_sim* s, int substep, const double* f, double* w);   /* beta * Curl(E^T f) */
void og_sim_C(og_sim* s, const double* w, double* f);                /* E Curl(L^-1(-w)) */
void og_sim_Ainv(og_sim* s, int substep, const double* a, double* w);
double og_sim_beta(const og_sim* s, int substep);
/* Cholesky factor of substep's M: diag (n) and strict lower L (n*n row-major). */
int  og_sim_cholesky(const og_sim* s, int substep, double* diag, double* lower);

#ifdef __cplusplus
}
#endif
#endif
