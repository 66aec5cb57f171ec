// ops.h -- per-flow descriptors and launchers of the projection-step kernels.
#pragma once

#include <functional>
#include "common.h"

namespace ibpmg {

// Stencil kernels (nonlinear term, right-hand side, flux refresh).
struct FlowOps {
    int nx, ny, L;
    double dx;                     // finest spacing
    double *q, *omega, *U, *V, *cross, *a, *nprev, *psi;
    double lap_coef, an, bn, dt;   // a = w_coef omega + dt (lap_coef Lap + an N + bn Nprev)
    double w_coef;
    int use_nprev, nprev_valid;
    double ubase, vbase;           // uniform base flow (magnitude 1, angle alpha)
    // unsteady base flow (ubf = 1): bf[0..2] = u, v, w of the current base
    // flux, X (u - w (y - yc)) d and Y (v + w (x - xc)) d at edge (x, y)
    int ubf;
    const double* bf;
    double xc, yc, xoff, yoff;     // centre of rotation; level-0 grid offsets
    // model (IBSolver.cc:203-246): 0 nonlinear, 1 linear, 2 adjoint, 3 linear
    // periodic, 4 SFD.  q0b, w0b: base state of the linear models; sff: the
    // adjoint's CrossProduct(q0, q), coarsified; what, sfd_prev: SFD's
    // filtered vorticity and its previous right-hand side (updated in place)
    int model;
    const double *q0b, *w0b, *sff;
    double *what, *sfd_prev;
    double sfd_chi, sfd_inv_delta;
    int sfd_valid;                 // what holds the filtered state (else: zero in N, omega at the update)
    // nonlinear-model flows of a batch: a and N are not formed at the nodes
    // strictly inside a coarse level's central region, which no kernel reads
    // (the Helmholtz solve and its coarsify chain read the coarsified
    // right-hand side there, and N there only feeds the next substep's a there)
    int skip_central;
};

// CrossProduct(Flux, Flux) of one flow before its coarsify: out = X(q1) Y(q2)
// - X(q2) Y(q1) at the nodes of every level.
struct CrossFFOps {
    int nx, ny, L;
    double dx;
    const double *q1, *q2;
    double* out;
};

// Regulariser gathers (E, and beta Curl E^T) of one flow.
struct RegOps {
    int nx, ny, L, nb;
    double dx, inv_dx, bydx2, beta;
    // E: per boundary component k, level-0 flux indices and weights, ascending
    const int* e_ptr;
    const int* e_f;
    const double* e_w;
    const double* psi;             // streamfunction whose curl is interpolated
    const double* b;               // subtracted constraint (or null)
    double* out;                   // 2 nb outputs
    // E^T then Curl: flux list (CSR over components, ascending) and nodes
    int nnodes;
    const int* t_ptr;
    const int* t_k;
    const double* t_w;
    const int4* node_nbr;          // flux-list slots of Y(i,j), Y(i-1,j), X(i,j-1), X(i,j)
    const int* node_idx;           // level-0 scalar index of each node
    const double* fin;             // boundary vector smeared
    double* cout;                  // scalar field written at the nodes
    int nold;
    const int* old_idx;            // nodes of the previous association to clear
    const double* skip;            // null, or the flow's force-solve flags
};

// Standalone regulariser (E and E^T on the finest level's flux field).
struct RegFluxOps {
    int nb, nslots;
    double dx;
    const int *e_ptr, *e_f;        // E: per component, ascending flux indices
    const double* e_w;
    const int *t_ptr, *t_k;        // E^T: per flux slot, components in ascending order
    const double* t_w;
    const int* flux_of_slot;       // level-0 flux index of each slot
};

struct MinvOps {
    int n;
    const double* minv;            // n x n row-major
    const double* rhs;
    double* x;                     // x = scale * minv * rhs
    double scale;
    const double* skip;            // null, or flags: skip[0] or skip[1] nonzero -> no-op
};

struct CgOps {
    int n;
    double *x, *r, *d, *q;
    const double* b;
    double* state;                 // delta, tol^2, iterations, converged
};

// Preconditioned CG on M f = rhs.  The preconditioner is
//   H = (I - S G Y^T) P (I - Y G S^T) + S G S^T,   G = (sym(S^T Y))^-1,
// with P = scale * (dense FP64 inverse of M at an earlier configuration) and
// (S, Y = M S) the first search directions of the previous solve of the same
// substep (a limited-memory low-rank update; no pairs -> H = P).  The
// residual test is the unpreconditioned r.r <= tol^2 of plain CG.
constexpr int kLmpMax = 8;
struct PcgOps {
    int n, kmax;                   // unknowns; pairs kept for the next solve
    double *x, *r, *d, *q;         // iterate, residual, direction, M d
    double *t, *u;                 // input and output of the dense P
    const double* b;
    double* state;                 // rho, tol^2, iterations, converged, breakdown, nk, delta
    const double *rs, *ry;         // pairs of the previous solve (read), count *rcnt
    const int* rcnt;
    double *ws, *wy;               // pairs of this solve (written), count *wcnt
    int* wcnt;
    double* g;                     // kLmpMax x kLmpMax inverse Gram matrix of the read pairs
    double* c;                     // kLmpMax: G S^T r of the current residual
};

// The unit vector of one flow's preconditioner build (n unknowns, one spare
// slot): column j moves the 1 from j - 1 to j (no 1 for j >= n).
struct UnitOps {
    double* u;
    int n;
};

// One contiguous copy of n 4-byte words, or n zero words with src null
// (k_copy_pieces: many copies in one launch instead of one runtime call each).
struct CopyPiece {
    const unsigned int* src;
    unsigned int* dst;
    long n;
};

// The preconditioner's matrix from a kernel table.  M = E G E^T with G the
// level-0 flux-to-flux operator inside M (Curl, A^-1, L^-1, Curl); on an
// unbounded lattice G is translation invariant, so G(a, b) is read from the
// response to one unit flux of b's component at the source node nearest to
// b's row of E, at a's node shifted by the source minus b's node (zero past
// the level).  Entry (r, c) = scale * sum over the E entries a of row r and b of
// row c of w_a w_b G(a, b); the upper triangle is formed and mirrored.  The
// rows come in separable form (weights u[a] v[b] over a rectangle of nodes),
// so the double sum is one over the 5 x 5 node offsets of the two rectangles,
// weighted by the correlations of their row and of their column factors.
struct PcAsmOps {
    int n, nx, ny;                 // unknowns (X components, then Y); level-0 cells
    int nsrc;                      // source nodes of the table
    const int* src;                // their (i, j)
    const double* sep;             // E rows in separable form, kESep per component (RegLists::e_sep)
    const double* tab;             // per source: level-0 flux responses to a unit X flux, then a unit Y flux
    double scale;
    double* out;                   // n x n, row-major
};

// C = alpha A B + beta D for n x n row-major FP64 matrices (D may be null);
// with sym, C is symmetric: the tiles on and above the diagonal are formed
// and mirrored.
struct GemmOps {
    int n;
    const double *A, *B, *D;
    double* C;
    double alpha, beta;
};

struct ForceOps {
    int nb;
    const double* f;
    double* hist;                  // (xF, yF) per step
    int* count;                    // steps recorded so far (device counter)
};

// Whether a pass launch of FFT length M over row_pairs row pairs (summed
// over its jobs) uses the register-held transform; the three calls below take
// that choice.
bool pass_in_registers(int M, long row_pairs);
int blocks_for_pass(const EllJob& J, int pass, bool regs);
int lds_for_pass(const EllJob& J, int pass, bool regs);
int lds_for_pass32(const EllJob& J, int pass);
void launch_ell_pass32(int pass, const EllJob* d_jobs, int njobs, int max_blocks, int lds_bytes, hipStream_t s,
                       bool force_solve, int M);
void launch_ell_pass(int pass, const EllJob* d_jobs, int njobs, int max_blocks, int lds_bytes, hipStream_t s,
                     bool force_solve, int M, bool regs);
void launch_coarsify(const CoarsifyJob* d_jobs, int njobs, int max_rows, hipStream_t s, bool force_solve = false);
// Split-BF16 GEMM-DST level solve of every job (square levels of n = N = nx - 1
// <= 511 unknowns per side, planes in the jobs' T1 / T2): Sh, Sl the split sine
// matrix of n with gemm_np(n) columns per row.
constexpr int gemm_np_host(int n) { return (n + 63) / 64 * 64; }
void launch_gemm_dst(const EllJob* d_jobs, int njobs, int n, const __bf16* Sh, const __bf16* Sl, hipStream_t s);
// The nsweeps Jacobi sweeps of the local Helmholtz solves (HelmJob), several
// per launch; the last launch writes each job's out.  max_tiles: helm_tiles
// of the largest box.  timed_launch(fn) runs fn, the launch (a caller may
// time it).
int helm_tiles(int h, int w);
int helm_launches(int nsweeps);   // launches of launch_helm_sweeps for nsweeps sweeps
// Row pairs per workgroup of register-held pass `pass` (1, 2, 3).
int r16_pass_group(int pass);
// eo = ei - out over each job's box; max_blocks: kThreads-point blocks of the
// largest box.
void launch_helm_epilogue(const HelmJob* d_jobs, int njobs, int max_blocks, hipStream_t s);
// out = tmp = 0 over each job's box.
void launch_helm_clear(const HelmJob* d_jobs, int njobs, int max_blocks, hipStream_t s);
void launch_helm_sweeps(const HelmJob* d_jobs, int njobs, int max_tiles, int nsweeps, hipStream_t s,
                        const std::function<void(const std::function<void()>&)>& timed_launch);
// The right-hand side's coarsify chain at level lev (>= 1) over each job's
// support (HelmJob); max_blocks: kThreads-point blocks of the largest support.
void launch_helm_coarsify(const HelmJob* d_jobs, int njobs, int max_blocks, int lev, hipStream_t s);
void launch_vel(const FlowOps* d, int nflows, long max_pts, hipStream_t s);
void launch_cross(const FlowOps* d, int nflows, long max_nf, int lev, hipStream_t s);
void launch_cross_ff(const CrossFFOps* d, int nflows, long max_pts, hipStream_t s);
void launch_rhs(const FlowOps* d, int nflows, long max_pts, hipStream_t s);
// max_cells_all: (nx + 1)(ny + 1) L of the largest flow
void launch_curl_q(const FlowOps* d, int nflows, long max_cells_all, hipStream_t s);
void launch_interp(const RegOps* d, int nflows, int max_comp, hipStream_t s);
void launch_smear_curl(const RegOps* d, int nflows, int max_nodes, hipStream_t s);
void launch_clear_nodes(const RegOps* d, int nflows, int max_nodes, hipStream_t s);
void launch_minv(const MinvOps* d, int nflows, int max_n, hipStream_t s);
void launch_cg(const CgOps* d, int nflows, int mode, hipStream_t s);
void launch_pcg_begin(const PcgOps* d, int nflows, hipStream_t s);
void launch_pcg_step(const PcgOps* d, int nflows, hipStream_t s);
void launch_pcg_dir(const PcgOps* d, int nflows, int begin, hipStream_t s);
void launch_forces(const ForceOps* d, int nflows, hipStream_t s);
void launch_unit_column(const UnitOps* d, int nflows, int j, hipStream_t s);
void launch_copy_pieces(const CopyPiece* d, int npieces, hipStream_t s);
void launch_pc_assemble(const PcAsmOps* d, int nflows, int max_n, hipStream_t s);
void launch_dgemm(const GemmOps* d, int nflows, int max_n, bool sym, hipStream_t s);
void launch_reg_to_flux(const RegFluxOps* d, int nslots, const double* f, double* q, hipStream_t s);
void launch_reg_to_boundary(const RegFluxOps* d, int nb, const double* q, double* f, hipStream_t s);

}  // namespace ibpmg
