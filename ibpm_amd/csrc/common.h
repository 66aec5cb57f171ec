// common.h -- shared definitions of the gfx950 projection-step library.
//
// Device data layout (HBM), per flow:
//   Scalar : [lev][i=1..nx-1][j=1..ny-1], j contiguous (reference Scalar.cc:81-89)
//   Flux   : [lev][X (nx+1)*ny | Y nx*(ny+1)]       (reference Flux.cc:42-50)
//   BoundaryVector : [X_0..X_{nb-1}, Y_0..Y_{nb-1}]  (reference BoundaryVector.h:65-69)
// identical to the reference so host<->device copies are flat memcpys.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ibpmg {

// Field pointers held in job descriptors are generic (flat) pointers; kernels
// address them as global memory so that loads and stores are global
// instructions, which do not share the LDS counter and may move 16 bytes at
// 8-byte alignment.
#define IBPM_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ const IBPM_GLOBAL double* gin(const double* p) { return (const IBPM_GLOBAL double*)p; }
__device__ __forceinline__ IBPM_GLOBAL double* gout(double* p) { return (IBPM_GLOBAL double*)p; }

constexpr int kThreads = 256;          // workgroup size of the stencil kernels (4 wave64)
#ifndef IBPM_FFT_THREADS
#define IBPM_FFT_THREADS 256
#endif
#ifndef IBPM_FFT_NP_MAX
#define IBPM_FFT_NP_MAX 1
#endif
constexpr int kFftThreads = IBPM_FFT_THREADS;   // workgroup size of the DST passes
constexpr int kFftNpMax = IBPM_FFT_NP_MAX;      // most row pairs per DST workgroup
constexpr int kFftLds = 2048;          // complex doubles staged in LDS per DST block (32 KB)
constexpr int kMaxStages = 16;

// Plan of one complex FFT of length M = 2(N+1) used for a DST-I of length N.
struct FftPlan {
    int N;          // DST length
    int M;          // FFT length
    int np;         // row pairs per workgroup (np * M <= kFftLds)
    int nst;        // number of Stockham stages
    int rad[kMaxStages];
    const double2* tw;   // tw[k] = exp(-2 pi i k / M), k < M
    const double2* tw16; // M = 1024, 2048: the second radix-16 stage's w^((M/256) jm r) at 16 r + jm (r, jm < 16)
};

// One level of one flow in a multi-domain elliptic solve.
struct EllJob {
    int nx, ny;              // cells; rows i=1..nx-1, columns j=1..ny-1
    int lev, ngrid;
    int kind;                // 0 Poisson, 1 Helmholtz
    int epilogue;            // 0: u = v ; 1: u = W - v
    double in_scale;         // multiplies the right-hand side (+1 or -1)
    double bc_coef;          // 1/dx^2 (Poisson) or alpha/dx^2 (Helmholtz)
    double alpha, bydx2, norm;
    const double* f;         // level slice of the right-hand side
    const double* R;         // level slice of the coarsified rhs (central region), lev >= 1
    const double* uc;        // level slice lev+1 of the solution (BC source), lev < ngrid-1
    const double* bcring;    // explicit BC ring 2(nx+ny) instead of uc (single-level API)
    double* T1;              // scratch (ny-1) rows of nx-1 values, row stride ld1
    double* T2;              // scratch (nx-1) rows of ny-1 values, row stride ld2
    int ld1, ld2;            // >= nx-1, ny-1; padded (t1_ld) where the scratch is the library's own
    double* u;               // level slice of the solution
    double* u2;              // null, or a second output of the same values (pass 3)
    const double* W;         // epilogue 1 operand (level slice)
    double* vside;           // epilogue solves, levels >= 1: v (before the epilogue) stored at
                             // the points the next finer level's getBC reads (level slice)
    const double* cosx;      // cos(pi i / nx), i = 1..nx-1
    const double* cosy;      // cos(pi j / ny), j = 1..ny-1
    FftPlan pj;              // transform along j (length ny-1)
    FftPlan pi;              // transform along i (length nx-1)
    const double* skip;      // null, or the flow's solver flags (see skipped())
    // Pass 3 over part of the level (the register-held pass 3 only): row pairs
    // in [full_lo, full_hi] and pairs fq0, fq1 are transformed; of the other
    // rows, those in [ring_lo, ring_hi] (1-based) get u at columns rj0 and rj1
    // alone, by direct sums against ringsin (2 sin(pi l j / ny), l = 1..ny-1,
    // for j = rj0 then rj1); the rest is not written.
    int full_lo, full_hi, fq0, fq1, ring_lo, ring_hi, rj0, rj1;
    const double* ringsin;
    // Pass 1 (register-held): rows i outside [sup_lo, sup_hi] (1-based) have
    // no interior right-hand side; a workgroup whose rows are all such rows,
    // none of them 1 or nx-1, writes their transform in closed form from the
    // two boundary terms, sin1[k - 1] = 2 sin(pi k / ny) (null: never).
    int sup_lo, sup_hi;
    const double* sin1;
    // Pass 1 (register-held): in the central region R is read inside this box
    // only and taken as 0 outside it, where the solve's coarsify chain, run
    // over the box alone, left other solves' values.
    int rb_i0, rb_i1, rb_j0, rb_j1;
    // Pass 2 (register-held): rows of T2 (0-based i) that pass 3 reads; the
    // others are not written.
    int t2_lo, t2_hi;
};

// Local solve of A u = r, A = 1 + alpha Lap (the force solve's A^-1 of one
// level of one flow) by Jacobi sweeps over a box: u <- (r - c S u) / (1 - 4 c),
// c = alpha / dx^2, S the sum of the four neighbours, u = 0 off the box.
struct HelmJob {
    int nx, ny, lev;
    int i0, i1, j0, j1;      // box (1-based, inclusive), inside 2..nx-2 x 2..ny-2
    int s0, s1, t0, t1;      // support of the right-hand side on this level (zero off it)
    int fs0, fs1, ft0, ft1;  // support on the next finer level (lev >= 1)
    double c, inv_diag;      // alpha / dx_lev^2 and 1 / (1 - 4 c)
    double in_scale;
    const double* f;         // level slice of the right-hand side
    double* R;               // level slice of the coarsified right-hand side (central region, lev >= 1)
    const double* f_fine;    // the next finer level's slices of f and R (lev >= 1)
    const double* R_fine;
    double* out;             // level slice written by the last sweep
    double* tmp;             // level slice of the other sweep buffer
    const double* skip;      // null, or the flow's solver flags
    double* eo;              // epilogue (launch_helm_epilogue): eo = ei - out over the box
    const double* ei;
};

// Coarsify job: one level (lev >= 1) of one flow (reference Scalar.cc:46-63).
struct CoarsifyJob {
    int nx, ny, lev;
    const double* f;         // whole field (all levels)
    double* R;               // whole coarsified buffer (all levels; central regions used)
    const double* skip;
};

// Force-solve flags of a flow: flags[0] converged, flags[1] broken down.  A
// kernel given a flow's flags does nothing for that flow once either is set,
// so the batched M applications of finished flows cost no bandwidth.
// The flags are uniform over a launch and written only by earlier launches
// (a launch starts with its caches invalidated), so both are read at once
// through the scalar cache: a workgroup's first loads then wait on no vector
// memory round trip, where the two short-circuited vector loads took two.
__device__ inline bool skipped(const double* flags) {
    if (!flags) return false;
    typedef const __attribute__((address_space(4))) double ConstDouble;
    ConstDouble* c = (ConstDouble*)flags;
    const double converged = c[0], broken = c[1];
    return (converged != 0.0) | (broken != 0.0);
}

__host__ __device__ inline int nx_ext(int nx) { return nx / 4; }   // Grid.h:89-91, zero shift

// Row stride of the transposed scratch rows of n values: a multiple of 8
// doubles, so the 64-byte pieces that the register-held pass 1 writes per
// column start on 64-byte boundaries (one write request each, not two).
inline int t1_ld(int n) { return (n + 7) & ~7; }

__device__ inline long sidx(int nx, int ny, int i, int j) {   // within a level slice
    return (long)(i - 1) * (ny - 1) + (j - 1);
}

}  // namespace ibpmg

#define IBPM_HIP_CHECK(expr)                                                        \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ibpmg::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,             \
                             hipGetErrorString(e_));                                \
            return -1;                                                              \
        }                                                                           \
    } while (0)

namespace ibpmg {
void set_error(const char* fmt, ...);
}
