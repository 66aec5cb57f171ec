// ops.hip -- memory-bound stencils, regulariser gathers and force-solve
// kernels of the projection step (FP64).  Every kernel takes an array of
// per-flow descriptors and is launched with blockIdx.y = flow, so one launch
// advances a whole batch of flows of uneven size (blocks past a flow's extent
// exit at once).
#include "geometry.h"
#include "ops.h"

namespace ibpmg {

// ------------------------------------------------------------ accessors
struct Lay {
    int nx, ny, L;
    __device__ long ns() const { return (long)(nx - 1) * (ny - 1); }
    __device__ long nf() const { return 2L * nx * ny + nx + ny; }
    __device__ long nfx() const { return (long)(nx + 1) * ny; }
    __device__ long s(int lev, int i, int j) const { return lev * ns() + (long)(i - 1) * (ny - 1) + (j - 1); }
    __device__ long fx(int lev, int i, int j) const { return lev * nf() + (long)i * ny + j; }
    __device__ long fy(int lev, int i, int j) const { return lev * nf() + nfx() + (long)i * (ny + 1) + j; }
};

// getBC of level lev from level lev+1 (reference Scalar.cc:91-127); zero on
// the coarsest level (VectorOperations.cc:74-77).
__device__ inline double bcb(const Lay& g, const double* F, int lev, int i) {
    if (lev == g.L - 1) return 0.0;
    const int ii = i / 2 + nx_ext(g.nx), jj = nx_ext(g.ny);
    const double a = F[g.s(lev + 1, ii, jj)];
    return (i & 1) ? 0.5 * (a + F[g.s(lev + 1, ii + 1, jj)]) : a;
}
__device__ inline double bct(const Lay& g, const double* F, int lev, int i) {
    if (lev == g.L - 1) return 0.0;
    const int ii = i / 2 + nx_ext(g.nx), jj = g.ny / 2 + nx_ext(g.ny);
    const double a = F[g.s(lev + 1, ii, jj)];
    return (i & 1) ? 0.5 * (a + F[g.s(lev + 1, ii + 1, jj)]) : a;
}
__device__ inline double bcl(const Lay& g, const double* F, int lev, int j) {
    if (lev == g.L - 1) return 0.0;
    const int ii = nx_ext(g.nx), jj = j / 2 + nx_ext(g.ny);
    const double a = F[g.s(lev + 1, ii, jj)];
    return (j & 1) ? 0.5 * (a + F[g.s(lev + 1, ii, jj + 1)]) : a;
}
__device__ inline double bcr(const Lay& g, const double* F, int lev, int j) {
    if (lev == g.L - 1) return 0.0;
    const int ii = g.nx / 2 + nx_ext(g.nx), jj = j / 2 + nx_ext(g.ny);
    const double a = F[g.s(lev + 1, ii, jj)];
    return (j & 1) ? 0.5 * (a + F[g.s(lev + 1, ii, jj + 1)]) : a;
}

// Curl(Scalar -> Flux) at one flux (reference VectorOperations.cc:63-123).
__device__ inline double curl_x(const Lay& g, const double* F, int lev, int i, int j) {
    if (i == 0) return bcl(g, F, lev, j + 1) - bcl(g, F, lev, j);
    if (i == g.nx) return bcr(g, F, lev, j + 1) - bcr(g, F, lev, j);
    if (j == 0) return F[g.s(lev, i, 1)] - bcb(g, F, lev, i);
    if (j == g.ny - 1) return bct(g, F, lev, i) - F[g.s(lev, i, g.ny - 1)];
    return F[g.s(lev, i, j + 1)] - F[g.s(lev, i, j)];
}
__device__ inline double curl_y(const Lay& g, const double* F, int lev, int i, int j) {
    if (j == 0) return bcb(g, F, lev, i) - bcb(g, F, lev, i + 1);
    if (j == g.ny) return bct(g, F, lev, i) - bct(g, F, lev, i + 1);
    if (i == 0) return bcl(g, F, lev, j) - F[g.s(lev, 1, j)];
    if (i == g.nx - 1) return F[g.s(lev, g.nx - 1, j)] - bcr(g, F, lev, j);
    return F[g.s(lev, i, j)] - F[g.s(lev, i + 1, j)];
}

// FluxToXVelocity / FluxToYVelocity at one node (VectorOperations.cc:488-665).
// The central nodes of coarse levels are never read downstream and are 0.
__device__ inline double xvel(const Lay& g, const double* q, double dx0, int lev, int i, int j) {
    if (lev == 0) return (q[g.fx(0, i, j)] + q[g.fx(0, i, j - 1)]) * (1. / (2 * dx0));
    const int ilo = nx_ext(g.nx), ihi = g.nx / 2 + nx_ext(g.nx);
    const int jlo = nx_ext(g.ny), jhi = g.ny / 2 + nx_ext(g.ny);
    const double bydx = 1. / (dx0 * (1 << lev));
    const bool inI = i > ilo && i < ihi, inJ = j > jlo && j < jhi;
    const bool onI = i == ilo || i == ihi, onJ = j == jlo || j == jhi;
    if (inI && inJ) return 0.0;
    const int f = lev - 1;
    if (onI && onJ) {   // corners F
        if (j == jlo) {
            const double c = (i == ilo) ? q[g.fx(f, 1, 0)] : q[g.fx(f, g.nx - 1, 0)];
            return (q[g.fx(lev, i, j - 1)] * 8. / 15 + q[g.fx(lev, i, j)] * 6. / 15 + c * 2. / 15) * bydx;
        }
        const double c = (i == ilo) ? q[g.fx(f, 1, g.ny - 1)] : q[g.fx(f, g.nx - 1, g.ny - 1)];
        return (q[g.fx(lev, i, j)] * 8. / 15 + q[g.fx(lev, i, j - 1)] * 6. / 15 + c * 2. / 15) * bydx;
    }
    if (inI && onJ) {   // interfaces E
        const int ii = (i - ilo) * 2;
        if (j == jlo) {
            const int jj = 0;
            return (q[g.fx(lev, i, jlo - 1)] * 2. / 3 + q[g.fx(f, ii, jj)] * 1. / 3
                    + (q[g.fx(f, ii - 1, jj)] + q[g.fx(f, ii + 1, jj)]) * 1. / 6) * bydx;
        }
        const int jj = (jhi - jlo) * 2;
        return (q[g.fx(lev, i, jhi)] * 2. / 3 + q[g.fx(f, ii, jj - 1)] * 1. / 3
                + (q[g.fx(f, ii - 1, jj - 1)] + q[g.fx(f, ii + 1, jj - 1)]) * 1. / 6) * bydx;
    }
    return (q[g.fx(lev, i, j)] + q[g.fx(lev, i, j - 1)]) * 0.5 * bydx;
}

__device__ inline double yvel(const Lay& g, const double* q, double dx0, int lev, int i, int j) {
    if (lev == 0) return (q[g.fy(0, i - 1, j)] + q[g.fy(0, i, j)]) * (1. / (2 * dx0));
    const int ilo = nx_ext(g.nx), ihi = g.nx / 2 + nx_ext(g.nx);
    const int jlo = nx_ext(g.ny), jhi = g.ny / 2 + nx_ext(g.ny);
    const double bydx = 1. / (dx0 * (1 << lev));
    const bool inI = i > ilo && i < ihi, inJ = j > jlo && j < jhi;
    const bool onI = i == ilo || i == ihi, onJ = j == jlo || j == jhi;
    if (inI && inJ) return 0.0;
    const int f = lev - 1;
    if (onI && onJ) {   // corners F
        if (i == ilo) {
            const double c = (j == jlo) ? q[g.fy(f, 0, 1)] : q[g.fy(f, 0, g.ny - 1)];
            return (q[g.fy(lev, i - 1, j)] * 8. / 15 + q[g.fy(lev, i, j)] * 6. / 15 + c * 2. / 15) * bydx;
        }
        const double c = (j == jlo) ? q[g.fy(f, g.nx - 1, 1)] : q[g.fy(f, g.nx - 1, g.ny - 1)];
        return (q[g.fy(lev, i, j)] * 8. / 15 + q[g.fy(lev, i - 1, j)] * 6. / 15 + c * 2. / 15) * bydx;
    }
    if (onI && inJ) {   // interfaces E
        const int jj = (j - jlo) * 2;
        if (i == ilo) {
            const int ii = 0;
            return (q[g.fy(lev, ilo - 1, j)] * 2. / 3 + q[g.fy(f, ii, jj)] * 1. / 3
                    + (q[g.fy(f, ii, jj - 1)] + q[g.fy(f, ii, jj + 1)]) * 1. / 6) * bydx;
        }
        const int ii = (ihi - ilo) * 2;
        return (q[g.fy(lev, ihi, j)] * 2. / 3 + q[g.fy(f, ii - 1, jj)] * 1. / 3
                + (q[g.fy(f, ii - 1, jj - 1)] + q[g.fy(f, ii - 1, jj + 1)]) * 1. / 6) * bydx;
    }
    return (q[g.fy(lev, i, j)] + q[g.fy(lev, i - 1, j)]) * 0.5 * bydx;
}

// XVelocityToFlux / YVelocityToFlux at one flux of level lev
// (VectorOperations.cc:669-915).  C holds the already finished finer level of
// the output, used by the coarse-level region G.
__device__ inline double xflux_from(const Lay& g, const double* S, const double* C, double dx0,
                                    int lev, int i, int j) {
    const double dx = dx0 * (1 << lev);
    const int nx2 = nx_ext(g.nx), ny2 = nx_ext(g.ny);
    if (i == 0 || i == g.nx) {   // A
        if (lev == g.L - 1) return 0.0;
        const int ic = (i == 0) ? nx2 : g.nx / 2 + nx2;
        const int jj = j / 2 + ny2;
        const double a = gin(S)[g.s(lev + 1, ic, jj)], b = gin(S)[g.s(lev + 1, ic, jj + 1)];
        return (j & 1) ? (0.25 * a + 0.75 * b) * dx : (0.75 * a + 0.25 * b) * dx;
    }
    if (j == 0 || j == g.ny - 1) {   // C
        const int jin = (j == 0) ? 1 : g.ny - 1;
        const double s0 = gin(S)[g.s(lev, i, jin)];
        if (lev == g.L - 1) return s0 * 0.5 * dx;
        const int jc = (j == 0) ? ny2 : g.ny / 2 + ny2;
        const int ii = i / 2 + nx2;
        if ((i & 1) == 0) return (s0 + gin(S)[g.s(lev + 1, ii, jc)]) * 0.5 * dx;
        return (0.5 * s0 + 0.25 * gin(S)[g.s(lev + 1, ii, jc)] + 0.25 * gin(S)[g.s(lev + 1, ii + 1, jc)]) * dx;
    }
    if (lev > 0 && i > nx2 && i < g.nx / 2 + nx2 && j >= ny2 && j < g.ny / 2 + ny2) {   // G
        const int ii = 2 * (i - nx2), jj = 2 * (j - ny2);
        return gin(C)[g.fx(lev - 1, ii, jj)] + gin(C)[g.fx(lev - 1, ii, jj + 1)];
    }
    return (gin(S)[g.s(lev, i, j)] + gin(S)[g.s(lev, i, j + 1)]) * 0.5 * dx;
}

__device__ inline double yflux_from(const Lay& g, const double* S, const double* C, double dx0,
                                    int lev, int i, int j) {
    const double dx = dx0 * (1 << lev);
    const int nx2 = nx_ext(g.nx), ny2 = nx_ext(g.ny);
    if (j == 0 || j == g.ny) {   // A
        if (lev == g.L - 1) return 0.0;
        const int jc = (j == 0) ? ny2 : g.ny / 2 + ny2;
        const int ii = i / 2 + nx2;
        const double a = gin(S)[g.s(lev + 1, ii, jc)], b = gin(S)[g.s(lev + 1, ii + 1, jc)];
        return (i & 1) ? (0.25 * a + 0.75 * b) * dx : (0.75 * a + 0.25 * b) * dx;
    }
    if (i == 0 || i == g.nx - 1) {   // C
        const int iin = (i == 0) ? 1 : g.nx - 1;
        const double s0 = gin(S)[g.s(lev, iin, j)];
        if (lev == g.L - 1) return s0 * 0.5 * dx;
        const int ic = (i == 0) ? nx2 : g.nx / 2 + nx2;
        const int jj = j / 2 + ny2;
        if ((j & 1) == 0) return (s0 + gin(S)[g.s(lev + 1, ic, jj)]) * 0.5 * dx;
        return (0.5 * s0 + 0.25 * gin(S)[g.s(lev + 1, ic, jj)] + 0.25 * gin(S)[g.s(lev + 1, ic, jj + 1)]) * dx;
    }
    if (lev > 0 && j > ny2 && j < g.ny / 2 + ny2 && i >= nx2 && i < g.nx / 2 + nx2) {   // G
        const int ii = 2 * (i - nx2), jj = 2 * (j - ny2);
        return gin(C)[g.fy(lev - 1, ii, jj)] + gin(C)[g.fy(lev - 1, ii + 1, jj)];
    }
    return (gin(S)[g.s(lev, i, j)] + gin(S)[g.s(lev, i + 1, j)]) * 0.5 * dx;
}

// ------------------------------------------------------------- kernels
// U = -omega * FluxToX(q), V = omega * FluxToY(q) on every level
// (CrossProduct, VectorOperations.cc:438-457).
__global__ void __launch_bounds__(kThreads) k_vel(const FlowOps* __restrict__ flows) {
    const FlowOps& F = flows[blockIdx.y];
    const Lay g{F.nx, F.ny, F.L};
    const long t = (long)blockIdx.x * kThreads + threadIdx.x;
    if (t >= g.L * g.ns()) return;
    const unsigned ns = (unsigned)g.ns(), ny1 = (unsigned)(g.ny - 1);   // sizes below 2^32
    const int lev = (int)((unsigned)t / ns);
    const unsigned r = (unsigned)t - lev * ns;
    const int i = (int)(r / ny1) + 1, j = (int)(r % ny1) + 1;
    // nodes strictly inside a coarse level's central region: xvel / yvel give
    // 0 there, and k_cross never reads them (its central fluxes take the finer
    // level's cross), so they are not written
    if (lev >= 1 && i > nx_ext(g.nx) && i < g.nx / 2 + nx_ext(g.nx) && j > nx_ext(g.ny) &&
        j < g.ny / 2 + nx_ext(g.ny))
        return;
    const double w = gin(F.omega)[t];
    double u, v;
    if (F.model == 1 || F.model == 3) {
        // LinearizedIBSolver::N (IBSolver.cc:211-216): the velocities of
        // CrossProduct(q0, omega) + CrossProduct(q, omega0), summed per node
        const double w0 = gin(F.w0b)[t];
        double u1 = xvel(g, F.q0b, F.dx, lev, i, j);
        u1 *= w;
        u1 *= -1;
        double u2 = xvel(g, F.q, F.dx, lev, i, j);
        u2 *= w0;
        u2 *= -1;
        double v1 = yvel(g, F.q0b, F.dx, lev, i, j);
        v1 *= w;
        double v2 = yvel(g, F.q, F.dx, lev, i, j);
        v2 *= w0;
        u = u1 + u2;
        v = v1 + v2;
    } else if (F.model == 2) {
        // AdjointIBSolver::N (IBSolver.cc:218-222): -CrossProduct(q, omega0),
        // so that k_rhs's Curl(cross) is the subtracted term
        const double w0 = gin(F.w0b)[t];
        u = xvel(g, F.q, F.dx, lev, i, j);
        u *= w0;
        v = yvel(g, F.q, F.dx, lev, i, j);
        v *= w0;
        v *= -1;
    } else {
        u = xvel(g, F.q, F.dx, lev, i, j);
        u *= w;
        u *= -1;
        v = yvel(g, F.q, F.dx, lev, i, j);
        v *= w;
    }
    gout(F.U)[t] = u;
    gout(F.V)[t] = v;
}

// CrossProduct(Flux, Flux) -> Scalar before its coarsify (VectorOperations.cc
// :465-486): u1 v2 - u2 v1 at every node of every level, with the two
// products rounded separately as the reference forms them.
__global__ void __launch_bounds__(kThreads) k_cross_ff(const CrossFFOps* __restrict__ ops) {
#pragma clang fp contract(off)
    const CrossFFOps& F = ops[blockIdx.y];
    const Lay g{F.nx, F.ny, F.L};
    const long t = (long)blockIdx.x * kThreads + threadIdx.x;
    if (t >= g.L * g.ns()) return;
    const unsigned ns = (unsigned)g.ns(), ny1 = (unsigned)(g.ny - 1);
    const int lev = (int)((unsigned)t / ns);
    const unsigned r = (unsigned)t - lev * ns;
    const int i = (int)(r / ny1) + 1, j = (int)(r % ny1) + 1;
    const double a = xvel(g, F.q1, F.dx, lev, i, j) * yvel(g, F.q2, F.dx, lev, i, j);
    const double b = xvel(g, F.q2, F.dx, lev, i, j) * yvel(g, F.q1, F.dx, lev, i, j);
    gout(F.out)[t] = a - b;
}

// One flux of cross = VelocityToFlux(V, U) on level lev (flat flux index t
// within the level), as xflux_from / yflux_from read it:
//  - next to the level's edge (their A and C cases): edge = true, whose value
//    those functions give;
//  - in the central region of a coarse level (G): src[ka] + src[kb] with src
//    the finer level's cross;
//  - otherwise (src[ka] + src[kb]) * 0.5 * dx with src = V (x-fluxes) or U.
// ka = kb = 0 with src = V or U for an edge flux.
struct CrossAt {
    const double* src;
    unsigned ka, kb;
    bool edge, central, xflux;
    int i, j;
};
__device__ inline CrossAt cross_at(const FlowOps& F, const Lay& g, int lev, unsigned t) {
    // (sizes below 2^32: 32-bit offsets)
    const unsigned nfx = (unsigned)g.nfx(), nf = (unsigned)g.nf(), ny1 = (unsigned)(g.ny - 1);
    const int nx2 = nx_ext(g.nx), ny2 = nx_ext(g.ny);
    CrossAt a;
    a.xflux = t < nfx;
    if (a.xflux) {
        a.i = (int)(t / (unsigned)g.ny);
        a.j = (int)(t % (unsigned)g.ny);
        a.edge = a.i == 0 || a.i == g.nx || a.j == 0 || a.j == g.ny - 1;
        a.central = !a.edge && lev > 0 && a.i > nx2 && a.i < g.nx / 2 + nx2 && a.j >= ny2 && a.j < g.ny / 2 + ny2;
        const unsigned kc = (lev - 1) * nf + (unsigned)(2 * (a.i - nx2)) * g.ny + 2 * (a.j - ny2);   // fx(lev-1, ii, jj)
        const unsigned ks = lev * (unsigned)g.ns() + (a.i - 1) * ny1 + (a.j - 1);                     // s(lev, i, j)
        a.ka = a.edge ? 0 : (a.central ? kc : ks);
        a.kb = a.edge ? 0 : a.ka + 1;
        a.src = a.central ? F.cross : F.V;
    } else {
        const unsigned u = t - nfx;
        a.i = (int)(u / (unsigned)(g.ny + 1));
        a.j = (int)(u % (unsigned)(g.ny + 1));
        a.edge = a.j == 0 || a.j == g.ny || a.i == 0 || a.i == g.nx - 1;
        a.central = !a.edge && lev > 0 && a.j > ny2 && a.j < g.ny / 2 + ny2 && a.i >= nx2 && a.i < g.nx / 2 + nx2;
        const unsigned kc = (lev - 1) * nf + nfx + (unsigned)(2 * (a.i - nx2)) * (g.ny + 1) + 2 * (a.j - ny2);   // fy
        const unsigned ks = lev * (unsigned)g.ns() + (a.i - 1) * ny1 + (a.j - 1);
        a.ka = a.edge ? 0 : (a.central ? kc : ks);
        a.kb = a.edge ? 0 : a.ka + (a.central ? (unsigned)(g.ny + 1) : ny1);
        a.src = a.central ? F.cross : F.U;
    }
    return a;
}

__device__ inline double cross_value(const FlowOps& F, const Lay& g, int lev, const CrossAt& a, double x0,
                                     double x1) {
    const double dx = F.dx * (1 << lev);
    const double sum = x0 + x1, mean = sum * 0.5 * dx;
    const double mc = a.central ? 1.0 : 0.0, md = (a.edge || a.central) ? 0.0 : 1.0;
    const double edge = a.edge ? (a.xflux ? xflux_from(g, F.V, F.cross, F.dx, lev, a.i, a.j)
                                          : yflux_from(g, F.U, F.cross, F.dx, lev, a.i, a.j))
                               : 0.0;
    return (sum * mc + mean * md) + edge;   // exactly one of the three terms is kept
}

// cross = VelocityToFlux(V, U) on one level (levels run fine -> coarse), two
// consecutive fluxes per thread with their four loads issued together.
__global__ void __launch_bounds__(kThreads) k_cross(const FlowOps* __restrict__ flows, int lev) {
    const FlowOps& F = flows[blockIdx.y];
    const Lay g{F.nx, F.ny, F.L};
    if (lev >= g.L) return;
    const long total = g.nf();
    const long t0 = 2 * ((long)blockIdx.x * kThreads + threadIdx.x);
    if (t0 >= total) return;
    const bool two = t0 + 1 < total;
    const CrossAt a = cross_at(F, g, lev, (unsigned)t0), b = cross_at(F, g, lev, (unsigned)(two ? t0 + 1 : t0));
    const double a0 = gin(a.src)[a.ka], a1 = gin(a.src)[a.kb], b0 = gin(b.src)[b.ka], b1 = gin(b.src)[b.kb];
    const double va = cross_value(F, g, lev, a, a0, a1), vb = cross_value(F, g, lev, b, b0, b1);
    IBPM_GLOBAL double* out = gout(F.cross) + lev * total;
    out[t0] = va;
    if (two) out[t0 + 1] = vb;
}

// N = Curl(cross) (IBSolver.cc:205-209) and the right-hand side
// a = omega + dt (c_L Lap(omega) + a_n N + b_n N_prev) (IBSolver.cc:169-183),
// Lap = -Curl(Curl(.)) (VectorOperations.cc:133-140).  N replaces N_prev.
__global__ void __launch_bounds__(kThreads) k_rhs(const FlowOps* __restrict__ flows) {
    const FlowOps& F = flows[blockIdx.y];
    const Lay g{F.nx, F.ny, F.L};
    const long t = (long)blockIdx.x * kThreads + threadIdx.x;
    if (t >= g.L * g.ns()) return;
    // (sizes below 2^32: 32-bit divisions)
    const unsigned ns = (unsigned)g.ns(), ny1 = (unsigned)(g.ny - 1);
    const int lev = (int)((unsigned)t / ns);
    const unsigned r = (unsigned)t - lev * ns;
    const int i = (int)(r / ny1) + 1, j = (int)(r % ny1) + 1;
    if (F.skip_central && lev >= 1 && i > nx_ext(g.nx) && i < g.nx / 2 + nx_ext(g.nx) && j > nx_ext(g.ny) &&
        j < g.ny / 2 + nx_ext(g.ny))
        return;
    const double d = F.dx * (1 << lev);
    const double bydx2 = 1. / (d * d);
    const IBPM_GLOBAL double* C = gin(F.cross);
    const double* W = F.omega;
    const IBPM_GLOBAL double* Wg = gin(W);
    // Away from the level's edge the curls of Lap(omega) are plain differences
    // of the four neighbours (the general expression, term for term and in the
    // same order).  Their offsets fall back to t itself next to the edge, so
    // all ten loads below are issued before any value is used; N_prev is read
    // from its buffer (always allocated) and used only when valid.
    const bool interior = i >= 2 && i <= g.nx - 2 && j >= 2 && j <= g.ny - 2;
    const unsigned tu = (unsigned)t;
    const unsigned te = interior ? tu + ny1 : tu, tw = interior ? tu - ny1 : tu;
    const unsigned tn = interior ? tu + 1 : tu, ts = interior ? tu - 1 : tu;
    const double cy1 = C[g.fy(lev, i, j)], cy0 = C[g.fy(lev, i - 1, j)];
    const double cx0 = C[g.fx(lev, i, j - 1)], cx1 = C[g.fx(lev, i, j)];
    const double c = Wg[tu], e = Wg[te], w = Wg[tw], up = Wg[tn], dn = Wg[ts];
    const double np = gin(F.nprev)[tu];
    double n = (cy1 - cy0 + cx0 - cx1) * bydx2;
    if (F.model == 2) {   // AdjointIBSolver::N = Laplacian(s) - Curl(CrossProduct(q, omega0))
        const IBPM_GLOBAL double* Sg = gin(F.sff);
        const double sc = Sg[tu], se = Sg[te], sw = Sg[tw], su = Sg[tn], sd = Sg[ts];
        const double sedge = interior ? 0.0
                                      : (curl_y(g, F.sff, lev, i, j) - curl_y(g, F.sff, lev, i - 1, j)
                                         + curl_x(g, F.sff, lev, i, j - 1) - curl_x(g, F.sff, lev, i, j)) * bydx2;
        double lap_s = ((sc - se) - (sw - sc) + (sc - sd) - (su - sc)) * bydx2 * (interior ? 1.0 : 0.0) + sedge;
        lap_s *= -1;
        n = lap_s + n;
    }
    double sfd_w = 0.0;   // SFD: the filtered vorticity (zero before the first substep)
    if (F.model == 4) {   // SFDSolver::N (IBSolver.cc:238-244): N - chi (omega - omega_hat)
        sfd_w = F.sfd_valid ? gin(F.what)[tu] : 0.0;
        n -= F.sfd_chi * (c - sfd_w);
    }
    // the interior difference times 1 plus 0, or times 0 plus the edge value
    const double edge = interior ? 0.0
                                 : (curl_y(g, W, lev, i, j) - curl_y(g, W, lev, i - 1, j)
                                    + curl_x(g, W, lev, i, j - 1) - curl_x(g, W, lev, i, j)) * bydx2;
    double lap = ((c - e) - (w - c) + (c - dn) - (up - c)) * bydx2 * (interior ? 1.0 : 0.0) + edge;
    lap *= -1;
    double a = lap * F.lap_coef;
    a += F.an * n;
    if (F.use_nprev) a += F.bn * (F.nprev_valid ? np : n);
    a *= F.dt;
    a += F.w_coef * c;
    gout(F.a)[t] = a;
    gout(F.nprev)[t] = n;
    if (F.model == 4) {   // the filtered state's substep (IBSolver.cc:247-279), omega_hat = omega at the first
        const double wh = F.sfd_valid ? sfd_w : c;
        const double rs = (c - wh) * F.sfd_inv_delta;
        double av = F.an * rs;
        if (F.bn != 0) av += F.bn * (F.nprev_valid ? gin(F.sfd_prev)[tu] : rs);
        av *= F.dt;
        gout(F.what)[tu] = wh + av;
        gout(F.sfd_prev)[tu] = rs;
    }
}

// q = Curl(psi) + q0 on every level (NavierStokesModel.cc:121-129): q0 is
// the uniform base flow (Flux.cc:79-99) or the unsteady one, free stream plus
// rotation about the body centre at the edge coordinates of each level
// (Flux.cc:101-129, Grid.cc:104-139).  Thread (lev, i, j) of each level's
// (nx + 1) x (ny + 1) cells writes X flux (i, j) (j < ny) and Y flux (i, j)
// (i < nx), so each psi value is fetched once for both: a cell away from the
// level's edge reads psi(i, j), psi(i, j + 1) and psi(i + 1, j) and takes
// curl_x's and curl_y's interior differences; the others take those
// functions (getBC at the edge).
__global__ void __launch_bounds__(kThreads) k_curl_q(const FlowOps* __restrict__ flows) {
    const FlowOps& F = flows[blockIdx.y];
    const Lay g{F.nx, F.ny, F.L};
    // (sizes below 2^32: 32-bit offsets and divisions)
    const unsigned nyc = (unsigned)(g.ny + 1), cells = (unsigned)(g.nx + 1) * nyc;
    const unsigned t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= cells * (unsigned)g.L) return;
    const int lev = (int)(t / cells);
    const unsigned r = t - lev * cells;
    const int i = (int)(r / nyc), j = (int)(r % nyc);
    double cx, cy;
    if (i >= 1 && i <= g.nx - 2 && j >= 1 && j <= g.ny - 2) {
        const unsigned ny1 = (unsigned)(g.ny - 1);
        const unsigned k = lev * (unsigned)g.ns() + (unsigned)(i - 1) * ny1 + (unsigned)(j - 1);
        const IBPM_GLOBAL double* P = gin(F.psi);
        const double p = P[k], pn = P[k + 1], pe = P[k + ny1];
        cx = pn - p;   // psi(i, j + 1) - psi(i, j)
        cy = p - pe;   // psi(i, j) - psi(i + 1, j)
    } else {
        cx = j < g.ny ? curl_x(g, F.psi, lev, i, j) : 0.0;
        cy = i < g.nx ? curl_y(g, F.psi, lev, i, j) : 0.0;
    }
    const double d = F.dx * (1 << lev);
    // level offsets: offset + 0.5 (shift - 1) (2^lev - 1) n dx with shift 0
    const double lofs = -0.5 * ((1 << lev) - 1);
    double qx0 = F.ubase * d, qy0 = F.vbase * d;
    if (F.ubf) {
        const double y = (F.yoff + lofs * (F.ny * F.dx)) + (j + 0.5) * d;
        qx0 = (F.bf[0] - F.bf[2] * (y - F.yc)) * d;
        const double x = (F.xoff + lofs * (F.nx * F.dx)) + (i + 0.5) * d;
        qy0 = (F.bf[1] + F.bf[2] * (x - F.xc)) * d;
    }
    IBPM_GLOBAL double* q = gout(F.q) + lev * (unsigned)g.nf();
    if (j < g.ny) q[(unsigned)i * (unsigned)g.ny + (unsigned)j] = cx + qx0;
    if (i < g.nx) q[(unsigned)g.nfx() + (unsigned)i * nyc + (unsigned)j] = cy + qy0;
}

// ---------------------------------------------------------- regulariser
// Level-0 flux of Curl(psi) at a flat flux index.
__device__ inline double curl_flat(const Lay& g, const double* psi, int fidx) {
    if (fidx < g.nfx()) return curl_x(g, psi, 0, fidx / g.ny, fidx % g.ny);
    const int u = fidx - (int)g.nfx();
    return curl_y(g, psi, 0, u / (g.ny + 1), u % (g.ny + 1));
}

// The two level-0 values whose difference is Curl(psi) at a flat flux index,
// Curl = psi[p0] - psi[p1], when the flux is interior (curl_x / curl_y take
// their last branch); false, with p0 = p1 = 0, otherwise.
__device__ inline bool interior_curl(const Lay& g, int fidx, int& p0, int& p1) {
    p0 = p1 = 0;
    const int nfx = (g.nx + 1) * g.ny, ny1 = g.ny - 1;
    if (fidx < nfx) {
        const int i = fidx / g.ny, j = fidx % g.ny;
        if (i < 1 || i > g.nx - 1 || j < 1 || j > g.ny - 2) return false;
        p1 = (i - 1) * ny1 + (j - 1);
        p0 = p1 + 1;
        return true;
    }
    const int u = fidx - nfx, i = u / (g.ny + 1), j = u % (g.ny + 1);
    if (j < 1 || j > g.ny - 1 || i < 1 || i > g.nx - 2) return false;
    p0 = (i - 1) * ny1 + (j - 1);
    p1 = p0 + ny1;
    return true;
}

// f[k] = E(Curl psi)[k] - b[k]: toBoundary (Regularizer.cc:111-128) of the
// flux of psi, each component summed in ascending flux order as the
// reference's association list does; b may be null.  Four entries of the
// list are read at a time, their psi values loaded together from valid
// addresses; a flux next to the level's edge takes curl_flat instead.
__global__ void __launch_bounds__(kThreads) k_interp(const RegOps* __restrict__ regs) {
    const RegOps& R = regs[blockIdx.y];
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= 2 * R.nb || skipped(R.skip)) return;
    const Lay g{R.nx, R.ny, R.L};
    const int a0 = R.e_ptr[k], a1 = R.e_ptr[k + 1];
    double s = 0.0;
    for (int a = a0; a < a1; a += 4) {
        int f[4];
        double w[4], c[4], m[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = min(a + u, a1 - 1);
            f[u] = R.e_f[e];
            w[u] = R.e_w[e];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int p0, p1;
            m[u] = interior_curl(g, f[u], p0, p1) ? 1.0 : 0.0;
            c[u] = R.psi[p0] - R.psi[p1];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double edge = (m[u] != 0.0) ? 0.0 : curl_flat(g, R.psi, f[u]);
            const double v = c[u] * m[u] + edge;   // exactly c or exactly edge
            if (a + u < a1) s += w[u] * v;
        }
    }
    s *= R.inv_dx;
    if (R.b) s -= R.b[k];
    R.out[k] = s;
}

// c = beta * Curl(E^T f) at the level-0 nodes touched by the regulariser
// (NavierStokesModel.cc:101-105, Regularizer.cc:92-109, ProjectionSolver.cc:85-88).
__device__ inline double smear_at(const RegOps& R, int fl) {
    if (fl < 0) return 0.0;
    double s = 0.0;
    for (int a = R.t_ptr[fl]; a < R.t_ptr[fl + 1]; ++a) s += R.t_w[a] * R.fin[R.t_k[a]];
    return s * R.dx;
}

__global__ void __launch_bounds__(kThreads) k_smear_curl(const RegOps* __restrict__ regs) {
    const RegOps& R = regs[blockIdx.y];
    const int m = blockIdx.x * kThreads + threadIdx.x;
    if (m >= R.nnodes || skipped(R.skip)) return;
    const int4 nb = R.node_nbr[m];   // flux-list slots of Y(i,j), Y(i-1,j), X(i,j-1), X(i,j)
    const double w = (smear_at(R, nb.x) - smear_at(R, nb.y) + smear_at(R, nb.z) - smear_at(R, nb.w)) * R.bydx2;
    R.cout[R.node_idx[m]] = w * R.beta;
}

__global__ void __launch_bounds__(kThreads) k_clear_nodes(const RegOps* __restrict__ regs) {
    const RegOps& R = regs[blockIdx.y];
    const int m = blockIdx.x * kThreads + threadIdx.x;
    if (m >= R.nold) return;
    R.cout[R.old_idx[m]] = 0.0;
}

// Regularizer::toFlux (Regularizer.cc:92-109): q[flux] = dx * sum of w f[k]
// over the components touching that flux, in ascending component order (the
// reference's accumulation order for a fixed flux); one thread per flux slot,
// every slot a distinct flux.  Fluxes no slot touches are left as they are.
__global__ void __launch_bounds__(kThreads) k_reg_to_flux(const RegFluxOps* __restrict__ R, const double* f,
                                                          double* q) {
    const int s = blockIdx.x * kThreads + threadIdx.x;
    if (s >= R->nslots) return;
    double v = 0.0;
    for (int a = R->t_ptr[s]; a < R->t_ptr[s + 1]; ++a) v += R->t_w[a] * f[R->t_k[a]];
    q[R->flux_of_slot[s]] = v * R->dx;
}

// Regularizer::toBoundary (Regularizer.cc:111-128): f[k] = (sum of w q[flux]
// in ascending flux order) / dx.
__global__ void __launch_bounds__(kThreads) k_reg_to_boundary(const RegFluxOps* __restrict__ R, const double* q,
                                                              double* f) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= 2 * R->nb) return;
    double v = 0.0;
    for (int a = R->e_ptr[k]; a < R->e_ptr[k + 1]; ++a) v += R->e_w[a] * q[R->e_f[a]];
    f[k] = v * (1.0 / R->dx);
}

// ------------------------------------------------------------ force solve
// x = Minv * rhs with Minv = (L L^T)^-1 assembled in FP64 from the Cholesky
// factor (CholeskySolver.cc:200-224 computes the same product by forward and
// back substitution).  One wave per row, rows spread over the whole GPU.
__global__ void __launch_bounds__(kThreads) k_minv(const MinvOps* __restrict__ mops) {
    const MinvOps& P = mops[blockIdx.y];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int row = blockIdx.x * (kThreads / 64) + wave;
    if (row >= P.n || skipped(P.skip)) return;
    const double* a = P.minv + (long)row * P.n;
    double s = 0.0;
#pragma unroll 4   // loads of four iterations in flight; each lane's sum keeps its order
    for (int k = lane; k < P.n; k += 64) s += a[k] * P.rhs[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) P.x[row] = P.scale * s;
}

// ------------------------------------------------------------------- CG
// Conjugate gradient on M f = rhs (ConjugateGradientSolver.cc:35-65), one
// workgroup per flow; M d is applied by the solve kernels between calls.
__device__ inline double block_sum(double v, double* red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kThreads / 64; ++w) s += red[w];
        red[kThreads / 64] = s;
    }
    __syncthreads();
    return red[kThreads / 64];
}

// mode 0: start   r = b - q (q = M x0), d = r, delta = r.r
// mode 1: iterate alpha = delta/(d.q), x += alpha d, r -= alpha q,
//                 delta' = r.r, beta = delta'/delta, d = r + beta d
__global__ void __launch_bounds__(kThreads) k_cg(const CgOps* __restrict__ cgs, int mode) {
    const CgOps& C = cgs[blockIdx.x];
    __shared__ double red[kThreads / 64 + 1];
    double* st = C.state;   // [delta, tol2, iterations, converged]
    const int n = C.n;
    if (mode == 0) {
        double s = 0.0;
        for (int k = threadIdx.x; k < n; k += kThreads) {
            double r = C.b[k];
            r -= C.q[k];
            C.r[k] = r;
            C.d[k] = r;
            s += r * r;
        }
        const double delta = block_sum(s, red);
        if (threadIdx.x == 0) {
            st[0] = delta;
            st[2] = 0;
            st[3] = (delta > st[1]) ? 0.0 : 1.0;
        }
        return;
    }
    if (st[3] != 0.0) return;   // converged earlier: this iteration is a no-op
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += kThreads) s += C.d[k] * C.q[k];
    const double dq = block_sum(s, red);
    const double delta = st[0];
    const double alpha = delta / dq;
    s = 0.0;
    for (int k = threadIdx.x; k < n; k += kThreads) {
        C.x[k] += alpha * C.d[k];
        double r = C.r[k] - alpha * C.q[k];
        C.r[k] = r;
        s += r * r;
    }
    const double dnew = block_sum(s, red);
    const double beta = dnew / delta;
    for (int k = threadIdx.x; k < n; k += kThreads) C.d[k] = C.r[k] + beta * C.d[k];
    if (threadIdx.x == 0) {
        st[0] = dnew;
        st[2] += 1;
        st[3] = (dnew > st[1]) ? 0.0 : 1.0;
    }
}

// ------------------------------------------------------------------ PCG
// Sums K per-thread values over the workgroup; every thread gets the totals.
// The order of the additions is fixed (lane tree, then waves in order).
template <int K>
__device__ inline void block_sums(double (&v)[K], double* red) {   // red: (kThreads/64 + 1) K
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[a] += __shfl_down(v[a], off, 64);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    constexpr int W = kThreads / 64;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int a = 0; a < K; ++a) red[wave * K + a] = v[a];
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0.0;
        for (int w = 0; w < W; ++w) s += red[w * K + threadIdx.x];
        red[W * K + threadIdx.x] = s;
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < K; ++a) v[a] = red[W * K + a];
}

constexpr int kRed = (kThreads / 64 + 1) * kLmpMax;

// Row of the pair arrays read for pair a when nk pairs are kept (nk >= 1):
// pairs past the last read its row again, so every pair's load is issued
// unconditionally; what they add is either discarded (sums for a >= nk are
// never used) or multiplied by a coefficient that is exactly 0.
__device__ inline int pair_row(int a, int nk) { return a < nk ? a : nk - 1; }

// First half of the preconditioner for the current residual:
//   c = G S^T r,  t = r - Y c   (t = r without pairs).
__device__ inline void pcg_project(const PcgOps& C, int nk, const double* G, double* red) {
    const int n = C.n;
    if (nk == 0) {
        for (int k = threadIdx.x; k < n; k += kThreads) C.t[k] = C.r[k];
        return;
    }
    double v[kLmpMax];
#pragma unroll
    for (int a = 0; a < kLmpMax; ++a) v[a] = 0.0;
    for (int k = threadIdx.x; k < n; k += kThreads) {
        const double r = C.r[k];
#pragma unroll
        for (int a = 0; a < kLmpMax; ++a) v[a] += C.rs[(long)pair_row(a, nk) * n + k] * r;
    }
    block_sums<kLmpMax>(v, red);
    double c[kLmpMax];
#pragma unroll
    for (int a = 0; a < kLmpMax; ++a) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < kLmpMax; ++b)
            if (a < nk && b < nk) s += G[a * kLmpMax + b] * v[b];
        c[a] = s;
    }
    if (threadIdx.x < kLmpMax) {
#pragma unroll
        for (int a = 0; a < kLmpMax; ++a)
            if (a == (int)threadIdx.x) C.c[a] = c[a];
    }
    for (int k = threadIdx.x; k < n; k += kThreads) {
        double s = C.r[k];
#pragma unroll
        for (int a = 0; a < kLmpMax; ++a) s -= C.ry[(long)pair_row(a, nk) * n + k] * c[a];
        C.t[k] = s;
    }
}

// Start of a solve: r = b - q (q = M x0), the residual test, and the inverse
// Gram matrix G of the previous solve's pairs (symmetrised S^T Y, Cholesky in
// FP64; the pairs are cut to the leading block that is safely positive
// definite).  Ends with the projection for the first preconditioner apply.
__global__ void __launch_bounds__(kThreads) k_pcg_begin(const PcgOps* __restrict__ ops) {
    const PcgOps& C = ops[blockIdx.x];
    __shared__ double red[kRed];
    __shared__ double W[kLmpMax * kLmpMax], G[kLmpMax * kLmpMax];
    __shared__ int nk_s;
    double* st = C.state;
    const int n = C.n;
    double v1[1] = {0.0};
    for (int k = threadIdx.x; k < n; k += kThreads) {
        const double r = C.b[k] - C.q[k];
        C.r[k] = r;
        v1[0] += r * r;
    }
    block_sums<1>(v1, red);
    const double delta = v1[0];
    int nk = min(*C.rcnt, C.kmax);
    for (int a = 0; a < nk; ++a) {
        double v[kLmpMax];
#pragma unroll
        for (int b = 0; b < kLmpMax; ++b) v[b] = 0.0;
        for (int k = threadIdx.x; k < n; k += kThreads) {
            const double sa = C.rs[(long)a * n + k], ya = C.ry[(long)a * n + k];
#pragma unroll
            for (int b = 0; b < kLmpMax; ++b) {
                const long o = (long)pair_row(b, nk) * n + k;
                v[b] += sa * C.ry[o] + C.rs[o] * ya;
            }
        }
        block_sums<kLmpMax>(v, red);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int b = 0; b < kLmpMax; ++b) W[a * kLmpMax + b] = 0.5 * v[b];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // W = L L^T (L in the lower triangle of W), stopping at the first
        // pivot that is not safely positive
        int m = 0;
        for (; m < nk; ++m) {
            double s = W[m * kLmpMax + m];
            for (int k = 0; k < m; ++k) s -= W[m * kLmpMax + k] * W[m * kLmpMax + k];
            if (!(s > 1e-10 * W[m * kLmpMax + m])) break;
            const double dm = sqrt(s);
            W[m * kLmpMax + m] = dm;
            for (int i = m + 1; i < nk; ++i) {
                double t = W[i * kLmpMax + m];
                for (int k = 0; k < m; ++k) t -= W[i * kLmpMax + k] * W[m * kLmpMax + k];
                W[i * kLmpMax + m] = t / dm;
            }
        }
        nk = m;
        // G = (L L^T)^-1, column by column
        for (int col = 0; col < nk; ++col) {
            double x[kLmpMax];
            for (int i = 0; i < nk; ++i) {
                double t = (i == col) ? 1.0 : 0.0;
                for (int k = 0; k < i; ++k) t -= W[i * kLmpMax + k] * x[k];
                x[i] = t / W[i * kLmpMax + i];
            }
            for (int i = nk - 1; i >= 0; --i) {
                double t = x[i];
                for (int k = i + 1; k < nk; ++k) t -= W[k * kLmpMax + i] * x[k];
                x[i] = t / W[i * kLmpMax + i];
            }
            for (int i = 0; i < nk; ++i) G[i * kLmpMax + col] = x[i];
        }
        for (int q = 0; q < kLmpMax * kLmpMax; ++q) C.g[q] = (q / kLmpMax < nk && q % kLmpMax < nk) ? G[q] : 0.0;
        nk_s = nk;
        st[2] = 0.0;
        st[3] = (delta > st[1]) ? 0.0 : 1.0;
        st[4] = 0.0;
        st[5] = nk;
        st[6] = delta;
        *C.wcnt = 0;
    }
    __syncthreads();
    nk = nk_s;
    pcg_project(C, nk, G, red);
}

// One iteration after q = M d:  alpha = rho/(d.q), x += alpha d,
// r -= alpha q, residual test on r.r; keeps (d, q) as a pair for the next
// solve while fewer than kmax are kept; then the projection of the new r.
__global__ void __launch_bounds__(kThreads) k_pcg_step(const PcgOps* __restrict__ ops) {
    const PcgOps& C = ops[blockIdx.x];
    __shared__ double red[kRed];
    double* st = C.state;
    if (st[3] != 0.0 || st[4] != 0.0) return;
    const int n = C.n, nk = (int)st[5], j = (int)st[2];
    const double rho = st[0];
    double v[1] = {0.0};
    for (int k = threadIdx.x; k < n; k += kThreads) v[0] += C.d[k] * C.q[k];
    block_sums<1>(v, red);
    const double dq = v[0];
    if (!(dq > 0.0)) {   // M not positive on d: plain CG takes over
        if (threadIdx.x == 0) st[4] = 1.0;
        return;
    }
    const double alpha = rho / dq;
    const bool keep = j < C.kmax;
    v[0] = 0.0;
    for (int k = threadIdx.x; k < n; k += kThreads) {
        const double dk = C.d[k], qk = C.q[k];
        if (keep) {
            C.ws[(long)j * n + k] = dk;
            C.wy[(long)j * n + k] = qk;
        }
        C.x[k] += alpha * dk;
        const double r = C.r[k] - alpha * qk;
        C.r[k] = r;
        v[0] += r * r;
    }
    block_sums<1>(v, red);
    if (threadIdx.x == 0) {
        st[2] = j + 1;
        st[3] = (v[0] > st[1]) ? 0.0 : 1.0;
        st[6] = v[0];
        if (keep) *C.wcnt = j + 1;
    }
    pcg_project(C, nk, C.g, red);
}

// After u = P t:  z = u + S (c - G Y^T u)  (stored over u), rho' = r.z,
// d = z (begin) or d = z + (rho'/rho) d.
__global__ void __launch_bounds__(kThreads) k_pcg_dir(const PcgOps* __restrict__ ops, int begin) {
    const PcgOps& C = ops[blockIdx.x];
    __shared__ double red[kRed];
    double* st = C.state;
    if (st[3] != 0.0 || st[4] != 0.0) return;
    const int n = C.n, nk = (int)st[5];
    const double rho = st[0];
    double w[kLmpMax];
#pragma unroll
    for (int a = 0; a < kLmpMax; ++a) w[a] = 0.0;
    if (nk > 0) {
        double v[kLmpMax];
#pragma unroll
        for (int a = 0; a < kLmpMax; ++a) v[a] = 0.0;
        for (int k = threadIdx.x; k < n; k += kThreads) {
            const double u = C.u[k];
#pragma unroll
            for (int a = 0; a < kLmpMax; ++a) v[a] += C.ry[(long)pair_row(a, nk) * n + k] * u;
        }
        block_sums<kLmpMax>(v, red);
#pragma unroll
        for (int a = 0; a < kLmpMax; ++a) {
            double s = 0.0;
#pragma unroll
            for (int b = 0; b < kLmpMax; ++b)
                if (a < nk && b < nk) s += C.g[a * kLmpMax + b] * v[b];
            w[a] = (a < nk) ? C.c[a] - s : 0.0;
        }
    }
    double v1[1] = {0.0};
    for (int k = threadIdx.x; k < n; k += kThreads) {
        double z = C.u[k];
        if (nk > 0) {
#pragma unroll
            for (int a = 0; a < kLmpMax; ++a) z += C.rs[(long)pair_row(a, nk) * n + k] * w[a];
        }
        C.u[k] = z;
        v1[0] += C.r[k] * z;
    }
    block_sums<1>(v1, red);
    const double rnew = v1[0];
    if (!(rnew > 0.0)) {   // preconditioner not positive on r: plain CG takes over
        if (threadIdx.x == 0) st[4] = 1.0;
        return;
    }
    const double beta = begin ? 0.0 : rnew / rho;
    for (int k = threadIdx.x; k < n; k += kThreads) C.d[k] = begin ? C.u[k] : C.u[k] + beta * C.d[k];
    if (threadIdx.x == 0) st[0] = rnew;
}

// Net force sum over the boundary points (State.cc:53-63; the dx^2 factor and
// the rotation to Cd/Cl are applied on the host).  One wave per flow: lane l
// sums points l, l+64, ... in order, then a fixed shuffle tree combines the
// lanes, so the result is deterministic.
__global__ void __launch_bounds__(64) k_forces(const ForceOps* __restrict__ fo) {
    const ForceOps& F = fo[blockIdx.x];
    double xf = 0, yf = 0;
    for (int k = threadIdx.x; k < F.nb; k += 64) {
        xf += F.f[k];
        yf += F.f[F.nb + k];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        xf += __shfl_down(xf, off, 64);
        yf += __shfl_down(yf, off, 64);
    }
    if (threadIdx.x == 0) {
        const int st = *F.count;
        F.hist[2 * st] = xf;
        F.hist[2 * st + 1] = yf;
        *F.count = st + 1;
    }
}

// ------------------------------------------------------------ host side
static int blocks_for(long n) { return (int)((n + kThreads - 1) / kThreads); }

void launch_vel(const FlowOps* d, int nflows, long max_pts, hipStream_t s) {
    hipLaunchKernelGGL(k_vel, dim3(blocks_for(max_pts), nflows), dim3(kThreads), 0, s, d);
}
void launch_cross(const FlowOps* d, int nflows, long max_nf, int lev, hipStream_t s) {
    hipLaunchKernelGGL(k_cross, dim3(blocks_for((max_nf + 1) / 2), nflows), dim3(kThreads), 0, s, d, lev);
}
void launch_cross_ff(const CrossFFOps* d, int nflows, long max_pts, hipStream_t s) {
    hipLaunchKernelGGL(k_cross_ff, dim3(blocks_for(max_pts), nflows), dim3(kThreads), 0, s, d);
}
void launch_rhs(const FlowOps* d, int nflows, long max_pts, hipStream_t s) {
    hipLaunchKernelGGL(k_rhs, dim3(blocks_for(max_pts), nflows), dim3(kThreads), 0, s, d);
}
void launch_curl_q(const FlowOps* d, int nflows, long max_cells_all, hipStream_t s) {
    hipLaunchKernelGGL(k_curl_q, dim3(blocks_for(max_cells_all), nflows), dim3(kThreads), 0, s, d);
}
// One thread per flow: e_{j-1} -> e_j.
__global__ void k_unit_column(const UnitOps* __restrict__ ops, int nflows, int j) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nflows) return;
    const UnitOps& o = ops[k];
    if (j > 0 && j - 1 < o.n) o.u[j - 1] = 0.0;
    if (j < o.n) o.u[j] = 1.0;
}

void launch_unit_column(const UnitOps* d, int nflows, int j, hipStream_t s) {
    hipLaunchKernelGGL(k_unit_column, dim3((nflows + 63) / 64), dim3(64), 0, s, d, nflows, j);
}

// Piece blockIdx.y, strided over gridDim.x workgroups.
__global__ void __launch_bounds__(kThreads) k_copy_pieces(const CopyPiece* __restrict__ pieces) {
    const CopyPiece P = pieces[blockIdx.y];
    for (long w = (long)blockIdx.x * kThreads + threadIdx.x; w < P.n; w += (long)gridDim.x * kThreads)
        P.dst[w] = P.src ? P.src[w] : 0u;
}

void launch_copy_pieces(const CopyPiece* d, int npieces, hipStream_t s) {
    if (npieces > 0) hipLaunchKernelGGL(k_copy_pieces, dim3(16, npieces), dim3(kThreads), 0, s, d);
}

// ---------------------------------------------------------------------------
// Kernel-table assembly of the preconditioner's matrix (PcAsmOps).  A
// workgroup forms a 16 x 16 tile on or above the diagonal; the separable rows
// of its 16 rows and 16 columns are staged in LDS.  For rows r (component cr,
// nodes i0r + a, j0r + b, weights ur[a] vr[b]) and c, the sum over node
// pairs is a sum over offsets (di, dj) = (a - a', b - b') of U(di) V(dj) G at
// offset (i0r - i0c + di, j0r - j0c + dj), with U the correlation of ur and
// uc and V that of vr and vc: 25 table reads instead of 81.
constexpr int kAsmTile = 16;

__global__ void __launch_bounds__(kAsmTile * kAsmTile) k_pc_assemble(const PcAsmOps* __restrict__ ops, int ntiles) {
    const PcAsmOps o = ops[blockIdx.y];
    const int nt = (o.n + kAsmTile - 1) / kAsmTile;
    const int br = blockIdx.x / ntiles, bc = blockIdx.x % ntiles;
    if (br >= nt || bc >= nt || br > bc) return;
    __shared__ double sp[2][kAsmTile][kESep];
    __shared__ int near[kAsmTile];   // each column's source
    const int t = threadIdx.x;
    for (int q = t; q < 2 * kAsmTile * kESep; q += kAsmTile * kAsmTile) {
        const int side = q / (kAsmTile * kESep), w = q % (kAsmTile * kESep);
        const int row = (side ? bc : br) * kAsmTile + w / kESep;
        sp[side][w / kESep][w % kESep] = row < o.n ? o.sep[(long)row * kESep + w % kESep] : 0.0;
    }
    __syncthreads();
    if (t < kAsmTile) {   // the source nearest to the centre of the column's nodes (first of equals)
        const int ci = (int)sp[1][t][0] + 1, cj = (int)sp[1][t][1] + 1;
        int best = 0;
        long bd = LONG_MAX;
        for (int a = 0; a < o.nsrc; ++a) {
            const long di = ci - o.src[2 * a], dj = cj - o.src[2 * a + 1], d = di * di + dj * dj;
            if (d < bd) {
                bd = d;
                best = a;
            }
        }
        near[t] = best;
    }
    __syncthreads();
    const int ty = t / kAsmTile, tx = t % kAsmTile;
    const int r = br * kAsmTile + ty, c = bc * kAsmTile + tx;
    if (r >= o.n || c >= o.n || (br == bc && ty > tx)) return;
    const double* R = sp[0][ty];
    const double* C = sp[1][tx];
    const int nir = (int)R[2], njr = (int)R[3], nic = (int)C[2], njc = (int)C[3];
    // U[k]: offset di = k - (nic - 1); V[m]: dj = m - (njc - 1)
    double U[5], V[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int di = k - (nic - 1), dj = k - (njc - 1);
        double su = 0.0, sv = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (a < nir && a - di >= 0 && a - di < nic) su += R[4 + a] * C[4 + a - di];
            if (a < njr && a - dj >= 0 && a - dj < njc) sv += R[7 + a] * C[7 + a - dj];
        }
        U[k] = su;
        V[k] = sv;
    }
    const int cr = r >= o.n / 2, cc = c >= o.n / 2;
    const long nfx = (long)(o.nx + 1) * o.ny, nf0 = nfx + (long)o.nx * (o.ny + 1);
    // the part of the column's source's response on a's component
    const int a = near[tx];
    const double* T = o.tab + (2L * a + cc) * nf0 + (cr ? nfx : 0);
    const int rows = cr ? o.nx : o.nx + 1, cols = cr ? o.ny + 1 : o.ny;
    const int i0 = (int)R[0] - (int)C[0] + o.src[2 * a] - (nic - 1);
    const int j0 = (int)R[1] - (int)C[1] + o.src[2 * a + 1] - (njc - 1);
    double sum = 0.0;
    for (int k = 0; k < nir + nic - 1; ++k) {
        const int i = i0 + k;
        if (i < 0 || i >= rows) continue;
        double part = 0.0;
        for (int m = 0; m < njr + njc - 1; ++m) {
            const int j = j0 + m;
            if (j >= 0 && j < cols) part += V[m] * T[(long)i * cols + j];
        }
        sum += U[k] * part;
    }
    const double v = o.scale * sum;
    o.out[(long)r * o.n + c] = v;
    o.out[(long)c * o.n + r] = v;
}

void launch_pc_assemble(const PcAsmOps* d, int nflows, int max_n, hipStream_t s) {
    const int nt = (max_n + kAsmTile - 1) / kAsmTile;
    if (nflows > 0 && nt > 0)
        hipLaunchKernelGGL(k_pc_assemble, dim3(nt * nt, nflows), dim3(kAsmTile * kAsmTile), 0, s, d, nt);
}

// ---------------------------------------------------------------------------
// Batched FP64 GEMM (GemmOps), one matrix per blockIdx.y, on the FP64 matrix
// cores: a 64 x 64 output tile per workgroup of 4 waves, each wave a 32 x 32
// quarter as 2 x 2 blocks of v_mfma_f64_16x16x4f64, k in steps of 16 staged in
// LDS (A's tile transposed); each step's global loads for the next step are
// issued before its products.  Fragments (lane l, checked on the GPU): A[l mod
// 16][l / 16], B[l / 16][l mod 16], D[l / 16 + 4 r][l mod 16] in register r.
// (Measured on 16 products of 600^2: 194 us, against 301 us for the same
// tiling with 4 x 4 VALU FMAs per thread.)
constexpr int kGemmT = 64, kGemmK = 16;
typedef double double4_v __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_dgemm(const GemmOps* __restrict__ ops, int ntiles, int sym) {
    const GemmOps o = ops[blockIdx.y];
    const int n = o.n, nt = (n + kGemmT - 1) / kGemmT;
    const int br = blockIdx.x / ntiles, bc = blockIdx.x % ntiles;
    if (br >= nt || bc >= nt || (sym && br > bc)) return;
    __shared__ double As[kGemmK][kGemmT + 1];
    __shared__ double Bs[kGemmK][kGemmT];
    const int t = threadIdx.x, w = t / 64, l = t & 63;
    const int r0 = br * kGemmT, c0 = bc * kGemmT;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;   // this wave's quarter
    const int lr = l & 15, lk = l >> 4;
    constexpr int kPer = kGemmT * kGemmK / 256;      // elements of each tile per thread
    double ra[kPer], rb[kPer];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = t + 256 * q;
            const int ar = r0 + e / kGemmK, ak = k0 + e % kGemmK;
            const int bk = k0 + e / kGemmT, bcol = c0 + e % kGemmT;
            ra[q] = (ar < n && ak < n) ? o.A[(long)ar * n + ak] : 0.0;
            rb[q] = (bk < n && bcol < n) ? o.B[(long)bk * n + bcol] : 0.0;
        }
    };
    double4_v acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = double4_v{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    for (int k0 = 0; k0 < n; k0 += kGemmK) {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = t + 256 * q;
            As[e % kGemmK][e / kGemmK] = ra[q];
            Bs[e / kGemmT][e % kGemmT] = rb[q];
        }
        __syncthreads();
        if (k0 + kGemmK < n) fetch(k0 + kGemmK);
#pragma unroll
        for (int kk = 0; kk < kGemmK; kk += 4) {
            const double a0 = As[kk + lk][wr + lr], a1 = As[kk + lk][wr + 16 + lr];
            const double b0 = Bs[kk + lk][wc + lr], b1 = Bs[kk + lk][wc + 16 + lr];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int rg = 0; rg < 4; ++rg) {
                const int r = r0 + wr + 16 * i + lk + 4 * rg, c = c0 + wc + 16 * j + lr;
                if (r >= n || c >= n) continue;
                const double v = o.alpha * acc[i][j][rg] + (o.D ? o.beta * o.D[(long)r * n + c] : 0.0);
                if (!sym) {
                    o.C[(long)r * n + c] = v;
                } else if (r <= c) {   // the upper triangle, mirrored
                    o.C[(long)r * n + c] = v;
                    o.C[(long)c * n + r] = v;
                }
            }
}

void launch_dgemm(const GemmOps* d, int nflows, int max_n, bool sym, hipStream_t s) {
    const int nt = (max_n + kGemmT - 1) / kGemmT;
    if (nflows > 0 && nt > 0)
        hipLaunchKernelGGL(k_dgemm, dim3(nt * nt, nflows), dim3(256), 0, s, d, nt, sym ? 1 : 0);
}

void launch_interp(const RegOps* d, int nflows, int max_comp, hipStream_t s) {
    hipLaunchKernelGGL(k_interp, dim3(blocks_for(max_comp), nflows), dim3(kThreads), 0, s, d);
}
void launch_smear_curl(const RegOps* d, int nflows, int max_nodes, hipStream_t s) {
    hipLaunchKernelGGL(k_smear_curl, dim3(blocks_for(max_nodes), nflows), dim3(kThreads), 0, s, d);
}
void launch_clear_nodes(const RegOps* d, int nflows, int max_nodes, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_nodes, dim3(blocks_for(max_nodes), nflows), dim3(kThreads), 0, s, d);
}
void launch_minv(const MinvOps* d, int nflows, int max_n, hipStream_t s) {
    const int rows_per_block = kThreads / 64;
    hipLaunchKernelGGL(k_minv, dim3((max_n + rows_per_block - 1) / rows_per_block, nflows),
                       dim3(kThreads), 0, s, d);
}
void launch_cg(const CgOps* d, int nflows, int mode, hipStream_t s) {
    hipLaunchKernelGGL(k_cg, dim3(nflows), dim3(kThreads), 0, s, d, mode);
}

void launch_pcg_begin(const PcgOps* d, int nflows, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_begin, dim3(nflows), dim3(kThreads), 0, s, d);
}
void launch_pcg_step(const PcgOps* d, int nflows, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_step, dim3(nflows), dim3(kThreads), 0, s, d);
}
void launch_pcg_dir(const PcgOps* d, int nflows, int begin, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_dir, dim3(nflows), dim3(kThreads), 0, s, d, begin);
}

void launch_reg_to_flux(const RegFluxOps* d, int nslots, const double* f, double* q, hipStream_t s) {
    hipLaunchKernelGGL(k_reg_to_flux, dim3(blocks_for(std::max(nslots, 1))), dim3(kThreads), 0, s, d, f, q);
}
void launch_reg_to_boundary(const RegFluxOps* d, int nb, const double* q, double* f, hipStream_t s) {
    hipLaunchKernelGGL(k_reg_to_boundary, dim3(blocks_for(std::max(2 * nb, 1))), dim3(kThreads), 0, s, d, q, f);
}

void launch_forces(const ForceOps* d, int nflows, hipStream_t s) {
    hipLaunchKernelGGL(k_forces, dim3(nflows), dim3(64), 0, s, d);
}

}  // namespace ibpmg
