/*
 * ibpm_oracle.c -- CPU FP64 restatement of the ibpm projection-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it; the product never does.
 *
 * Each function cites the reference file:line it restates.  The reference's
 * FFTW dependency (EllipticSolver2d.cc:34-35,65) is replaced by an FFT-based
 * DST-I written here (og_dst2d), pinned against direct summation
 * (og_dst2d_direct) and scipy.fft.dstn(type=1) in tests/test_oracle.py.
 */
#define _GNU_SOURCE   /* sincos */
#include "ibpm_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>

/* sin and cos of one argument from glibc's sincos: gcc compiles the
   reference's sin(a) ... cos(a) pairs into one sincos call, which differs from
   separate sin and cos by an ulp for about one argument in 700. */
static void sin_cos(double a, double* s, double* c) { sincos(a, s, c); }

/* ======================================================================= */
/* Complex FFT (mixed-radix Stockham, exact twiddle table)                 */
/* ======================================================================= */
typedef struct { double re, im; } cpx;

typedef struct {
    int M;
    int nst;
    int rad[40];
    cpx* tw;     /* tw[k] = exp(-2 pi i k / M) */
    cpx* a;      /* work buffers */
    cpx* b;
} fftplan;

static fftplan* g_plans[64];
static int g_nplans = 0;

static fftplan* plan_get(int M) {
    for (int i = 0; i < g_nplans; ++i)
        if (g_plans[i]->M == M) return g_plans[i];
    fftplan* p = (fftplan*)calloc(1, sizeof(fftplan));
    p->M = M;
    int m = M;
    while (m % 4 == 0) { p->rad[p->nst++] = 4; m /= 4; }
    while (m % 2 == 0) { p->rad[p->nst++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) { p->rad[p->nst++] = f; m /= f; }
    }
    p->tw = (cpx*)malloc(sizeof(cpx) * M);
    const double twopi = 8.0 * atan(1.0);
    for (int k = 0; k < M; ++k) {
        double ang = -twopi * (double)k / (double)M;
        p->tw[k].re = cos(ang);
        p->tw[k].im = sin(ang);
    }
    p->a = (cpx*)malloc(sizeof(cpx) * M);
    p->b = (cpx*)malloc(sizeof(cpx) * M);
    if (g_nplans < 64) g_plans[g_nplans++] = p;
    return p;
}

/* In-place forward DFT of p->a (result left in p->a). */
static void fft_exec(fftplan* p) {
    const int M = p->M;
    cpx* x = p->a;
    cpx* y = p->b;
    int Ns = 1;
    cpx v[64];
    for (int s = 0; s < p->nst; ++s) {
        const int R = p->rad[s];
        const int MR = M / R;
        const int tstep = M / (Ns * R);
        for (int j = 0; j < MR; ++j) {
            const int jm = j % Ns;
            for (int r = 0; r < R; ++r) {
                cpx in = x[j + r * MR];
                int ti = (int)(((long)r * jm * tstep) % M);
                cpx w = p->tw[ti];
                v[r].re = in.re * w.re - in.im * w.im;
                v[r].im = in.re * w.im + in.im * w.re;
            }
            const int idxD = (j / Ns) * Ns * R + jm;
            if (R == 2) {
                cpx a0 = v[0], a1 = v[1];
                y[idxD].re = a0.re + a1.re;           y[idxD].im = a0.im + a1.im;
                y[idxD + Ns].re = a0.re - a1.re;      y[idxD + Ns].im = a0.im - a1.im;
            } else if (R == 4) {
                cpx a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
                double s02r = a0.re + a2.re, s02i = a0.im + a2.im;
                double d02r = a0.re - a2.re, d02i = a0.im - a2.im;
                double s13r = a1.re + a3.re, s13i = a1.im + a3.im;
                double d13r = a1.re - a3.re, d13i = a1.im - a3.im;
                y[idxD].re = s02r + s13r;           y[idxD].im = s02i + s13i;
                /* -i * d13 = (d13i, -d13r) */
                y[idxD + Ns].re = d02r + d13i;      y[idxD + Ns].im = d02i - d13r;
                y[idxD + 2 * Ns].re = s02r - s13r;  y[idxD + 2 * Ns].im = s02i - s13i;
                y[idxD + 3 * Ns].re = d02r - d13i;  y[idxD + 3 * Ns].im = d02i + d13r;
            } else {
                const int wstep = M / R;
                for (int qd = 0; qd < R; ++qd) {
                    double sr = 0.0, si = 0.0;
                    for (int r = 0; r < R; ++r) {
                        cpx w = p->tw[(int)(((long)r * qd * wstep) % M)];
                        sr += v[r].re * w.re - v[r].im * w.im;
                        si += v[r].re * w.im + v[r].im * w.re;
                    }
                    y[idxD + qd * Ns].re = sr;
                    y[idxD + qd * Ns].im = si;
                }
            }
        }
        cpx* t = x; x = y; y = t;
        Ns *= R;
    }
    if (x != p->a) memcpy(p->a, x, sizeof(cpx) * M);
}

/* DST-I (RODFT00, unnormalised: Y_k = 2 sum_j X_j sin(pi (j+1)(k+1)/(N+1)))
   of `count` vectors of length N; element k of vector v is at
   data[v*vstride + k*estride].  Two vectors share one complex FFT of length
   2(N+1) (odd extension; real part -> vector a, imaginary part -> vector b). */
static void dst1_many(int N, int count, double* data, long vstride, long estride) {
    const int n = N + 1;
    const int M = 2 * n;
    fftplan* p = plan_get(M);
    for (int v = 0; v < count; v += 2) {
        const int hasb = (v + 1 < count);
        double* ra = data + (long)v * vstride;
        double* rb = hasb ? data + (long)(v + 1) * vstride : NULL;
        cpx* z = p->a;
        z[0].re = z[0].im = 0.0;
        z[n].re = z[n].im = 0.0;
        for (int m = 1; m <= N; ++m) {
            double av = ra[(long)(m - 1) * estride];
            double bv = hasb ? rb[(long)(m - 1) * estride] : 0.0;
            z[m].re = av;        z[m].im = bv;
            z[M - m].re = -av;   z[M - m].im = -bv;
        }
        fft_exec(p);
        for (int k = 0; k < N; ++k) {
            ra[(long)k * estride] = -z[k + 1].im;
            if (hasb) rb[(long)k * estride] = z[k + 1].re;
        }
    }
}

void og_dst2d(int n1, int n2, const double* in, double* out) {
    if (out != in) memcpy(out, in, sizeof(double) * (size_t)n1 * n2);
    dst1_many(n2, n1, out, n2, 1);   /* along j (contiguous) for every row */
    dst1_many(n1, n2, out, 1, n2);   /* along i for every column */
}

void og_dst2d_direct(int n1, int n2, const double* in, double* out) {
    const double pi = 4.0 * atan(1.0);
    double* s1 = (double*)malloc(sizeof(double) * 2 * (n1 + 1));
    double* s2 = (double*)malloc(sizeof(double) * 2 * (n2 + 1));
    for (int t = 0; t < 2 * (n1 + 1); ++t) s1[t] = sin(pi * t / (n1 + 1));
    for (int t = 0; t < 2 * (n2 + 1); ++t) s2[t] = sin(pi * t / (n2 + 1));
    double* tmp = (double*)malloc(sizeof(double) * (size_t)n1 * n2);
    for (int i = 0; i < n1; ++i)
        for (int l = 0; l < n2; ++l) {
            double s = 0.0;
            for (int j = 0; j < n2; ++j)
                s += in[(long)i * n2 + j] * s2[((j + 1) * (l + 1)) % (2 * (n2 + 1))];
            tmp[(long)i * n2 + l] = 2.0 * s;
        }
    for (int k = 0; k < n1; ++k)
        for (int l = 0; l < n2; ++l) {
            double s = 0.0;
            for (int i = 0; i < n1; ++i)
                s += tmp[(long)i * n2 + l] * s1[((i + 1) * (k + 1)) % (2 * (n1 + 1))];
            out[(long)k * n2 + l] = 2.0 * s;
        }
    free(tmp); free(s1); free(s2);
}

/* ======================================================================= */
/* Index helpers                                                           */
/* ======================================================================= */
#define SN(nx, ny) ((long)((nx) - 1) * ((ny) - 1))
#define SI(nx, ny, lev, i, j) ((long)(lev) * SN(nx, ny) + (long)((i) - 1) * ((ny) - 1) + ((j) - 1))
#define NFX(nx, ny) ((long)((nx) + 1) * (ny))
#define NF(nx, ny) (2L * (nx) * (ny) + (nx) + (ny))
#define FXI(nx, ny, lev, i, j) ((long)(lev) * NF(nx, ny) + (long)(i) * (ny) + (j))
#define FYI(nx, ny, lev, i, j) ((long)(lev) * NF(nx, ny) + NFX(nx, ny) + (long)(i) * ((ny) + 1) + (j))

/* BC ring accessors, BC.h:22-79 */
#define BC_LEFT(bc, nx, ny, j)   ((bc)[(j)])
#define BC_TOP(bc, nx, ny, i)    ((bc)[(ny) + (i)])
#define BC_RIGHT(bc, nx, ny, j)  ((bc)[2 * (ny) + (nx) - (j)])
#define BC_BOTTOM(bc, nx, ny, i) ((bc)[(i) == 0 ? 0 : 2 * ((nx) + (ny)) - (i)])

static inline int nxext(int nx) { return (int)floor(nx / 4 * (1 - 0.0) + 0.5); } /* Grid.h:89-91, shift 0 */

/* Scalar::coarsify, Scalar.cc:46-63 */
void og_coarsify(int nx, int ny, int ngrid, double* s) {
    const int ex = nxext(nx), ey = nxext(ny);
    for (int lev = 1; lev < ngrid; ++lev) {
        for (int i = ex + 1; i < nx / 2 + ex; ++i) {
            for (int j = ey + 1; j < ny / 2 + ey; ++j) {
                const int ii = (i - ex) * 2, jj = (j - ey) * 2;   /* Grid::c2f */
                const int lf = lev - 1;
                s[SI(nx, ny, lev, i, j)] = 0.25 * s[SI(nx, ny, lf, ii, jj)]
                    + 0.125 * (s[SI(nx, ny, lf, ii + 1, jj)] + s[SI(nx, ny, lf, ii, jj + 1)]
                             + s[SI(nx, ny, lf, ii - 1, jj)] + s[SI(nx, ny, lf, ii, jj - 1)])
                    + 0.0625 * (s[SI(nx, ny, lf, ii + 1, jj + 1)] + s[SI(nx, ny, lf, ii + 1, jj - 1)]
                              + s[SI(nx, ny, lf, ii - 1, jj + 1)] + s[SI(nx, ny, lf, ii - 1, jj - 1)]);
            }
        }
    }
}

/* Scalar::getBC, Scalar.cc:91-127 : boundary of level lev from level lev+1 */
void og_getbc(int nx, int ny, int ngrid, const double* s, int lev, double* bc) {
    (void)ngrid;
    const int ex = nxext(nx), ey = nxext(ny);
    const int c = lev + 1;
    for (int i = 0; i <= nx; ++i) {
        const int ii = i / 2 + ex, jj = 0 / 2 + ey;   /* Grid::f2c(i,0) */
        BC_BOTTOM(bc, nx, ny, i) = s[SI(nx, ny, c, ii, jj)];
        BC_TOP(bc, nx, ny, i) = s[SI(nx, ny, c, ii, ny / 2 + jj)];
        if (++i <= nx) {
            BC_BOTTOM(bc, nx, ny, i) = 0.5 * (s[SI(nx, ny, c, ii, jj)] + s[SI(nx, ny, c, ii + 1, jj)]);
            BC_TOP(bc, nx, ny, i) = 0.5 * (s[SI(nx, ny, c, ii, ny / 2 + jj)] + s[SI(nx, ny, c, ii + 1, ny / 2 + jj)]);
        }
    }
    for (int j = 0; j <= ny; ++j) {
        const int ii = 0 / 2 + ex, jj = j / 2 + ey;   /* Grid::f2c(0,j) */
        BC_LEFT(bc, nx, ny, j) = s[SI(nx, ny, c, ii, jj)];
        BC_RIGHT(bc, nx, ny, j) = s[SI(nx, ny, c, nx / 2 + ii, jj)];
        if (++j <= ny) {
            BC_LEFT(bc, nx, ny, j) = 0.5 * (s[SI(nx, ny, c, ii, jj)] + s[SI(nx, ny, c, ii, jj + 1)]);
            BC_RIGHT(bc, nx, ny, j) = 0.5 * (s[SI(nx, ny, c, nx / 2 + ii, jj)] + s[SI(nx, ny, c, nx / 2 + ii, jj + 1)]);
        }
    }
}

/* ======================================================================= */
/* Elliptic solvers (EllipticSolver2d.cc, EllipticSolver.cc)                */
/* ======================================================================= */
typedef struct { int kind, nx, ny; double dx, alpha; double* eig; } eigtab;
static eigtab g_eig[128];
static int g_neig = 0;

/* EllipticSolver2d.cc:42-54 (Laplacian eigenvalues), :115-122 (Poisson 1/lambda),
   :156-163 (Helmholtz 1/(1+alpha lambda)) */
static const double* eig_get(int kind, int nx, int ny, double dx, double alpha) {
    for (int k = 0; k < g_neig; ++k) {
        eigtab* e = &g_eig[k];
        if (e->kind == kind && e->nx == nx && e->ny == ny && e->dx == dx && e->alpha == alpha)
            return e->eig;
    }
    double* t = (double*)malloc(sizeof(double) * SN(nx, ny));
    const double pi = 4. * atan(1.0);
    const double bydx2 = 1. / (dx * dx);
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j) {
            double lam = 2. * (cos((pi * i) / nx) + cos((pi * j) / ny) - 2.) * bydx2;
            t[(long)(i - 1) * (ny - 1) + (j - 1)] = (kind == 0) ? 1. / lam : 1. / (1 + alpha * lam);
        }
    if (g_neig < 128) {
        eigtab* e = &g_eig[g_neig++];
        e->kind = kind; e->nx = nx; e->ny = ny; e->dx = dx; e->alpha = alpha; e->eig = t;
    }
    return t;
}

void og_elliptic2d_solve(int kind, int nx, int ny, double dx, double alpha,
                         const double* f, const double* bc, double* u) {
    const long n = SN(nx, ny);
    if (u != f) memcpy(u, f, sizeof(double) * n);
    if (bc) {   /* getRHS, EllipticSolver2d.cc:125-144 / :166-185 */
        const double c = (kind == 0) ? 1 / (dx * dx) : alpha / (dx * dx);
        for (int i = 1; i < nx; ++i) {
            u[(long)(i - 1) * (ny - 1) + 0] -= BC_BOTTOM(bc, nx, ny, i) * c;
            u[(long)(i - 1) * (ny - 1) + (ny - 2)] -= BC_TOP(bc, nx, ny, i) * c;
        }
        for (int j = 1; j < ny; ++j) {
            u[(long)0 * (ny - 1) + (j - 1)] -= BC_LEFT(bc, nx, ny, j) * c;
            u[(long)(nx - 2) * (ny - 1) + (j - 1)] -= BC_RIGHT(bc, nx, ny, j) * c;
        }
    }
    const double* eig = eig_get(kind, nx, ny, dx, alpha);
    og_dst2d(nx - 1, ny - 1, u, u);                 /* sinTransform */
    for (long k = 0; k < n; ++k) u[k] *= eig[k];   /* :89-93 */
    og_dst2d(nx - 1, ny - 1, u, u);                 /* sinTransformInv :76-84 */
    const double norm = 1. / (2 * nx * 2 * ny);
    for (long k = 0; k < n; ++k) u[k] *= norm;
}

void og_elliptic_solve(int kind, int nx, int ny, int ngrid, double dx, double alpha,
                       const double* f, double* u) {
    const long n = SN(nx, ny);
    double* rhs = (double*)malloc(sizeof(double) * n * ngrid);
    memcpy(rhs, f, sizeof(double) * n * ngrid);
    og_coarsify(nx, ny, ngrid, rhs);
    double* bc = (double*)malloc(sizeof(double) * 2 * (nx + ny));
    for (int lev = ngrid - 1; lev >= 0; --lev) {
        const double dxl = dx * (1 << lev);
        if (lev == ngrid - 1) {
            og_elliptic2d_solve(kind, nx, ny, dxl, alpha, rhs + lev * n, NULL, u + lev * n);
        } else {
            og_getbc(nx, ny, ngrid, u, lev, bc);
            og_elliptic2d_solve(kind, nx, ny, dxl, alpha, rhs + lev * n, bc, u + lev * n);
        }
    }
    free(bc);
    free(rhs);
}

/* ======================================================================= */
/* Grid operators (VectorOperations.cc)                                     */
/* ======================================================================= */
void og_curl_flux(int nx, int ny, int ngrid, double dx, const double* q, double* w) {
    for (int lev = 0; lev < ngrid; ++lev) {
        const double d = dx * (1 << lev);
        const double bydx2 = 1. / (d * d);
        for (int i = 1; i < nx; ++i)
            for (int j = 1; j < ny; ++j)
                w[SI(nx, ny, lev, i, j)] = (q[FYI(nx, ny, lev, i, j)] - q[FYI(nx, ny, lev, i - 1, j)]
                    + q[FXI(nx, ny, lev, i, j - 1)] - q[FXI(nx, ny, lev, i, j)]) * bydx2;
    }
}

void og_curl_scalar(int nx, int ny, int ngrid, const double* f, double* q) {
    double* bc = (double*)calloc(2 * (nx + ny), sizeof(double));
    for (int lev = ngrid - 1; lev >= 0; --lev) {
        if (lev == ngrid - 1) memset(bc, 0, sizeof(double) * 2 * (nx + ny));
        else og_getbc(nx, ny, ngrid, f, lev, bc);
        for (int j = 1; j < ny - 1; ++j)
            for (int i = 1; i < nx; ++i)
                q[FXI(nx, ny, lev, i, j)] = f[SI(nx, ny, lev, i, j + 1)] - f[SI(nx, ny, lev, i, j)];
        for (int i = 1; i < nx; ++i) {
            q[FXI(nx, ny, lev, i, 0)] = f[SI(nx, ny, lev, i, 1)] - BC_BOTTOM(bc, nx, ny, i);
            q[FXI(nx, ny, lev, i, ny - 1)] = BC_TOP(bc, nx, ny, i) - f[SI(nx, ny, lev, i, ny - 1)];
        }
        for (int j = 0; j < ny; ++j) {
            q[FXI(nx, ny, lev, 0, j)] = BC_LEFT(bc, nx, ny, j + 1) - BC_LEFT(bc, nx, ny, j);
            q[FXI(nx, ny, lev, nx, j)] = BC_RIGHT(bc, nx, ny, j + 1) - BC_RIGHT(bc, nx, ny, j);
        }
        for (int j = 1; j < ny; ++j)
            for (int i = 1; i < nx - 1; ++i)
                q[FYI(nx, ny, lev, i, j)] = f[SI(nx, ny, lev, i, j)] - f[SI(nx, ny, lev, i + 1, j)];
        for (int j = 1; j < ny; ++j) {
            q[FYI(nx, ny, lev, 0, j)] = BC_LEFT(bc, nx, ny, j) - f[SI(nx, ny, lev, 1, j)];
            q[FYI(nx, ny, lev, nx - 1, j)] = f[SI(nx, ny, lev, nx - 1, j)] - BC_RIGHT(bc, nx, ny, j);
        }
        for (int i = 0; i < nx; ++i) {
            q[FYI(nx, ny, lev, i, 0)] = BC_BOTTOM(bc, nx, ny, i) - BC_BOTTOM(bc, nx, ny, i + 1);
            q[FYI(nx, ny, lev, i, ny)] = BC_TOP(bc, nx, ny, i) - BC_TOP(bc, nx, ny, i + 1);
        }
    }
    free(bc);
}

void og_laplacian(int nx, int ny, int ngrid, double dx, const double* w, double* out) {
    double* q = (double*)malloc(sizeof(double) * NF(nx, ny) * ngrid);
    og_curl_scalar(nx, ny, ngrid, w, q);
    og_curl_flux(nx, ny, ngrid, dx, q, out);
    for (long k = 0; k < SN(nx, ny) * ngrid; ++k) out[k] *= -1;
    free(q);
}

/* Single-level 5-point Laplacian with BC, VectorOperations.cc:176-214 */
void og_laplacian2d_bc(int nx, int ny, double dx, const double* f, const double* bc, double* g) {
    const double bydx2 = 1. / (dx * dx);
#define F(i, j) f[(long)((i) - 1) * (ny - 1) + ((j) - 1)]
#define G(i, j) g[(long)((i) - 1) * (ny - 1) + ((j) - 1)]
    for (int i = 2; i < nx - 1; ++i)
        for (int j = 2; j < ny - 1; ++j)
            G(i, j) = (F(i - 1, j) + F(i + 1, j) + F(i, j - 1) + F(i, j + 1) - 4 * F(i, j)) * bydx2;
    for (int j = 2; j < ny - 1; ++j) {
        G(1, j) = (BC_LEFT(bc, nx, ny, j) + F(2, j) + F(1, j + 1) + F(1, j - 1) - 4 * F(1, j)) * bydx2;
        G(nx - 1, j) = (BC_RIGHT(bc, nx, ny, j) + F(nx - 2, j) + F(nx - 1, j + 1) + F(nx - 1, j - 1) - 4 * F(nx - 1, j)) * bydx2;
    }
    for (int i = 2; i < nx - 1; ++i) {
        G(i, 1) = (BC_BOTTOM(bc, nx, ny, i) + F(i, 2) + F(i + 1, 1) + F(i - 1, 1) - 4 * F(i, 1)) * bydx2;
        G(i, ny - 1) = (BC_TOP(bc, nx, ny, i) + F(i, ny - 2) + F(i + 1, ny - 1) + F(i - 1, ny - 1) - 4 * F(i, ny - 1)) * bydx2;
    }
    G(1, 1) = (BC_LEFT(bc, nx, ny, 1) + BC_BOTTOM(bc, nx, ny, 1) + F(1, 2) + F(2, 1) - 4 * F(1, 1)) * bydx2;
    G(nx - 1, 1) = (BC_RIGHT(bc, nx, ny, 1) + BC_BOTTOM(bc, nx, ny, nx - 1) + F(nx - 1, 2) + F(nx - 2, 1) - 4 * F(nx - 1, 1)) * bydx2;
    G(1, ny - 1) = (BC_LEFT(bc, nx, ny, ny - 1) + BC_TOP(bc, nx, ny, 1) + F(1, ny - 2) + F(2, ny - 1) - 4 * F(1, ny - 1)) * bydx2;
    G(nx - 1, ny - 1) = (BC_RIGHT(bc, nx, ny, ny - 1) + BC_TOP(bc, nx, ny, nx - 1) + F(nx - 2, ny - 1) + F(nx - 1, ny - 2) - 4 * F(nx - 1, ny - 1)) * bydx2;
#undef F
#undef G
}

/* InnerProduct(Scalar, Scalar), VectorOperations.cc:223-292: fine grid plus
   the coarse-level regions outside the finer domain, interface nodes weighted
   3/4 and interface corners 15/16. */
double og_ip_scalar(int nx, int ny, int ngrid, double dx0, const double* f, const double* g) {
    const int nx2 = nxext(nx), ny2 = nxext(ny);
#define P(l, i, j) (f[SI(nx, ny, l, i, j)] * g[SI(nx, ny, l, i, j)])
    double ip = 0.;
    double dx2 = dx0 * dx0;
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j) ip += P(0, i, j) * dx2;
    for (int lev = 1; lev < ngrid; ++lev) {
        const double d = dx0 * (1 << lev);
        dx2 = d * d;
        ip += P(lev, nx2, ny2) * dx2 * 15. / 16;
        ip += P(lev, nx / 2 + nx2, ny2) * dx2 * 15. / 16;
        ip += P(lev, nx2, ny / 2 + ny2) * dx2 * 15. / 16;
        ip += P(lev, nx / 2 + nx2, ny / 2 + ny2) * dx2 * 15. / 16;
        for (int j = ny2 + 1; j < ny / 2 + ny2; ++j) {
            ip += P(lev, nx2, j) * dx2 * 0.75;
            ip += P(lev, nx / 2 + nx2, j) * dx2 * 0.75;
        }
        for (int i = nx2 + 1; i < nx / 2 + nx2; ++i) {
            ip += P(lev, i, ny2) * dx2 * 0.75;
            ip += P(lev, i, ny / 2 + ny2) * dx2 * 0.75;
        }
        for (int i = 1; i < nx2; ++i)
            for (int j = 1; j < ny; ++j) ip += P(lev, i, j) * dx2;
        for (int i = nx / 2 + nx2 + 1; i < nx; ++i)
            for (int j = 1; j < ny; ++j) ip += P(lev, i, j) * dx2;
        for (int i = nx2; i < nx / 2 + nx2 + 1; ++i) {
            for (int j = 1; j < ny2; ++j) ip += P(lev, i, j) * dx2;
            for (int j = ny / 2 + ny2 + 1; j < ny; ++j) ip += P(lev, i, j) * dx2;
        }
    }
#undef P
    return ip;
}

/* InnerProduct(Flux, Flux), VectorOperations.cc:296-405 */
double og_ip_flux(int nx, int ny, int ngrid, const double* p, const double* q) {
    const int nx2 = nxext(nx), ny2 = nxext(ny);
#define PX(l, i, j) (p[FXI(nx, ny, l, i, j)] * q[FXI(nx, ny, l, i, j)])
#define PY(l, i, j) (p[FYI(nx, ny, l, i, j)] * q[FYI(nx, ny, l, i, j)])
    double ip = 0.;
    for (int j = 0; j < ny; ++j)
        for (int i = 1; i < nx; ++i) ip += PX(0, i, j);
    for (int i = 0; i < nx; ++i)
        for (int j = 1; j < ny; ++j) ip += PY(0, i, j);
    for (int lev = 1; lev < ngrid; ++lev) {
        for (int j = ny2; j < ny / 2 + ny2; ++j) {
            ip += PX(lev, nx2, j) * 0.75;
            ip += PX(lev, nx / 2 + nx2, j) * 0.75;
        }
        for (int j = 0; j < ny; ++j) {
            for (int i = 1; i < nx2; ++i) ip += PX(lev, i, j);
            for (int i = nx / 2 + nx2 + 1; i < nx; ++i) ip += PX(lev, i, j);
        }
        for (int i = nx2; i < nx / 2 + nx2 + 1; ++i) {
            for (int j = 0; j < ny2; ++j) ip += PX(lev, i, j);
            for (int j = ny / 2 + ny2; j < ny; ++j) ip += PX(lev, i, j);
        }
    }
    for (int lev = 1; lev < ngrid; ++lev) {
        for (int i = nx2; i < nx / 2 + nx2; ++i) {
            ip += PY(lev, i, ny2) * 0.75;
            ip += PY(lev, i, ny / 2 + ny2) * 0.75;
        }
        for (int i = 0; i < nx; ++i) {
            for (int j = 1; j < ny2; ++j) ip += PY(lev, i, j);
            for (int j = ny / 2 + ny2 + 1; j < ny; ++j) ip += PY(lev, i, j);
        }
        for (int j = ny2; j < ny / 2 + ny2 + 1; ++j) {
            for (int i = 0; i < nx2; ++i) ip += PY(lev, i, j);
            for (int i = nx / 2 + nx2; i < nx; ++i) ip += PY(lev, i, j);
        }
    }
#undef PX
#undef PY
    return ip;
}

/* FluxToXVelocity, VectorOperations.cc:488-577 */
void og_flux_to_xvel(int nx, int ny, int ngrid, double dx, const double* q, double* u) {
    const int nx2 = nxext(nx), ny2 = nxext(ny);
    const double oneOver2Delta = 1. / (2 * dx);
#define Q(l, i, j) q[FXI(nx, ny, l, i, j)]
#define U(l, i, j) u[SI(nx, ny, l, i, j)]
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            U(0, i, j) = (Q(0, i, j) + Q(0, i, j - 1)) * oneOver2Delta;
    for (int lev = 1; lev < ngrid; ++lev) {
        const double bydx = 1. / (dx * (1 << lev));
        for (int j = 1; j < ny; ++j) {
            for (int i = 1; i < nx2; ++i) U(lev, i, j) = (Q(lev, i, j) + Q(lev, i, j - 1)) * 0.5 * bydx;
            for (int i = nx / 2 + nx2 + 1; i < nx; ++i) U(lev, i, j) = (Q(lev, i, j) + Q(lev, i, j - 1)) * 0.5 * bydx;
        }
        for (int i = nx2; i <= nx / 2 + nx2; ++i) {
            for (int j = 1; j < ny2; ++j) U(lev, i, j) = (Q(lev, i, j) + Q(lev, i, j - 1)) * 0.5 * bydx;
            for (int j = ny / 2 + ny2 + 1; j < ny; ++j) U(lev, i, j) = (Q(lev, i, j) + Q(lev, i, j - 1)) * 0.5 * bydx;
        }
        for (int j = ny2 + 1; j < ny / 2 + ny2; ++j) {
            U(lev, nx2, j) = (Q(lev, nx2, j) + Q(lev, nx2, j - 1)) * 0.5 * bydx;
            U(lev, nx / 2 + nx2, j) = (Q(lev, nx / 2 + nx2, j) + Q(lev, nx / 2 + nx2, j - 1)) * 0.5 * bydx;
        }
        for (int i = nx2 + 1; i < nx / 2 + nx2; ++i) {
            int ii = (i - nx2) * 2, jj = (ny2 - ny2) * 2;
            U(lev, i, ny2) = (Q(lev, i, ny2 - 1) * 2. / 3 + Q(lev - 1, ii, jj) * 1. / 3
                + (Q(lev - 1, ii - 1, jj) + Q(lev - 1, ii + 1, jj)) * 1. / 6) * bydx;
            jj = (ny / 2 + ny2 - ny2) * 2;
            U(lev, i, ny / 2 + ny2) = (Q(lev, i, ny / 2 + ny2) * 2. / 3 + Q(lev - 1, ii, jj - 1) * 1. / 3
                + (Q(lev - 1, ii - 1, jj - 1) + Q(lev - 1, ii + 1, jj - 1)) * 1. / 6) * bydx;
        }
        int i = nx2, j = ny2;
        U(lev, i, j) = (Q(lev, i, j - 1) * 8. / 15 + Q(lev, i, j) * 6. / 15 + Q(lev - 1, 1, 0) * 2. / 15) * bydx;
        i = nx / 2 + nx2;
        U(lev, i, j) = (Q(lev, i, j - 1) * 8. / 15 + Q(lev, i, j) * 6. / 15 + Q(lev - 1, nx - 1, 0) * 2. / 15) * bydx;
        i = nx2; j = ny / 2 + ny2;
        U(lev, i, j) = (Q(lev, i, j) * 8. / 15 + Q(lev, i, j - 1) * 6. / 15 + Q(lev - 1, 1, ny - 1) * 2. / 15) * bydx;
        i = nx / 2 + nx2;
        U(lev, i, j) = (Q(lev, i, j) * 8. / 15 + Q(lev, i, j - 1) * 6. / 15 + Q(lev - 1, nx - 1, ny - 1) * 2. / 15) * bydx;
    }
#undef Q
#undef U
}

/* FluxToYVelocity, VectorOperations.cc:579-665 */
void og_flux_to_yvel(int nx, int ny, int ngrid, double dx, const double* q, double* v) {
    const int nx2 = nxext(nx), ny2 = nxext(ny);
    const double oneOver2Delta = 1. / (2 * dx);
#define Q(l, i, j) q[FYI(nx, ny, l, i, j)]
#define V(l, i, j) v[SI(nx, ny, l, i, j)]
    for (int j = 1; j < ny; ++j)
        for (int i = 1; i < nx; ++i)
            V(0, i, j) = (Q(0, i - 1, j) + Q(0, i, j)) * oneOver2Delta;
    for (int lev = 1; lev < ngrid; ++lev) {
        const double bydx = 1. / (dx * (1 << lev));
        for (int i = 1; i < nx; ++i) {
            for (int j = 1; j < ny2; ++j) V(lev, i, j) = (Q(lev, i, j) + Q(lev, i - 1, j)) * 0.5 * bydx;
            for (int j = ny / 2 + ny2 + 1; j < ny; ++j) V(lev, i, j) = (Q(lev, i, j) + Q(lev, i - 1, j)) * 0.5 * bydx;
        }
        for (int j = ny2; j <= ny / 2 + ny2; ++j) {
            for (int i = 1; i < nx2; ++i) V(lev, i, j) = (Q(lev, i, j) + Q(lev, i - 1, j)) * 0.5 * bydx;
            for (int i = nx / 2 + nx2 + 1; i < nx; ++i) V(lev, i, j) = (Q(lev, i, j) + Q(lev, i - 1, j)) * 0.5 * bydx;
        }
        for (int i = nx2 + 1; i < nx / 2 + nx2; ++i) {
            V(lev, i, ny2) = (Q(lev, i, ny2) + Q(lev, i - 1, ny2)) * 0.5 * bydx;
            V(lev, i, ny / 2 + ny2) = (Q(lev, i, ny / 2 + ny2) + Q(lev, i - 1, ny / 2 + ny2)) * 0.5 * bydx;
        }
        for (int j = ny2 + 1; j < ny / 2 + ny2; ++j) {
            int ii = (nx2 - nx2) * 2, jj = (j - ny2) * 2;
            V(lev, nx2, j) = (Q(lev, nx2 - 1, j) * 2. / 3 + Q(lev - 1, ii, jj) * 1. / 3
                + (Q(lev - 1, ii, jj - 1) + Q(lev - 1, ii, jj + 1)) * 1. / 6) * bydx;
            ii = (nx / 2 + nx2 - nx2) * 2;
            V(lev, nx / 2 + nx2, j) = (Q(lev, nx / 2 + nx2, j) * 2. / 3 + Q(lev - 1, ii - 1, jj) * 1. / 3
                + (Q(lev - 1, ii - 1, jj - 1) + Q(lev - 1, ii - 1, jj + 1)) * 1. / 6) * bydx;
        }
        int j = ny2, i = nx2;
        V(lev, i, j) = (Q(lev, i - 1, j) * 8. / 15 + Q(lev, i, j) * 6. / 15 + Q(lev - 1, 0, 1) * 2. / 15) * bydx;
        j = ny / 2 + ny2;
        V(lev, i, j) = (Q(lev, i - 1, j) * 8. / 15 + Q(lev, i, j) * 6. / 15 + Q(lev - 1, 0, ny - 1) * 2. / 15) * bydx;
        j = ny2; i = nx / 2 + nx2;
        V(lev, i, j) = (Q(lev, i, j) * 8. / 15 + Q(lev, i - 1, j) * 6. / 15 + Q(lev - 1, nx - 1, 1) * 2. / 15) * bydx;
        j = ny / 2 + ny2;
        V(lev, i, j) = (Q(lev, i, j) * 8. / 15 + Q(lev, i - 1, j) * 6. / 15 + Q(lev - 1, nx - 1, ny - 1) * 2. / 15) * bydx;
    }
#undef Q
#undef V
}

/* XVelocityToFlux, VectorOperations.cc:669-791 (the redundant outer loop of
   :760-787 is executed once; it rewrites identical values in the reference) */
void og_xvel_to_flux(int nx, int ny, int ngrid, double dx0, const double* u, double* q) {
    const int nx2 = nxext(nx), ny2 = nxext(ny);
#define Q(l, i, j) q[FXI(nx, ny, l, i, j)]
#define U(l, i, j) u[SI(nx, ny, l, i, j)]
    for (int lev = 0; lev < ngrid; ++lev) {
        const double dx = dx0 * (1 << lev);
        if (lev == 0) {
            for (int i = 1; i < nx; ++i)
                for (int j = 1; j < ny - 1; ++j) Q(0, i, j) = (U(0, i, j) + U(0, i, j + 1)) * 0.5 * dx;
        } else {
            for (int i = 1; i < nx; ++i) {
                for (int j = 1; j < ny2; ++j) Q(lev, i, j) = (U(lev, i, j) + U(lev, i, j + 1)) * 0.5 * dx;
                for (int j = ny / 2 + ny2; j < ny - 1; ++j) Q(lev, i, j) = (U(lev, i, j) + U(lev, i, j + 1)) * 0.5 * dx;
            }
            for (int j = ny2; j < ny / 2 + ny2; ++j) {
                for (int i = 1; i <= nx2; ++i) Q(lev, i, j) = (U(lev, i, j) + U(lev, i, j + 1)) * 0.5 * dx;
                for (int i = nx / 2 + nx2; i < nx; ++i) Q(lev, i, j) = (U(lev, i, j) + U(lev, i, j + 1)) * 0.5 * dx;
            }
            for (int i = nx2 + 1; i < nx / 2 + nx2; ++i)
                for (int j = ny2; j < ny / 2 + ny2; ++j) {
                    int ii = (i - nx2) * 2, jj = (j - ny2) * 2;
                    Q(lev, i, j) = Q(lev - 1, ii, jj) + Q(lev - 1, ii, jj + 1);
                }
        }
        if (lev == ngrid - 1) {
            for (int j = 0; j < ny; ++j) { Q(lev, 0, j) = 0; Q(lev, nx, j) = 0; }
        } else {
            for (int j = 0; j < ny - 1; j += 2) {
                int jj = j / 2 + ny2;
                Q(lev, 0, j) = (0.75 * U(lev + 1, nx2, jj) + 0.25 * U(lev + 1, nx2, jj + 1)) * dx;
                Q(lev, nx, j) = (0.75 * U(lev + 1, nx / 2 + nx2, jj) + 0.25 * U(lev + 1, nx / 2 + nx2, jj + 1)) * dx;
                Q(lev, 0, j + 1) = (0.25 * U(lev + 1, nx2, jj) + 0.75 * U(lev + 1, nx2, jj + 1)) * dx;
                Q(lev, nx, j + 1) = (0.25 * U(lev + 1, nx / 2 + nx2, jj) + 0.75 * U(lev + 1, nx / 2 + nx2, jj + 1)) * dx;
            }
        }
        if (lev == ngrid - 1) {
            for (int i = 1; i < nx; ++i) {
                Q(lev, i, 0) = U(lev, i, 1) * 0.5 * dx;
                Q(lev, i, ny - 1) = U(lev, i, ny - 1) * 0.5 * dx;
            }
        } else {
            for (int i = 2; i < nx; i += 2) {
                int ii = i / 2 + nx2;
                Q(lev, i, 0) = (U(lev, i, 1) + U(lev + 1, ii, ny2)) * 0.5 * dx;
                Q(lev, i, ny - 1) = (U(lev, i, ny - 1) + U(lev + 1, ii, ny / 2 + ny2)) * 0.5 * dx;
            }
            for (int i = 1; i < nx; i += 2) {
                int ii = i / 2 + nx2;
                Q(lev, i, 0) = (0.5 * U(lev, i, 1) + 0.25 * U(lev + 1, ii, ny2) + 0.25 * U(lev + 1, ii + 1, ny2)) * dx;
                Q(lev, i, ny - 1) = (0.5 * U(lev, i, ny - 1) + 0.25 * U(lev + 1, ii, ny / 2 + ny2)
                    + 0.25 * U(lev + 1, ii + 1, ny / 2 + ny2)) * dx;
            }
        }
    }
#undef Q
#undef U
}

/* YVelocityToFlux, VectorOperations.cc:793-915 */
void og_yvel_to_flux(int nx, int ny, int ngrid, double dx0, const double* v, double* q) {
    const int nx2 = nxext(nx), ny2 = nxext(ny);
#define Q(l, i, j) q[FYI(nx, ny, l, i, j)]
#define V(l, i, j) v[SI(nx, ny, l, i, j)]
    for (int lev = 0; lev < ngrid; ++lev) {
        const double dx = dx0 * (1 << lev);
        if (lev == 0) {
            for (int j = 1; j < ny; ++j)
                for (int i = 1; i < nx - 1; ++i) Q(0, i, j) = (V(0, i, j) + V(0, i + 1, j)) * 0.5 * dx;
        } else {
            for (int j = 1; j < ny; ++j) {
                for (int i = 1; i < nx2; ++i) Q(lev, i, j) = (V(lev, i, j) + V(lev, i + 1, j)) * 0.5 * dx;
                for (int i = nx / 2 + nx2; i < nx - 1; ++i) Q(lev, i, j) = (V(lev, i, j) + V(lev, i + 1, j)) * 0.5 * dx;
            }
            for (int i = nx2; i < nx / 2 + nx2; ++i) {
                for (int j = 1; j <= ny2; ++j) Q(lev, i, j) = (V(lev, i, j) + V(lev, i + 1, j)) * 0.5 * dx;
                for (int j = ny / 2 + ny2; j < ny; ++j) Q(lev, i, j) = (V(lev, i, j) + V(lev, i + 1, j)) * 0.5 * dx;
            }
            for (int j = ny2 + 1; j < ny / 2 + ny2; ++j)
                for (int i = nx2; i < nx / 2 + nx2; ++i) {
                    int ii = (i - nx2) * 2, jj = (j - ny2) * 2;
                    Q(lev, i, j) = Q(lev - 1, ii, jj) + Q(lev - 1, ii + 1, jj);
                }
        }
        if (lev == ngrid - 1) {
            for (int i = 0; i < nx; ++i) { Q(lev, i, 0) = 0; Q(lev, i, ny) = 0; }
        } else {
            for (int i = 0; i < nx - 1; i += 2) {
                int ii = i / 2 + nx2;
                Q(lev, i, 0) = (0.75 * V(lev + 1, ii, ny2) + 0.25 * V(lev + 1, ii + 1, ny2)) * dx;
                Q(lev, i, ny) = (0.75 * V(lev + 1, ii, ny / 2 + ny2) + 0.25 * V(lev + 1, ii + 1, ny / 2 + ny2)) * dx;
                Q(lev, i + 1, 0) = (0.25 * V(lev + 1, ii, ny2) + 0.75 * V(lev + 1, ii + 1, ny2)) * dx;
                Q(lev, i + 1, ny) = (0.25 * V(lev + 1, ii, ny / 2 + ny2) + 0.75 * V(lev + 1, ii + 1, ny / 2 + ny2)) * dx;
            }
        }
        if (lev == ngrid - 1) {
            for (int j = 1; j < ny; ++j) {
                Q(lev, 0, j) = V(lev, 1, j) * 0.5 * dx;
                Q(lev, nx - 1, j) = V(lev, nx - 1, j) * 0.5 * dx;
            }
        } else {
            for (int j = 2; j < ny; j += 2) {
                int jj = j / 2 + ny2;
                Q(lev, 0, j) = (V(lev, 1, j) + V(lev + 1, nx2, jj)) * 0.5 * dx;
                Q(lev, nx - 1, j) = (V(lev, nx - 1, j) + V(lev + 1, nx / 2 + nx2, jj)) * 0.5 * dx;
            }
            for (int j = 1; j < ny; j += 2) {
                int jj = j / 2 + ny2;
                Q(lev, 0, j) = (0.5 * V(lev, 1, j) + 0.25 * V(lev + 1, nx2, jj) + 0.25 * V(lev + 1, nx2, jj + 1)) * dx;
                Q(lev, nx - 1, j) = (0.5 * V(lev, nx - 1, j) + 0.25 * V(lev + 1, nx / 2 + nx2, jj)
                    + 0.25 * V(lev + 1, nx / 2 + nx2, jj + 1)) * dx;
            }
        }
    }
#undef Q
#undef V
}

/* CrossProduct(Flux, Scalar), VectorOperations.cc:438-457 */
void og_cross_product(int nx, int ny, int ngrid, double dx, const double* q, const double* w, double* out) {
    const long n = SN(nx, ny) * ngrid;
    double* u = (double*)calloc(n, sizeof(double));
    double* v = (double*)calloc(n, sizeof(double));
    og_flux_to_xvel(nx, ny, ngrid, dx, q, u);
    for (long k = 0; k < n; ++k) { u[k] *= w[k]; u[k] *= -1; }
    og_flux_to_yvel(nx, ny, ngrid, dx, q, v);
    for (long k = 0; k < n; ++k) v[k] *= w[k];
    og_xvel_to_flux(nx, ny, ngrid, dx, v, out);   /* VelocityToFlux(v, u, cross) */
    og_yvel_to_flux(nx, ny, ngrid, dx, u, out);
    free(u); free(v);
}

/* CrossProduct(Flux, Flux) -> Scalar, VectorOperations.cc:465-486:
   u1 v2 - u2 v1 at the nodes of every level, then coarsify over the overlap. */
void og_cross_flux_flux(int nx, int ny, int ngrid, double dx, const double* q1, const double* q2, double* out) {
    const long n = SN(nx, ny) * ngrid;
    double* u = (double*)calloc(n, sizeof(double));
    double* v = (double*)calloc(n, sizeof(double));
    og_flux_to_xvel(nx, ny, ngrid, dx, q1, u);
    og_flux_to_yvel(nx, ny, ngrid, dx, q2, v);
    for (long k = 0; k < n; ++k) out[k] = u[k] * v[k];
    og_flux_to_xvel(nx, ny, ngrid, dx, q2, u);
    og_flux_to_yvel(nx, ny, ngrid, dx, q1, v);
    for (long k = 0; k < n; ++k) out[k] -= u[k] * v[k];
    og_coarsify(nx, ny, ngrid, out);
    free(u); free(v);
}

/* NonlinearIBSolver::N, IBSolver.cc:205-209 */
void og_nonlinear(int nx, int ny, int ngrid, double dx, const double* q, const double* w, double* n_out) {
    double* cross = (double*)calloc(NF(nx, ny) * ngrid, sizeof(double));
    og_cross_product(nx, ny, ngrid, dx, q, w, cross);
    og_curl_flux(nx, ny, ngrid, dx, cross, n_out);
    free(cross);
}

/* LinearizedIBSolver::N, IBSolver.cc:211-216: Curl(q0 x omega + q x omega0),
   the two cross products formed and summed flux by flux as the reference does */
void og_linear_N(int nx, int ny, int ngrid, double dx, const double* q0, const double* w0, const double* q,
                 const double* w, double* n_out) {
    const long nf = NF(nx, ny) * ngrid;
    double* v = (double*)calloc(nf, sizeof(double));
    double* v2 = (double*)calloc(nf, sizeof(double));
    og_cross_product(nx, ny, ngrid, dx, q0, w, v);
    og_cross_product(nx, ny, ngrid, dx, q, w0, v2);
    for (long k = 0; k < nf; ++k) v[k] += v2[k];
    og_curl_flux(nx, ny, ngrid, dx, v, n_out);
    free(v); free(v2);
}

/* AdjointIBSolver::N, IBSolver.cc:218-222:
   Laplacian(CrossProduct(q0, q)) - Curl(CrossProduct(q, omega0)) */
void og_adjoint_N(int nx, int ny, int ngrid, double dx, const double* q0, const double* w0, const double* q,
                  double* n_out) {
    const long n = SN(nx, ny) * ngrid, nf = NF(nx, ny) * ngrid;
    double* s = (double*)calloc(n, sizeof(double));
    double* v = (double*)calloc(nf, sizeof(double));
    double* c = (double*)calloc(n, sizeof(double));
    og_cross_flux_flux(nx, ny, ngrid, dx, q0, q, s);
    og_laplacian(nx, ny, ngrid, dx, s, n_out);
    og_cross_product(nx, ny, ngrid, dx, q, w0, v);
    og_curl_flux(nx, ny, ngrid, dx, v, c);
    for (long k = 0; k < n; ++k) n_out[k] -= c[k];
    free(s); free(v); free(c);
}

/* ======================================================================= */
/* Regulariser (Regularizer.cc)                                             */
/* ======================================================================= */
static double delta_fn(double r) {   /* Regularizer.cc:37-49 (Roma-Peskin-Berger) */
    if (r > 1.5) return 0;
    if (r <= 0.5) return (1 + sqrt(1 - 3 * r * r)) / 3;
    return (5 - 3 * r - sqrt(1 - 3 * (1 - r) * (1 - r))) / 6;
}

/* Regularizer::update, Regularizer.cc:54-90.  The reference scans every
   level-0 flux; this scans a window of +-3 cells around each point, in the
   same ascending flux-index order and with the same floating-point test, so
   the list (order and members) is identical. */
int og_reg_build(int nx, int ny, double h, double xoff, double yoff, int npts,
                 const double* px, const double* py, int* bidx, int* fidx, double* w) {
    int na = 0;
    for (int dir = 0; dir <= 1; ++dir) {
        for (int p = 0; p < npts; ++p) {
            const int ci = (int)floor((px[p] - xoff) / h);
            const int cj = (int)floor((py[p] - yoff) / h);
            const int imax = (dir == 0) ? nx : nx - 1;
            const int jmax = (dir == 0) ? ny - 1 : ny;
            int i0 = ci - 3, i1 = ci + 3, j0 = cj - 3, j1 = cj + 3;
            if (i0 < 0) i0 = 0;
            if (j0 < 0) j0 = 0;
            if (i1 > imax) i1 = imax;
            if (j1 > jmax) j1 = jmax;
            for (int i = i0; i <= i1; ++i) {
                for (int j = j0; j <= j1; ++j) {
                    double fx, fy;
                    long ind;
                    if (dir == 0) {   /* X flux: x at edge, y at centre (Flux.h:96-150) */
                        fx = xoff + i * h;
                        fy = yoff + (j + 0.5) * h;
                        ind = (long)i * ny + j;
                    } else {          /* Y flux: x at centre, y at edge */
                        fx = xoff + (i + 0.5) * h;
                        fy = yoff + j * h;
                        ind = NFX(nx, ny) + (long)i * (ny + 1) + j;
                    }
                    const double ddx = fabs(fx - px[p]) / h;
                    const double ddy = fabs(fy - py[p]) / h;
                    if (ddx < 1.5 && ddy < 1.5) {
                        w[na] = delta_fn(ddx) * delta_fn(ddy);
                        fidx[na] = (int)ind;
                        bidx[na] = dir * npts + p;
                        ++na;
                    }
                }
            }
        }
    }
    return na;
}

void og_reg_toflux(int nx, int ny, int ngrid, double dx, int na, const int* bidx, const int* fidx,
                   const double* w, const double* f, double* q) {
    const long n = NF(nx, ny) * ngrid;
    memset(q, 0, sizeof(double) * n);
    for (int a = 0; a < na; ++a) q[fidx[a]] += w[a] * f[bidx[a]];
    for (long k = 0; k < n; ++k) q[k] *= dx;
}

void og_reg_toboundary(int npts, double dx, int na, const int* bidx, const int* fidx,
                       const double* w, const double* q, double* f) {
    memset(f, 0, sizeof(double) * 2 * npts);
    for (int a = 0; a < na; ++a) f[bidx[a]] += w[a] * q[fidx[a]];
    const double inv = 1.0 / dx;   /* Array.h:335-338 divides by multiplying with 1/a */
    for (int k = 0; k < 2 * npts; ++k) f[k] *= inv;
}

/* ======================================================================= */
/* Geometry (Geometry.cc, RigidBody.cc, PitchPlunge.h, TangentSE2.h)         */
/* ======================================================================= */
/* motion codes: the "motion" types of RigidBody.cc:351-492 */
enum {
    OG_NONE = 0, OG_FIXED, OG_PITCHPLUNGE, OG_FIXEDVEL, OG_SIGMOIDALSTEP, OG_LAGSTEP1, OG_LAGSTEP2, OG_ELDREDGE,
    OG_ELDREDGE1, OG_ELDREDGE2, OG_ELDREDGECOMBINED2, OG_MOTIONFILE, OG_MOTIONFILEPERIODIC
};

typedef struct {
    int start, npts;
    double xc, yc;
    int motion;
    double mp[14];       /* parameters in the constructor's order, then derived constants */
    double* table;       /* motion files: nrows x (t, x, y, theta, dx, dy, dtheta) */
    int nrows;
    double acc[3], oldtime;   /* state of FixedVelocity, Eldredge2, EldredgeCombined2 */
} og_body;

struct og_sim {
    int nx, ny, ngrid;
    double dx, xoff, yoff;
    double Re, dt, alpha_rad;
    int scheme, nsub;
    double an[4], bn[4], cn[4];
    double tol;
    int force_cg, use_cg, stationary;
    /* geometry */
    int nb, nbody;
    og_body bodies[64];
    double *refx, *refy, *px, *py, *pu, *pv;
    /* state */
    double time; int timestep;
    double *omega, *q, *f, *q0, *nprev;
    int old_saved;
    /* regulariser */
    int na; int *abi, *afi; double* aw;
    /* Cholesky per substep */
    double *chol_diag[4], *chol_low[4];
    int cg_iters[4];
    /* unsteady base flow (ibpm -ubf): the motion moved from the only body */
    int ubf;
    og_body bf;
    double bf_alpha, bf_mag, bf_u, bf_v, bf_w;
    double force_alpha;
    /* models of ibpm.cc:185-235: 0 nonlinear, 1 linear, 2 adjoint,
       3 linear periodic, 4 SFD */
    int model, period;
    double *bq[64], *bw[64];   /* base states (LinearizedPeriodic: one per phase) */
    double chi, delta;         /* SFD gain and filter width (SFDSolver, IBSolver.h:189-222) */
    double *what, *sfd_prev;   /* filtered vorticity and the filter's previous right-hand side */
    int what_saved, sfd_saved;
};

static char g_err[256];
const char* og_sim_error(void) { return g_err; }

static void add_point(og_sim* s, og_body* b, double x, double y) {
    s->refx = (double*)realloc(s->refx, sizeof(double) * (s->nb + 1));
    s->refy = (double*)realloc(s->refy, sizeof(double) * (s->nb + 1));
    s->refx[s->nb] = x; s->refy[s->nb] = y;
    s->nb++; b->npts++;
}

/* MotionFile / MotionFilePeriodic constructors (MotionFile.h:57-81,
   MotionFilePeriodic.h:56-81): "filename" or "filename period"; the file holds
   n, then n rows t x y theta dx dy dtheta with t non-decreasing. */
static int load_motion_file(og_body* B, const char* args, int periodic) {
    char fname[512] = {0};
    int used = 0;
    if (sscanf(args, "%511s%n", fname, &used) != 1) { snprintf(g_err, sizeof g_err, "geom: motion file name missing"); return -1; }
    if (periodic && sscanf(args + used, "%lf", &B->mp[0]) != 1) {
        snprintf(g_err, sizeof g_err, "geom: motionfileperiodic needs a period"); return -1;
    }
    FILE* fp = fopen(fname, "r");
    if (!fp) { snprintf(g_err, sizeof g_err, "geom: cannot open %s", fname); return -1; }
    int n = 0;
    if (fscanf(fp, "%d", &n) != 1 || n < 0) { fclose(fp); snprintf(g_err, sizeof g_err, "geom: bad motion file"); return -1; }
    B->table = (double*)malloc(sizeof(double) * (7 * (size_t)n + 1));
    B->nrows = n;
    for (int r = 0; r < 7 * n; ++r)
        if (fscanf(fp, "%lf", &B->table[r]) != 1) { fclose(fp); snprintf(g_err, sizeof g_err, "geom: bad motion file"); return -1; }
    fclose(fp);
    for (int r = 1; r < n; ++r)
        if (B->table[7 * r] < B->table[7 * (r - 1)]) { snprintf(g_err, sizeof g_err, "geom: motion file times decrease"); return -1; }
    return 0;
}

static int parse_geom(og_sim* s, const char* text) {
    const char* p = text;
    char line[1024];
    og_body* cur = NULL;
    while (*p) {
        int k = 0;
        while (*p && *p != '\n' && k < 1023) line[k++] = *p++;
        line[k] = 0;
        if (*p == '\n') ++p;
        char cmd[64] = {0};
        if (sscanf(line, "%63s", cmd) != 1) continue;
        for (char* c = cmd; *c; ++c) *c = (char)tolower(*c);
        if (cmd[0] == '#') continue;
        const char* rest = strstr(line, cmd[0] ? line : line);
        /* move past the command word */
        rest = line; while (*rest && isspace((unsigned char)*rest)) ++rest;
        while (*rest && !isspace((unsigned char)*rest)) ++rest;
        if (!cur) {
            if (!strcmp(cmd, "body")) {
                cur = &s->bodies[s->nbody++];
                memset(cur, 0, sizeof(*cur));
                cur->start = s->nb;
            } else if (!strcmp(cmd, "end")) {
                break;
            } else { snprintf(g_err, sizeof g_err, "geom: unexpected '%s'", cmd); return -1; }
            continue;
        }
        double a[12]; int n;
        if (!strcmp(cmd, "end")) { cur = NULL; continue; }
        else if (!strcmp(cmd, "center")) {
            if (sscanf(rest, "%lf %lf", &a[0], &a[1]) != 2) return -1;
            cur->xc = a[0]; cur->yc = a[1];
        } else if (!strcmp(cmd, "circle")) {   /* RigidBody.cc:92-101 */
            if (sscanf(rest, "%lf %lf %lf %lf", &a[0], &a[1], &a[2], &a[3]) != 4) return -1;
            double dTheta = a[3] / a[2];
            double twopi = 8. * atan(1.);
            int np = (int)floor(twopi / dTheta + 1 + 0.5);
            double dt = twopi / np;
            for (int i = 0; i < np; ++i) {
                double si, ci;
                sin_cos(i * dt, &si, &ci);
                add_point(s, cur, a[0] + a[2] * ci, a[1] + a[2] * si);
            }
        } else if (!strcmp(cmd, "circle_n")) { /* RigidBody.cc:103-118 */
            if (sscanf(rest, "%lf %lf %lf %d", &a[0], &a[1], &a[2], &n) != 4) return -1;
            double twopi = 8. * atan(1.);
            double dt = twopi / n;
            for (int i = 0; i < n; ++i) {
                double si, ci;
                sin_cos(i * dt, &si, &ci);
                add_point(s, cur, a[0] + a[2] * ci, a[1] + a[2] * si);
            }
        } else if (!strcmp(cmd, "line") || !strcmp(cmd, "line_n")) { /* RigidBody.cc:120-146 */
            int np;
            if (!strcmp(cmd, "line")) {
                if (sscanf(rest, "%lf %lf %lf %lf %lf", &a[0], &a[1], &a[2], &a[3], &a[4]) != 5) return -1;
                double len = sqrt((a[2] - a[0]) * (a[2] - a[0]) + (a[3] - a[1]) * (a[3] - a[1]));
                np = (int)floor(len / a[4] + 1 + 0.5);
            } else {
                if (sscanf(rest, "%lf %lf %lf %lf %d", &a[0], &a[1], &a[2], &a[3], &np) != 5) return -1;
            }
            double ddx = (a[2] - a[0]) / (np - 1), ddy = (a[3] - a[1]) / (np - 1);
            for (int i = 0; i < np; ++i) add_point(s, cur, a[0] + i * ddx, a[1] + i * ddy);
        } else if (!strcmp(cmd, "line_aoa")) { /* RigidBody.cc:148-171 */
            int np;
            if (sscanf(rest, "%lf %lf %lf %lf %d", &a[0], &a[1], &a[2], &a[3], &np) != 5) return -1;
            double pi = 3.141592653589793238462643383279502884197169399375;
            double al = a[3] / 180 * pi, ca, sa, del = a[0] / (np - 1);
            sin_cos(al, &sa, &ca);
            for (int i = 0; i < np; ++i) {
                double x0r = i * del - a[1], y0r = 0. - a[2];
                add_point(s, cur, x0r * ca - y0r * sa + a[1], x0r * sa + y0r * ca + a[2]);
            }
            cur->xc = a[1]; cur->yc = a[2];
        } else if (!strcmp(cmd, "point")) {
            if (sscanf(rest, "%lf %lf", &a[0], &a[1]) != 2) return -1;
            add_point(s, cur, a[0], a[1]);
        } else if (!strcmp(cmd, "name")) {
        } else if (!strcmp(cmd, "motion")) {
            char mt[64] = {0};
            if (sscanf(rest, "%63s", mt) != 1) return -1;
            for (char* c = mt; *c; ++c) *c = (char)tolower(*c);
            const char* r2 = rest;   /* past the motion word */
            while (*r2 && isspace((unsigned char)*r2)) ++r2;
            while (*r2 && !isspace((unsigned char)*r2)) ++r2;
            int want = -1, code = OG_NONE;
            if (!strcmp(mt, "fixed")) { want = 3; code = OG_FIXED; }
            else if (!strcmp(mt, "fixedvel")) { want = 3; code = OG_FIXEDVEL; }
            else if (!strcmp(mt, "pitchplunge")) { want = 6; code = OG_PITCHPLUNGE; }
            else if (!strcmp(mt, "sigmoidalstep")) { want = 3; code = OG_SIGMOIDALSTEP; }
            else if (!strcmp(mt, "lagstep1")) { want = 4; code = OG_LAGSTEP1; }
            else if (!strcmp(mt, "lagstep2")) { want = 4; code = OG_LAGSTEP2; }
            else if (!strcmp(mt, "eldredge")) { want = 6; code = OG_ELDREDGE; }
            else if (!strcmp(mt, "eldredge1")) { want = 6; code = OG_ELDREDGE1; }
            else if (!strcmp(mt, "eldredge2")) { want = 6; code = OG_ELDREDGE2; }
            else if (!strcmp(mt, "eldredgecombined2")) { want = 12; code = OG_ELDREDGECOMBINED2; }
            else if (!strcmp(mt, "motionfile")) { want = 0; code = OG_MOTIONFILE; }
            else if (!strcmp(mt, "motionfileperiodic")) { want = 0; code = OG_MOTIONFILEPERIODIC; }
            if (want < 0) { snprintf(g_err, sizeof g_err, "geom: unsupported motion %s", mt); return -1; }
            if (want > 0) {
                const char* q = r2;
                for (int v = 0; v < want; ++v) {
                    int used = 0;
                    if (sscanf(q, "%lf%n", &cur->mp[v], &used) != 1) {
                        snprintf(g_err, sizeof g_err, "geom: motion %s needs %d values", mt, want); return -1;
                    }
                    q += used;
                }
            } else if (load_motion_file(cur, r2, code == OG_MOTIONFILEPERIODIC) != 0) return -1;
            cur->motion = code;
            double* m = cur->mp;
            if (code == OG_PITCHPLUNGE) {   /* PitchPlunge.h:45-47: frequencies in radians */
                double twopi = 8. * atan(1.);
                m[1] *= twopi; m[4] *= twopi;
            } else if (code == OG_ELDREDGE || code == OG_ELDREDGE1 || code == OG_ELDREDGE2) {
                /* maxG of EldredgeManeuver.h:42-43 (same in Eldredge1.h, Eldredge2.h) */
                double tt = (m[3] + m[4]) / 2.;
                m[6] = log((cosh(m[1] * (tt - m[2])) * cosh(m[1] * (tt - m[5]))) /
                           (cosh(m[1] * (tt - m[3])) * cosh(m[1] * (tt - m[4]))));
            } else if (code == OG_ELDREDGECOMBINED2) {
                /* EldredgeCombined2.h:53-59; its maxGb takes cosh(a*(ttb-b4)) with a, not b */
                double tta = (m[3] + m[4]) / 2., ttb = (m[9] + m[10]) / 2.;
                m[12] = log((cosh(m[1] * (tta - m[2])) * cosh(m[1] * (tta - m[5]))) /
                            (cosh(m[1] * (tta - m[3])) * cosh(m[1] * (tta - m[4]))));
                m[13] = log((cosh(m[7] * (ttb - m[8])) * cosh(m[1] * (ttb - m[11]))) /
                            (cosh(m[7] * (ttb - m[9])) * cosh(m[7] * (ttb - m[10]))));
            }
        } else { snprintf(g_err, sizeof g_err, "geom: unsupported command %s", cmd); return -1; }
    }
    return 0;
}

/* Motion::getTransformation(t) -> (x, y, theta, xdot, ydot, thetadot) for each
   motion type, in the arithmetic order of the reference's headers. */
static void body_transform(og_body* B, double t, double* o) {
    const double* m = B->mp;
    const double pi = 4. * atan(1.);
    o[0] = o[1] = o[2] = o[3] = o[4] = o[5] = 0;
    switch (B->motion) {
    case OG_FIXED: o[0] = m[0]; o[1] = m[1]; o[2] = m[2]; break;    /* FixedPosition.h */
    case OG_PITCHPLUNGE: {                                            /* PitchPlunge.h:52-58 */
        double sp, cp, sq, cq;
        sin_cos(m[4] * t + m[5], &sp, &cp);
        sin_cos(m[1] * t + m[2], &sq, &cq);
        o[1] = m[3] * sp;
        o[4] = m[3] * m[4] * cp;
        o[2] = m[0] * sq;
        o[5] = m[0] * m[1] * cq;
        break;
    }
    case OG_FIXEDVEL: {                                               /* FixedVelocity.h:42-50 */
        double dtt = t - B->oldtime;
        if (dtt > 1) dtt = 0.;
        for (int k = 0; k < 3; ++k) B->acc[k] = B->acc[k] + m[k] * dtt;
        B->oldtime = t;
        o[0] = B->acc[0]; o[1] = B->acc[1]; o[2] = B->acc[2];
        o[3] = m[0]; o[4] = m[1]; o[5] = m[2];
        break;
    }
    case OG_SIGMOIDALSTEP: {                                          /* SigmoidalStep.h:41-52 */
        double td = t - m[2];
        double arg = ((12. * td / m[1]) - 6.);
        double coeff = 12. / (m[1] * sqrt(pi));
        o[2] = (1. / 2) * m[0] * (pi / 180) * (1. + erf(arg));
        o[5] = m[0] * (pi / 180) * coeff * exp(-arg * arg);
        break;
    }
    case OG_LAGSTEP1: case OG_LAGSTEP2: {                             /* LagStep1.h, LagStep2.h:41-63 */
        double amp = m[0] * (pi / 180), pw = m[1], tau = m[2], td = t - m[3];
        double th = 0., thd = 0.;
        if (td > 0 && B->motion == OG_LAGSTEP1) {
            if (td <= pw) { th = td + tau * exp(-td / tau) - tau; thd = 1 - exp(-td / tau); }
            else {
                th = pw + tau * exp(-td / tau) - tau * exp(-(td - pw) / tau);
                thd = -1. * exp(-td / tau) + exp(-(td - pw) / tau);
            }
        } else if (td > 0) {
            if (td <= pw) {
                th = td + (2 * tau + td) * exp(-td / tau) - 2 * tau;
                thd = 1. - (1. + td / tau) * exp(-td / tau);
            } else {
                th = pw + (2 * tau + td) * exp(-td / tau) + (pw - 2 * tau) * exp(-(td - pw) / tau) -
                     td * exp(-(td - pw) / tau);
                thd = -(1. + td / tau) * exp(-td / tau) + (1. + (td - pw) / tau) * exp(-(td - pw) / tau);
            }
        }
        o[2] = th * amp / pw; o[5] = thd * amp / pw;
        break;
    }
    case OG_ELDREDGE: case OG_ELDREDGE1: case OG_ELDREDGE2: {         /* EldredgeManeuver.h, Eldredge1.h, Eldredge2.h */
        double a = m[1], d1 = t - m[2], d2 = t - m[3], d3 = t - m[4], d4 = t - m[5];
        double G = log(((cosh(a * d1) * cosh(a * d4)) / (cosh(a * d2) * cosh(a * d3))));
        double dG = a * (tanh(a * d1) - tanh(a * d2) - tanh(a * d3) + tanh(a * d4));
        if (B->motion == OG_ELDREDGE) {
            double amp = m[0] * (pi / 180);
            o[2] = G * amp / m[6]; o[5] = dG * amp / m[6];
        } else if (B->motion == OG_ELDREDGE1) {
            o[1] = G * m[0] / m[6]; o[4] = dG * m[0] / m[6];
        } else {   /* running integral of G: Eldredge2.h:62-64 */
            B->acc[0] = B->acc[0] + G * (t - B->oldtime);
            B->oldtime = t;
            o[1] = B->acc[0] * m[0] / m[6]; o[4] = G * m[0] / m[6];
        }
        break;
    }
    case OG_ELDREDGECOMBINED2: {                                      /* EldredgeCombined2.h:64-88 */
        double amp = m[6] * (pi / 180), a = m[1], b = m[7];
        double Ga = log(((cosh(a * (t - m[2])) * cosh(a * (t - m[5]))) / (cosh(a * (t - m[3])) * cosh(a * (t - m[4])))));
        double Gb = log(((cosh(b * (t - m[8])) * cosh(b * (t - m[11]))) / (cosh(b * (t - m[9])) * cosh(b * (t - m[10])))));
        double dGb = b * (tanh(b * (t - m[8])) - tanh(b * (t - m[9])) - tanh(b * (t - m[10])) + tanh(b * (t - m[11])));
        B->acc[0] = B->acc[0] + Ga * (t - B->oldtime);
        B->oldtime = t;
        o[1] = B->acc[0] * m[0] / m[12]; o[2] = Gb * amp / m[13];
        o[4] = Ga * m[0] / m[12]; o[5] = dGb * amp / m[13];
        break;
    }
    case OG_MOTIONFILE: case OG_MOTIONFILEPERIODIC: {                 /* MotionFile.h:90-116, MotionFilePeriodic.h:91-117 */
        double tt = B->motion == OG_MOTIONFILEPERIODIC ? fmod(t, m[0]) : t;
        const double* d = B->table;
        int idx = -1;
        for (int i = 0; i + 1 < B->nrows; ++i)
            if (tt >= d[7 * i] && tt < d[7 * (i + 1)]) idx = i;
        if (idx >= 0) {
            const double* p = d + 7 * idx;
            double al = (tt - p[0]) / (p[7] - p[0]);
            for (int k = 0; k < 6; ++k) o[k] = al * p[8 + k] + (1 - al) * p[1 + k];
        }
        break;
    }
    default: break;
    }
}

/* Geometry::moveBodies -> RigidBody::moveBody (RigidBody.cc:586-609) with
   TangentSE2::mapPosition/mapVelocity (TangentSE2.h:83-107) */
static void move_bodies(og_sim* s, double t) {
    for (int b = 0; b < s->nbody; ++b) {
        og_body* B = &s->bodies[b];
        if (B->motion == OG_NONE) continue;
        double o[6];
        body_transform(B, t, o);
        const double x = o[0], y = o[1], th = o[2], xd = o[3], yd = o[4], thd = o[5];
        double ct, st;
        sin_cos(th, &st, &ct);
        for (int k = B->start; k < B->start + B->npts; ++k) {
            double a = s->refx[k] - B->xc, c = s->refy[k] - B->yc;
            s->px[k] = (x + a * ct - c * st) + B->xc;
            s->py[k] = (y + a * st + c * ct) + B->yc;
            s->pu[k] = -a * st * thd - c * ct * thd + xd;
            s->pv[k] = a * ct * thd - c * st * thd + yd;
        }
    }
}

/* ======================================================================= */
/* Model, projection, time stepping                                         */
/* ======================================================================= */
static long s_n(const og_sim* s) { return SN(s->nx, s->ny) * s->ngrid; }
static long f_n(const og_sim* s) { return NF(s->nx, s->ny) * s->ngrid; }

static void reg_update(og_sim* s) {   /* NavierStokesModel::init / updateOperators */
    s->na = og_reg_build(s->nx, s->ny, s->dx, s->xoff, s->yoff, s->nb, s->px, s->py, s->abi, s->afi, s->aw);
}

/* NavierStokesModel::B, NavierStokesModel.cc:101-105 */
static void model_B(og_sim* s, const double* f, double* w) {
    double* q = (double*)malloc(sizeof(double) * f_n(s));
    og_reg_toflux(s->nx, s->ny, s->ngrid, s->dx, s->na, s->abi, s->afi, s->aw, f, q);
    og_curl_flux(s->nx, s->ny, s->ngrid, s->dx, q, w);
    free(q);
}

/* vorticityToStreamfunction + Curl, NavierStokesModel.cc:107-140 */
static void flux_without_base(og_sim* s, const double* w, double* q) {
    const long n = s_n(s);
    double* psi = (double*)malloc(sizeof(double) * n);
    for (long k = 0; k < n; ++k) psi[k] = -1. * w[k];
    og_coarsify(s->nx, s->ny, s->ngrid, psi);
    og_elliptic_solve(0, s->nx, s->ny, s->ngrid, s->dx, 0.0, psi, psi);
    og_curl_scalar(s->nx, s->ny, s->ngrid, psi, q);
    free(psi);
}

void og_sim_C(og_sim* s, const double* w, double* f) {
    double* q = (double*)malloc(sizeof(double) * f_n(s));
    flux_without_base(s, w, q);
    og_reg_toboundary(s->nb, s->dx, s->na, s->abi, s->afi, s->aw, q, f);
    free(q);
}

double og_sim_beta(const og_sim* s, int i) { return (s->an[i] + s->bn[i]) * s->dt; }

void og_sim_Ainv(og_sim* s, int i, const double* a, double* w) {   /* ProjectionSolver.cc:81-83 */
    const double beta = og_sim_beta(s, i);
    og_elliptic_solve(1, s->nx, s->ny, s->ngrid, s->dx, -beta / 2. * (1. / s->Re), a, w);
}

void og_sim_B(og_sim* s, int i, const double* f, double* w) {      /* ProjectionSolver.cc:85-88 */
    model_B(s, f, w);
    const double beta = og_sim_beta(s, i);
    for (long k = 0; k < s_n(s); ++k) w[k] *= beta;
}

void og_sim_M(og_sim* s, int i, const double* f, double* y) {      /* ProjectionSolver.cc:96-101 */
    double* u = (double*)malloc(sizeof(double) * s_n(s));
    og_sim_B(s, i, f, u);
    og_sim_Ainv(s, i, u, u);
    og_sim_C(s, u, y);
    free(u);
}

/* CholeskySolver::computeMatrixM + computeFactorization, CholeskySolver.cc:63-109 */
static void chol_build(og_sim* s, int i) {
    const int n = 2 * s->nb;
    double* M = (double*)malloc(sizeof(double) * (size_t)n * n);
    double* e = (double*)calloc(n, sizeof(double));
    double* x = (double*)malloc(sizeof(double) * n);
    for (int j = 0; j < n; ++j) {
        memset(e, 0, sizeof(double) * n);
        e[j] = 1;
        og_sim_M(s, i, e, x);
        for (int r = 0; r < n; ++r) M[(size_t)r * n + j] = x[r];
    }
    double* L = (double*)calloc((size_t)n * n, sizeof(double));
    double* d = (double*)calloc(n, sizeof(double));
    /* L(j,i) = (M(i,j) - sum_{k=i-1..0} L(i,k) L(j,k)) / d(i) for j > i: the
       reference reads row i of M's upper part (CholeskySolver.cc:93-107). */
    for (int r = 0; r < n; ++r) {
        for (int c = r; c < n; ++c) {
            double sum = M[(size_t)r * n + c];
            for (int k = r - 1; k >= 0; --k) sum -= L[(size_t)r * n + k] * L[(size_t)c * n + k];
            if (c == r) d[r] = sqrt(sum);
            else L[(size_t)c * n + r] = sum / d[r];
        }
    }
    s->chol_diag[i] = d;
    s->chol_low[i] = L;
    free(M); free(e); free(x);
}

/* CholeskySolver::Minv, CholeskySolver.cc:200-224 */
static void chol_minv(og_sim* s, int i, const double* b, double* x) {
    const int n = 2 * s->nb;
    const double* L = s->chol_low[i];
    const double* d = s->chol_diag[i];
    for (int r = 0; r < n; ++r) {
        double sum = b[r];
        for (int k = r - 1; k >= 0; --k) sum -= L[(size_t)r * n + k] * x[k];
        x[r] = sum / d[r];
    }
    for (int r = n - 1; r >= 0; --r) {
        double sum = x[r];
        for (int k = r + 1; k < n; ++k) sum -= L[(size_t)k * n + r] * x[k];
        x[r] = sum / d[r];
    }
}

static double bv_dot(int n, const double* a, const double* b) {   /* BoundaryVector.h:240-252 */
    double s = 0;
    for (int k = 0; k < n; ++k) s += a[k] * b[k];
    return s;
}

/* ConjugateGradientSolver::Minv, ConjugateGradientSolver.cc:35-65; the loop
   counts iterations (the reference never increments its counter) and stops at
   100000 as a guard against a non-converging run. */
static int cg_minv(og_sim* s, int i, const double* b, double* x) {
    const int n = 2 * s->nb;
    double* r = (double*)malloc(sizeof(double) * n);
    double* q = (double*)malloc(sizeof(double) * n);
    double* d = (double*)malloc(sizeof(double) * n);
    og_sim_M(s, i, x, q);
    for (int k = 0; k < n; ++k) { r[k] = b[k]; r[k] -= q[k]; d[k] = r[k]; }
    double delta = bv_dot(n, r, r);
    const double tol2 = s->tol * s->tol;
    int it = 0;
    while (delta > tol2 && it < 100000) {
        og_sim_M(s, i, d, q);
        double alpha = delta / bv_dot(n, d, q);
        for (int k = 0; k < n; ++k) x[k] += alpha * d[k];
        for (int k = 0; k < n; ++k) r[k] -= alpha * q[k];
        double dold = delta;
        delta = bv_dot(n, r, r);
        double beta = delta / dold;
        for (int k = 0; k < n; ++k) d[k] = r[k] + beta * d[k];
        ++it;
    }
    free(r); free(q); free(d);
    return it;
}

/* ProjectionSolver::solve, ProjectionSolver.cc:57-79 */
void og_sim_projection_solve(og_sim* s, int i, const double* a, const double* b, double* omega, double* f) {
    const long n = s_n(s);
    const int nbv = 2 * s->nb;
    double* ws = (double*)malloc(sizeof(double) * n);
    og_sim_Ainv(s, i, a, ws);
    double* rhs = (double*)malloc(sizeof(double) * (nbv ? nbv : 1));
    og_sim_C(s, ws, rhs);
    for (int k = 0; k < nbv; ++k) rhs[k] -= b[k];
    if (s->use_cg) s->cg_iters[i] = cg_minv(s, i, rhs, f);
    else { chol_minv(s, i, rhs, f); s->cg_iters[i] = -1; }
    double* c = (double*)malloc(sizeof(double) * n);
    og_sim_B(s, i, f, c);
    og_sim_Ainv(s, i, c, c);
    for (long k = 0; k < n; ++k) omega[k] = ws[k] - c[k];
    free(ws); free(rhs); free(c);
}

/* NavierStokesModel::refreshState / computeFlux, NavierStokesModel.cc:121-129 */
static void refresh_state(og_sim* s) {
    flux_without_base(s, s->omega, s->q);
    for (long k = 0; k < f_n(s); ++k) s->q[k] += s->q0[k];
}

/* BaseFlow::setAlphaMag (BaseFlow.cc:111-125): angle and magnitude of the
   free stream relative to the moving body; free stream 1 at alpha_rad */
static void bf_set_alpha_mag(og_sim* s, double t) {
    double o[6];
    body_transform(&s->bf, t, o);
    double ca, sa;
    sin_cos(s->alpha_rad, &sa, &ca);
    const double xdot_bf = 1. * ca, ydot_bf = 1. * sa;
    const double xdot_t = o[3] - xdot_bf, ydot_t = o[4] - ydot_bf;
    const double gamma = atan2(ydot_t, -1. * xdot_t);
    s->bf_alpha = -1. * o[2] - gamma;
    s->bf_mag = sqrt(xdot_t * xdot_t + ydot_t * ydot_t);
}

/* BaseFlow::moveFlow (BaseFlow.cc:127-150) and Flux::setFlow (Flux.cc:101-129):
   q0 = free stream of magnitude mag at angle alpha plus the rotation -thetadot
   about the body centre, at the edge coordinates of every level
   (Grid.cc:104-139 with zero shift) */
static void bf_move(og_sim* s, double t) {
    double o[6];
    body_transform(&s->bf, t, o);
    bf_set_alpha_mag(s, t);
    double ca, sa;
    sin_cos(s->bf_alpha, &sa, &ca);
    s->bf_u = s->bf_mag * ca;
    s->bf_v = s->bf_mag * sa;
    s->bf_w = -1. * o[5];
    const int nx = s->nx, ny = s->ny;
    for (int lev = 0; lev < s->ngrid; ++lev) {
        const double d = s->dx * (1 << lev);
        const double x0 = s->xoff + 0.5 * (0. - 1) * ((1 << lev) - 1) * (nx * s->dx);
        const double y0 = s->yoff + 0.5 * (0. - 1) * ((1 << lev) - 1) * (ny * s->dx);
        double* q = s->q0 + lev * NF(nx, ny);
        for (int i = 0; i <= nx; ++i)
            for (int j = 0; j < ny; ++j) {
                const double yd = (y0 + (j + 0.5) * d) - s->bf.yc;
                q[(long)i * ny + j] = (s->bf_u - s->bf_w * yd) * d;
            }
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j <= ny; ++j) {
                const double xd = (x0 + (i + 0.5) * d) - s->bf.xc;
                q[NFX(nx, ny) + (long)i * (ny + 1) + j] = (s->bf_v + s->bf_w * xd) * d;
            }
    }
}

/* getConstraints, NavierStokesModel.cc:86-91,142-146 */
static void get_constraints(og_sim* s, double* b) {
    double* b0 = (double*)malloc(sizeof(double) * 2 * s->nb);
    og_reg_toboundary(s->nb, s->dx, s->na, s->abi, s->afi, s->aw, s->q0, b0);
    for (int k = 0; k < s->nb; ++k) { b[k] = s->pu[k]; b[s->nb + k] = s->pv[k]; }
    for (int k = 0; k < 2 * s->nb; ++k) b[k] -= b0[k];
    free(b0);
}

og_sim* og_sim_create(int nx, int ny, int ngrid, double length, double xoff, double yoff,
                      double Re, double dt, int scheme, double alpha_deg,
                      const char* geom_text, double cg_tol, int force_cg) {
    og_sim* s = (og_sim*)calloc(1, sizeof(og_sim));
    s->nx = nx; s->ny = ny; s->ngrid = ngrid;
    s->dx = length / nx; s->xoff = xoff; s->yoff = yoff;
    s->Re = Re; s->dt = dt; s->scheme = scheme;
    double pi = 4. * atan(1.);
    s->alpha_rad = alpha_deg * pi / 180.;
    s->force_alpha = s->alpha_rad;
    s->tol = cg_tol; s->force_cg = force_cg;
    switch (scheme) {   /* Scheme.h:14-56 */
    case 0: s->nsub = 1; s->an[0] = 1.; s->bn[0] = 0.; s->cn[0] = 1.; break;
    case 1: s->nsub = 1; s->an[0] = 3. / 2.; s->bn[0] = -1. / 2.; s->cn[0] = 1.; break;
    case 3:
        s->nsub = 4;
        s->an[0] = 8. / 17.; s->an[1] = 17. / 60.; s->an[2] = 5. / 12.; s->an[3] = 3. / 4.;
        s->bn[0] = 0.; s->bn[1] = -15. / 68.; s->bn[2] = -17. / 60.; s->bn[3] = -5. / 12.;
        s->cn[0] = 8. / 17.; s->cn[1] = 8. / 15.; s->cn[2] = 2. / 3.; s->cn[3] = 1.;
        break;
    default:
        s->nsub = 3;
        s->an[0] = 8. / 15.; s->an[1] = 5. / 12.; s->an[2] = 3. / 4.;
        s->bn[0] = 0.; s->bn[1] = -17. / 60.; s->bn[2] = -5. / 12.;
        s->cn[0] = 8. / 15.; s->cn[1] = 2. / 3.; s->cn[2] = 1.;
    }
    if (parse_geom(s, geom_text ? geom_text : "") != 0) { og_sim_destroy(s); return NULL; }
    s->stationary = 1;
    for (int b = 0; b < s->nbody; ++b)   /* only FixedPosition is stationary (Motion.h:28, FixedPosition.h:39) */
        if (s->bodies[b].motion != OG_NONE && s->bodies[b].motion != OG_FIXED) s->stationary = 0;
    s->use_cg = (!s->stationary) || force_cg;
    const int nb = s->nb;
    s->px = (double*)calloc(nb + 1, sizeof(double)); s->py = (double*)calloc(nb + 1, sizeof(double));
    s->pu = (double*)calloc(nb + 1, sizeof(double)); s->pv = (double*)calloc(nb + 1, sizeof(double));
    for (int k = 0; k < nb; ++k) { s->px[k] = s->refx[k]; s->py[k] = s->refy[k]; }
    s->abi = (int*)malloc(sizeof(int) * (36 * nb + 1));
    s->afi = (int*)malloc(sizeof(int) * (36 * nb + 1));
    s->aw = (double*)malloc(sizeof(double) * (36 * nb + 1));
    s->omega = (double*)calloc(s_n(s), sizeof(double));
    s->nprev = (double*)calloc(s_n(s), sizeof(double));
    s->q = (double*)calloc(f_n(s), sizeof(double));
    s->q0 = (double*)calloc(f_n(s), sizeof(double));
    s->f = (double*)calloc(2 * nb + 1, sizeof(double));
    /* BaseFlow(grid, 1, alpha) -> Flux::UniformFlow, Flux.cc:79-99 */
    double ca, sa;
    sin_cos(s->alpha_rad, &sa, &ca);
    const double u = 1. * ca, v = 1. * sa;
    for (int lev = 0; lev < ngrid; ++lev) {
        const double d = s->dx * (1 << lev);
        for (long k = 0; k < NFX(nx, ny); ++k) s->q0[lev * NF(nx, ny) + k] = u * d;
        for (long k = NFX(nx, ny); k < NF(nx, ny); ++k) s->q0[lev * NF(nx, ny) + k] = v * d;
    }
    return s;
}

void og_sim_destroy(og_sim* s) {
    if (!s) return;
    for (int b = 0; b < s->nbody; ++b) free(s->bodies[b].table);
    free(s->bf.table);
    free(s->refx); free(s->refy); free(s->px); free(s->py); free(s->pu); free(s->pv);
    free(s->abi); free(s->afi); free(s->aw);
    free(s->omega); free(s->nprev); free(s->q); free(s->q0); free(s->f);
    for (int k = 0; k < 64; ++k) { free(s->bq[k]); free(s->bw[k]); }
    free(s->what); free(s->sfd_prev);
    for (int i = 0; i < 4; ++i) { free(s->chol_diag[i]); free(s->chol_low[i]); }
    free(s);
}

/* ibpm.cc:166-178 with -ubf: a single body's motion moves to the base flow
   (Geometry::transferMotion / transferCenter, Geometry.cc:46-77); the body
   stays at its reference points and the projection uses Cholesky.  Call
   before og_sim_init.  Returns 1 if the motion moved, 0 for a stationary
   body, -1 for more than one body (the reference exits). */
int og_sim_set_ubf(og_sim* s) {
    if (s->nbody != 1) { snprintf(g_err, sizeof g_err, "ubf: only for a single moving body"); return -1; }
    if (s->stationary) return 0;
    s->bf = s->bodies[0];
    s->bodies[0].table = NULL;
    s->bodies[0].motion = OG_NONE;
    s->stationary = 1;
    s->use_cg = s->force_cg;
    s->ubf = 1;
    return 1;
}

/* The base flux after moveFlow at each time of a sequence (all levels, the
   Flux layout) and (alpha, mag) after it, for a geometry with one moving
   body under -ubf. */
int og_baseflow_sequence(const char* geom, int nx, int ny, int ngrid, double length, double xoff, double yoff,
                         double alpha_deg, int ntimes, const double* times, double* q_out, double* am_out) {
    og_sim* s = og_sim_create(nx, ny, ngrid, length, xoff, yoff, 100.0, 0.01, 2, alpha_deg, geom, 1e-7, 0);
    if (!s) return -1;
    if (og_sim_set_ubf(s) != 1) { og_sim_destroy(s); return -1; }
    const long n = f_n(s);
    for (int c = 0; c < ntimes; ++c) {
        bf_move(s, times[c]);
        memcpy(q_out + n * c, s->q0, sizeof(double) * n);
        am_out[2 * c] = s->bf_alpha;
        am_out[2 * c + 1] = s->bf_mag;
    }
    og_sim_destroy(s);
    return 0;
}

/* The model of ibpm.cc:185-235 (call before og_sim_init).  The linear,
   adjoint and linear periodic models use NavierStokesModel(grid, geom, Re),
   with no potential-flow part (q0 = 0); SFD keeps it.  period: base states of
   the linear periodic model. */
int og_sim_set_model(og_sim* s, int model, int period, double chi, double delta) {
    if (model < 0 || model > 4 || period < 1 || period > 64) {
        snprintf(g_err, sizeof g_err, "bad model %d or period %d", model, period);
        return -1;
    }
    s->model = model;
    s->period = model == 3 ? period : 1;
    s->chi = chi;
    s->delta = delta;
    if (model >= 1 && model <= 3) memset(s->q0, 0, sizeof(double) * f_n(s));
    if (model == 4) {
        s->what = (double*)calloc(s_n(s), sizeof(double));   /* the reference's is unset memory; here zero */
        s->sfd_prev = (double*)calloc(s_n(s), sizeof(double));
    }
    return 0;
}

/* Base state k (vorticity and flux) of the linear models (x00.load in
   ibpm.cc:191-218). */
int og_sim_set_base_state(og_sim* s, int k, const double* w, const double* q) {
    if (k < 0 || k >= 64) return -1;
    if (!s->bw[k]) s->bw[k] = (double*)malloc(sizeof(double) * s_n(s));
    if (!s->bq[k]) s->bq[k] = (double*)malloc(sizeof(double) * f_n(s));
    memcpy(s->bw[k], w, sizeof(double) * s_n(s));
    memcpy(s->bq[k], q, sizeof(double) * f_n(s));
    return 0;
}

/* SFDSolver::loadFilteredState (IBSolver.cc:289-311): the filtered state is
   set (and marked saved); call after og_sim_init. */
int og_sim_set_filtered(og_sim* s, const double* w) {
    if (s->model != 4) return -1;
    memcpy(s->what, w, sizeof(double) * s_n(s));
    s->what_saved = 1;
    return 0;
}

int og_sim_get_filtered(const og_sim* s, double* w) {
    if (s->model != 4) return -1;
    memcpy(w, s->what, sizeof(double) * s_n(s));
    return 0;
}

int og_sim_num_points(const og_sim* s) { return s->nb; }
int og_sim_uses_cg(const og_sim* s) { return s->use_cg; }

int og_sim_init(og_sim* s) {
    move_bodies(s, s->time);   /* ibpm.cc:285 */
    reg_update(s);             /* model->init(), NavierStokesModel.cc:52-57 */
    if (!s->use_cg)
        for (int i = 0; i < s->nsub; ++i) chol_build(s, i);   /* solver->init(), ibpm.cc:291-297 */
    if (s->ubf) bf_move(s, s->time);                                  /* updateOperators(t) */
    if (!s->stationary) { move_bodies(s, s->time); reg_update(s); }
    refresh_state(s);          /* ibpm.cc:304-305 */
    return 0;
}

/* Timing variant of og_sim_init for the CPU baseline: the Cholesky factors
   are set to the identity instead of being built (the triangular solves then
   run the same O(n^2) loops per substep), so a bounded sample of the step
   loop can be timed without the one-time factorisation. */
int og_sim_init_timing(og_sim* s) {
    move_bodies(s, s->time);
    reg_update(s);
    if (!s->use_cg) {
        const int n = 2 * s->nb;
        for (int i = 0; i < s->nsub; ++i) {
            s->chol_diag[i] = (double*)malloc(sizeof(double) * (n + 1));
            s->chol_low[i] = (double*)calloc((size_t)n * n + 1, sizeof(double));
            for (int k = 0; k < n; ++k) s->chol_diag[i][k] = 1.0;
        }
    }
    refresh_state(s);
    return 0;
}

/* IBSolver::advance + advanceSubstep, IBSolver.cc:142-198 */
int og_sim_advance(og_sim* s) {
    const long n = s_n(s);
    const double nu = 1. / s->Re;
    double* nl = (double*)malloc(sizeof(double) * n);
    double* a = (double*)malloc(sizeof(double) * n);
    double* b = (double*)malloc(sizeof(double) * (2 * s->nb + 1));
    double* wtemp = s->model == 4 ? (double*)malloc(sizeof(double) * n) : NULL;
    for (int i = 0; i < s->nsub; ++i) {
        switch (s->model) {   /* IBSolver::advance: N(x) of the model, IBSolver.cc:142-150, 203-246 */
        case 1: og_linear_N(s->nx, s->ny, s->ngrid, s->dx, s->bq[0], s->bw[0], s->q, s->omega, nl); break;
        case 2: og_adjoint_N(s->nx, s->ny, s->ngrid, s->dx, s->bq[0], s->bw[0], s->q, nl); break;
        case 3: {   /* phase k = timestep % period, IBSolver.cc:224-231 */
            const int k = s->timestep % s->period;
            og_linear_N(s->nx, s->ny, s->ngrid, s->dx, s->bq[k], s->bw[k], s->q, s->omega, nl);
            break;
        }
        case 4:   /* SFDSolver::N, IBSolver.cc:238-244 */
            og_nonlinear(s->nx, s->ny, s->ngrid, s->dx, s->q, s->omega, nl);
            for (long k = 0; k < n; ++k) nl[k] -= s->chi * (s->omega[k] - s->what[k]);
            break;
        default: og_nonlinear(s->nx, s->ny, s->ngrid, s->dx, s->q, s->omega, nl);
        }
        if (s->model == 4) {   /* SFDSolver::advanceSubstep, IBSolver.cc:247-252 */
            if (!s->what_saved) { memcpy(s->what, s->omega, sizeof(double) * n); s->what_saved = 1; }
            memcpy(wtemp, s->omega, sizeof(double) * n);
        }
        if (s->ubf) bf_move(s, s->time + s->cn[i] * s->dt);   /* NavierStokesModel.cc:93-99 */
        if (!s->stationary) {
            move_bodies(s, s->time + s->cn[i] * s->dt);
            reg_update(s);
        }
        og_laplacian(s->nx, s->ny, s->ngrid, s->dx, s->omega, a);
        const double an = s->an[i], bn = s->bn[i];
        for (long k = 0; k < n; ++k) a[k] *= 0.5 * nu * (an + bn);
        for (long k = 0; k < n; ++k) a[k] += an * nl[k];
        if (bn != 0) {
            if (!s->old_saved) memcpy(s->nprev, nl, sizeof(double) * n);
            for (long k = 0; k < n; ++k) a[k] += bn * s->nprev[k];
        }
        for (long k = 0; k < n; ++k) a[k] *= s->dt;
        for (long k = 0; k < n; ++k) a[k] += s->omega[k];
        get_constraints(s, b);
        og_sim_projection_solve(s, i, a, b, s->omega, s->f);
        refresh_state(s);
        memcpy(s->nprev, nl, sizeof(double) * n);
        s->old_saved = 1;
        if (s->model == 4) {   /* the filtered state, IBSolver.cc:257-279 */
            const double inv = 1.0 / s->delta;   /* Scalar / double multiplies by the reciprocal (Array.h:335-340) */
            for (long k = 0; k < n; ++k) {
                const double rhs = (wtemp[k] - s->what[k]) * inv;
                double av = an * rhs;
                if (bn != 0) {
                    if (!s->sfd_saved) s->sfd_prev[k] = rhs;
                    av += bn * s->sfd_prev[k];
                }
                av *= s->dt;
                s->what[k] += av;
                s->sfd_prev[k] = rhs;
            }
            s->sfd_saved = 1;
        }
    }
    free(wtemp);
    s->time += s->dt;
    s->timestep++;
    if (s->ubf) {   /* ibpm.cc:347-351: forces rotated by the base-flow angle at the new time */
        bf_set_alpha_mag(s, s->time);
        s->force_alpha = s->bf_alpha;
    }
    free(nl); free(a); free(b);
    return 0;
}

/* State::computeNetForce (State.cc:53-63) and ibpm.cc:346-355 */
void og_sim_forces(const og_sim* s, double* cd, double* cl) {
    double xF = 0, yF = 0;
    for (int k = 0; k < s->nb; ++k) { xF += s->f[k]; yF += s->f[s->nb + k]; }
    const double dx2 = s->dx * s->dx;
    xF *= dx2; yF *= dx2;
    double ca, sa;
    sin_cos(s->force_alpha, &sa, &ca);
    const double drag = xF * ca + yF * sa;
    const double lift = xF * -1. * sa + yF * ca;
    *cd = 2 * drag; *cl = 2 * lift;
}

int og_geom_points(const char* geom, double t, double* xy, double* uv, int cap) {
    og_sim* s = og_sim_create(8, 8, 1, 1.0, 0.0, 0.0, 100.0, 0.01, 2, 0.0, geom, 1e-7, 0);
    if (!s) return -1;
    if (s->nb > cap) { og_sim_destroy(s); return -2; }
    move_bodies(s, t);
    for (int k = 0; k < s->nb; ++k) {
        xy[k] = s->px[k]; xy[s->nb + k] = s->py[k];
        uv[k] = s->pu[k]; uv[s->nb + k] = s->pv[k];
    }
    int n = s->nb;
    og_sim_destroy(s);
    return n;
}

/* moveBodies at each time of a sequence, in order (the motions that integrate
   keep their state between calls); points [x.., y..] and velocities
   [u.., v..] of call k at offset 2 n k. */
int og_geom_sequence(const char* geom, int ntimes, const double* times, double* xy, double* uv, int cap) {
    og_sim* s = og_sim_create(8, 8, 1, 1.0, 0.0, 0.0, 100.0, 0.01, 2, 0.0, geom, 1e-7, 0);
    if (!s) return -1;
    const int n = s->nb;
    if (2L * n * ntimes > cap) { og_sim_destroy(s); return -2; }
    for (int c = 0; c < ntimes; ++c) {
        move_bodies(s, times[c]);
        for (int k = 0; k < n; ++k) {
            xy[2L * n * c + k] = s->px[k]; xy[2L * n * c + n + k] = s->py[k];
            uv[2L * n * c + k] = s->pu[k]; uv[2L * n * c + n + k] = s->pv[k];
        }
    }
    og_sim_destroy(s);
    return n;
}

void og_sim_get_omega(const og_sim* s, double* out) { memcpy(out, s->omega, sizeof(double) * s_n(s)); }
void og_sim_set_omega(og_sim* s, const double* in) { memcpy(s->omega, in, sizeof(double) * s_n(s)); refresh_state(s); }
void og_sim_get_q(const og_sim* s, double* out) { memcpy(out, s->q, sizeof(double) * f_n(s)); }
void og_sim_get_f(const og_sim* s, double* out) { memcpy(out, s->f, sizeof(double) * 2 * s->nb); }
void og_sim_get_points(const og_sim* s, double* xy) {
    for (int k = 0; k < s->nb; ++k) { xy[k] = s->px[k]; xy[s->nb + k] = s->py[k]; }
}
double og_sim_time(const og_sim* s) { return s->time; }
int og_sim_last_cg_iters(const og_sim* s, int* it) {
    for (int i = 0; i < s->nsub; ++i) it[i] = s->cg_iters[i];
    return s->nsub;
}
int og_sim_cholesky(const og_sim* s, int i, double* diag, double* lower) {
    if (s->use_cg || !s->chol_diag[i]) return -1;
    const int n = 2 * s->nb;
    memcpy(diag, s->chol_diag[i], sizeof(double) * n);
    memcpy(lower, s->chol_low[i], sizeof(double) * (size_t)n * n);
    return 0;
}
