// boundary_test.cpp -- the reference's solver-class tests, restated against
// the C++ classes of ibpm_amd/host/ibpm_gpu.hpp (GPU path):
//   * EllipticSolver2dTest (reference test/EllipticSolver2dTest.cc:43-172):
//     Poisson and Helmholtz solves of every single-wavenumber field on a 4 x 8
//     grid, and the constant and linear boundary-value closed forms;
//   * EllipticSolverTest's multi-domain identity L(solve f) = coarsify(f)
//     (test/EllipticSolverTest.cc:86-112), here on a 3-level grid;
//   * ProjectionSolverTest (test/ProjectionSolverTest.cc:70-165): the
//     projection equations (1 - h/2 L) omega + h B f = a, C omega = b for the
//     CG and Cholesky solvers, with and without a body, and the Cholesky
//     factor's save / load round trip with its size and timestep checks.
// Prints one line per case and exits 1 on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <unistd.h>

#include "../../ibpm_amd/host/ibpm_gpu.hpp"

using namespace ibpm;

static int g_fail = 0;
static void check(bool ok, const std::string& what, double err) {
    std::printf("%-60s %s (%.3e)\n", what.c_str(), ok ? "ok" : "FAIL", err);
    std::fflush(stdout);
    if (!ok) g_fail = 1;
}

// test/SingleWavenumber.h:62-79 (Array2 form)
static void single_wavenumber(int kx, int ky, Array2d& f) {
    const double pi = 4 * std::atan(1.);
    const int nx = f.n1() + 1, ny = f.n2() + 1;
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j) f(i, j) = std::sin(kx * i * pi / nx) * std::sin(ky * j * pi / ny);
}

// test/SingleWavenumber.h:30-59 (Scalar form)
static void single_wavenumber(int kx, int ky, const Grid& g, Scalar& f) {
    const double pi = 4 * std::atan(1.);
    const int L = g.Ngrid();
    const double x0 = g.getXEdge(L - 1, 0), y0 = g.getYEdge(L - 1, 0);
    const double kxx = kx * pi / (g.getXEdge(L - 1, g.Nx()) - x0), kyy = ky * pi / (g.getYEdge(L - 1, g.Ny()) - y0);
    for (int lev = 0; lev < L; ++lev)
        for (int i = 1; i < g.Nx(); ++i)
            for (int j = 1; j < g.Ny(); ++j)
                f(lev, i, j) = std::sin(kxx * (g.getXEdge(lev, i) - x0)) * std::sin(kyy * (g.getYEdge(lev, j) - y0));
}

// 5-point Laplacian of one level with boundary values (the reference's
// Laplacian(u, dx, bc, Lu) of a 2-D array, VectorOperations.cc)
static void laplacian2d(const Array2d& u, double dx, BC& bc, Array2d& Lu) {
    const int nx = u.n1() + 1, ny = u.n2() + 1;
    auto at = [&](int i, int j) -> double {
        if (i == 0) return bc.left(j);
        if (i == nx) return bc.right(j);
        if (j == 0) return bc.bottom(i);
        if (j == ny) return bc.top(i);
        return u(i, j);
    };
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            Lu(i, j) = (at(i + 1, j) + at(i - 1, j) + at(i, j + 1) + at(i, j - 1) - 4 * u(i, j)) / (dx * dx);
}

static double maxdiff(const std::vector<double>& a, const std::vector<double>& b) {
    double m = 0;
    for (size_t k = 0; k < a.size(); ++k) m = std::max(m, std::fabs(a[k] - b[k]));
    return m;
}

static void linear_bc(BC& bc, int nx, int ny, double dx) {   // u = x + 2 y
    for (int i = 0; i <= nx; ++i) {
        bc.bottom(i) = i * dx;
        bc.top(i) = i * dx + 2 * ny * dx;
    }
    for (int j = 0; j <= ny; ++j) {
        bc.left(j) = 2 * j * dx;
        bc.right(j) = nx * dx + 2 * j * dx;
    }
}

static void elliptic2d_tests() {
    const int nx = 4, ny = 8;
    const double dx = 0.1, alpha = 0.2, tol = 1e-10;
    PoissonSolver2d poisson(nx, ny, dx);
    HelmholtzSolver2d helmholtz(nx, ny, dx, alpha);
    Array2d u(nx - 1, ny - 1), f(nx - 1, ny - 1), Lu(nx - 1, ny - 1);
    BC zero(nx, ny);
    double e1 = 0, e2 = 0;
    for (int kx = 0; kx < nx; ++kx)
        for (int ky = 0; ky < ny; ++ky) {
            single_wavenumber(kx, ky, f);
            poisson.solve(f, u);
            laplacian2d(u, dx, zero, Lu);
            e1 = std::max(e1, maxdiff(f, Lu));
            helmholtz.solve(f, u);
            laplacian2d(u, dx, zero, Lu);
            for (size_t k = 0; k < Lu.size(); ++k) Lu[k] = u[k] + alpha * Lu[k];
            e2 = std::max(e2, maxdiff(f, Lu));
        }
    check(e1 <= tol, "EllipticSolver2d Poisson2d: L solve f = f", e1);
    check(e2 <= tol, "EllipticSolver2d Helmholtz2d: (1 + a L) solve f = f", e2);

    BC bc(nx, ny);
    bc = 8.0;
    f = 0.0;
    poisson.solve(f, bc, u);
    Array2d c8(nx - 1, ny - 1);
    c8 = 8.0;
    check(maxdiff(u, c8) <= tol, "EllipticSolver2d Poisson2dConstWithBC", maxdiff(u, c8));
    helmholtz.solve(f, bc, u);
    laplacian2d(u, dx, bc, Lu);
    for (size_t k = 0; k < Lu.size(); ++k) Lu[k] *= -alpha;
    check(maxdiff(u, Lu) <= tol, "EllipticSolver2d Helmholtz2dWithConstBC", maxdiff(u, Lu));

    BC lin(nx, ny);
    linear_bc(lin, nx, ny, dx);
    poisson.solve(f, lin, u);
    Array2d ex(nx - 1, ny - 1);
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j) ex(i, j) = i * dx + 2 * j * dx;
    check(maxdiff(u, ex) <= tol, "EllipticSolver2d Poisson2dLinearWithBC", maxdiff(u, ex));
    helmholtz.solve(f, lin, u);
    laplacian2d(u, dx, lin, Lu);
    for (size_t k = 0; k < Lu.size(); ++k) Lu[k] *= -alpha;
    check(maxdiff(u, Lu) <= tol, "EllipticSolver2d Helmholtz2dWithLinearBC", maxdiff(u, Lu));
}

// Scalar Laplacian (VectorOperations.cc:133-140) through the C ABI
static Scalar laplacian(const Grid& g, const Scalar& w) {
    Scalar out(g);
    DeviceBuffer dw(w), dout(out.size());
    gpu_check(ibpm_laplacian(default_context(), g.Nx(), g.Ny(), g.Ngrid(), g.Dx(), dw.get(), dout.get()));
    dout.to(out);
    return out;
}

static void multidomain_test() {
    // EllipticSolverTest.cc:86-112: L(solve f) = coarsify(f) on every level
    Grid g(16, 24, 3, 4.0, -2.0, -3.0);
    Scalar f(g);
    single_wavenumber(1, 1, g, f);
    for (size_t k = 0; k < f.size(); ++k) f[k] += 0.1 * std::sin(3.7 * k);   // not coarsify-consistent
    PoissonSolver poisson(g);
    Scalar u = poisson.solve(f);
    Scalar Lu = laplacian(g, u);
    // coarsify(f) (Scalar.cc:46-63): central regions of levels 1.. from the finer level
    Scalar cf = f;
    const int nx = g.Nx(), ny = g.Ny(), ex = nx / 4, ey = ny / 4;
    for (int lev = 1; lev < g.Ngrid(); ++lev)
        for (int i = ex + 1; i < nx / 2 + ex; ++i)
            for (int j = ey + 1; j < ny / 2 + ey; ++j) {
                const int ii = 2 * (i - ex), jj = 2 * (j - ey);
                cf(lev, i, j) = 0.25 * cf(lev - 1, ii, jj) +
                                0.125 * (cf(lev - 1, ii + 1, jj) + cf(lev - 1, ii, jj + 1) + cf(lev - 1, ii - 1, jj) +
                                         cf(lev - 1, ii, jj - 1)) +
                                0.0625 * (cf(lev - 1, ii + 1, jj + 1) + cf(lev - 1, ii + 1, jj - 1) +
                                          cf(lev - 1, ii - 1, jj + 1) + cf(lev - 1, ii - 1, jj - 1));
            }
    double e = maxdiff(Lu, cf), scale = 0;
    for (double v : cf) scale = std::max(scale, std::fabs(v));
    check(e <= 1e-10 * scale, "EllipticSolver PoissonMultiDomain: L solve f = coarsify f", e / scale);
}

// ProjectionSolverTest::verify (test/ProjectionSolverTest.cc:80-112)
static void verify(const Grid& g, NavierStokesModel& model, ProjectionSolver& solver, double h,
                   const std::string& name) {
    const int n = model.getNumPoints();
    Scalar a(g), omega(g);
    single_wavenumber(1, 1, g, a);
    BoundaryVector b(n), f(n);
    b = 3.0;
    solver.solve(a, b, omega, f);
    Scalar lhs = laplacian(g, omega);
    lhs *= model.getAlpha() * (-h / 2.);
    lhs += omega;
    Scalar forcing(g);
    model.B(f, forcing);
    forcing *= h;
    lhs += forcing;
    check(maxdiff(a, lhs) <= 1e-10, name + ": (1 - h/2 L) omega + h B f = a", maxdiff(a, lhs));
    BoundaryVector cons(n);
    model.C(omega, cons);
    check(maxdiff(b, cons) <= 1e-10, name + ": C omega = b", n ? maxdiff(b, cons) : 0.0);
}

static void projection_tests() {
    const double h = 0.123, Re = 100.0;
    Grid g(6, 6, 1, 2.0, -1.0, -1.0);
    Geometry empty, body;
    empty.loadText("body Empty\nend\n");
    body.loadText("body Line\n  line_n -0.75 0 0.75 0 4\nend\n");
    BaseFlow q0(g, 1.0, 0.0);
    NavierStokesModel noBodies(g, empty, Re, q0), withBodies(g, body, Re, q0);
    noBodies.init();
    withBodies.init();
    {
        ConjugateGradientSolver s(g, noBodies, h, 1e-10);
        verify(g, noBodies, s, h, "CGSolver NoConstraints");
    }
    {
        ConjugateGradientSolver s(g, withBodies, h, 1e-10);
        verify(g, withBodies, s, h, "CGSolver WithConstraints");
    }
    {
        CholeskySolver s(g, noBodies, h);
        s.init();
        verify(g, noBodies, s, h, "CholeskySolver NoConstraints");
    }
    {
        CholeskySolver s(g, withBodies, h);
        s.init();
        verify(g, withBodies, s, h, "CholeskySolver WithConstraints");
    }
    {   // CholeskySolverTest.SaveFile
        const std::string base = "boundary_test_solver_" + std::to_string((long)getpid());
        CholeskySolver s(g, withBodies, h);
        s.init();
        check(s.save(base), "CholeskySolver SaveFile: save", 0.0);
        CholeskySolver loaded(g, withBodies, h);
        check(loaded.load(base), "CholeskySolver SaveFile: load", 0.0);
        verify(g, withBodies, loaded, h, "CholeskySolver SaveFile (loaded)");
        CholeskySolver wrongSize(g, noBodies, h);
        check(!wrongSize.load(base), "CholeskySolver SaveFile: wrong size refused", 0.0);
        CholeskySolver wrongStep(g, withBodies, 2 * h);
        check(!wrongStep.load(base), "CholeskySolver SaveFile: wrong timestep refused", 0.0);
        unlink((base + ".cholesky").c_str());
    }
}

int main() {
    try {
        elliptic2d_tests();
        multidomain_test();
        projection_tests();
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 2;
    }
    std::printf("%s\n", g_fail ? "FAILED" : "all ok");
    return g_fail;
}
