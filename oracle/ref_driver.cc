// ref_driver.cc -- extern "C" shim over the FFTW-free part of the reference
// (Grid, Scalar, Flux, BC, BoundaryVector, Regularizer, Geometry, RigidBody,
// State, BaseFlow, the Motion classes), compiled from the sources where they lie under the reference's src/
// by oracle/Makefile into oracle/_ref/libibpm_refpart.so.
//
// TEST INFRASTRUCTURE ONLY: it pins the CPU restatement (ibpm_oracle.c)
// against the reference's own code for coarsify/getBC, the regulariser, the
// .geom parser with every body motion, the uniform and the unsteady base
// flow and the net force.
// The elliptic solvers, VectorOperations and the projection solvers include
// FFTW3, which this image lacks, so they are not part of this build.
#include "Grid.h"
#include "Scalar.h"
#include "Flux.h"
#include "BC.h"
#include "BoundaryVector.h"
#include "Regularizer.h"
#include "Geometry.h"
#include "State.h"
#include "BaseFlow.h"
#include "Motion.h"
#include <cstring>
#include <string>

using namespace ibpm;

namespace {
void load_scalar(Scalar& s, const double* d) {
    const Grid& g = s.getGrid();
    long k = 0;
    for (int lev = 0; lev < g.Ngrid(); ++lev)
        for (int i = 1; i < g.Nx(); ++i)
            for (int j = 1; j < g.Ny(); ++j) s(lev, i, j) = d[k++];
}
void store_scalar(const Scalar& s, double* d) {
    const Grid& g = s.getGrid();
    long k = 0;
    for (int lev = 0; lev < g.Ngrid(); ++lev)
        for (int i = 1; i < g.Nx(); ++i)
            for (int j = 1; j < g.Ny(); ++j) d[k++] = s(lev, i, j);
}
}  // namespace

extern "C" {

void ref_coarsify(int nx, int ny, int ngrid, double length, double xoff, double yoff, double* data) {
    Grid grid(nx, ny, ngrid, length, xoff, yoff);
    Scalar s(grid);
    load_scalar(s, data);
    s.coarsify();
    store_scalar(s, data);
}

void ref_getbc(int nx, int ny, int ngrid, double length, double xoff, double yoff, const double* data,
               int lev, double* bc_out) {
    Grid grid(nx, ny, ngrid, length, xoff, yoff);
    Scalar s(grid);
    load_scalar(s, data);
    BC bc(nx, ny);
    s.getBC(lev, bc);
    // ring order of BC.h: left(0..ny), top(1..nx), right(ny-1..0), bottom(nx-1..1)
    for (int j = 0; j <= ny; ++j) bc_out[j] = bc.left(j);
    for (int i = 0; i <= nx; ++i) bc_out[ny + i] = bc.top(i);
    for (int j = 0; j <= ny; ++j) bc_out[2 * ny + nx - j] = bc.right(j);
    for (int i = 1; i < nx; ++i) bc_out[2 * (nx + ny) - i] = bc.bottom(i);
}

// Loads a .geom file, moves bodies to time t, writes points [x.., y..] and
// velocities [u.., v..]; returns the number of points or -1.
int ref_geometry(const char* path, double t, double* xy, double* uv, int cap) {
    Geometry geom;
    if (!geom.load(std::string(path))) return -1;
    geom.moveBodies(t);
    int n = geom.getNumPoints();
    if (n > cap) return -2;
    BoundaryVector p = geom.getPoints();
    BoundaryVector v = geom.getVelocities();
    for (int i = 0; i < n; ++i) {
        xy[i] = p(X, i); xy[n + i] = p(Y, i);
        uv[i] = v(X, i); uv[n + i] = v(Y, i);
    }
    return n;
}

// Loads a .geom file and calls moveBodies at each time of a sequence, in order
// (FixedVelocity, Eldredge2 and EldredgeCombined2 integrate over the calls);
// call k writes points [x.., y..] and velocities [u.., v..] at offset 2 n k.
// Returns the number of points per call or -1.
int ref_geometry_sequence(const char* path, int ntimes, const double* times, double* xy, double* uv, int cap) {
    Geometry geom;
    if (!geom.load(std::string(path))) return -1;
    int n = geom.getNumPoints();
    if (2L * n * ntimes > cap) return -2;
    for (int c = 0; c < ntimes; ++c) {
        geom.moveBodies(times[c]);
        BoundaryVector p = geom.getPoints();
        BoundaryVector v = geom.getVelocities();
        for (int i = 0; i < n; ++i) {
            xy[2L * n * c + i] = p(X, i); xy[2L * n * c + n + i] = p(Y, i);
            uv[2L * n * c + i] = v(X, i); uv[2L * n * c + n + i] = v(Y, i);
        }
    }
    return n;
}

int ref_geometry_stationary(const char* path) {
    Geometry geom;
    if (!geom.load(std::string(path))) return -1;
    return geom.isStationary() ? 1 : 0;
}

// Regularizer on the geometry in `path` at time t: q_out = toFlux(f_in) (all
// levels) and f_out = toBoundary(q_in).
int ref_regularizer(int nx, int ny, int ngrid, double length, double xoff, double yoff,
                    const char* path, double t, const double* f_in, double* q_out,
                    const double* q_in, double* f_out) {
    Grid grid(nx, ny, ngrid, length, xoff, yoff);
    Geometry geom;
    if (!geom.load(std::string(path))) return -1;
    geom.moveBodies(t);
    Regularizer reg(grid, geom);
    reg.update();
    int n = geom.getNumPoints();
    BoundaryVector f(n);
    for (int k = 0; k < 2 * n; ++k) f(k) = f_in[k];
    Flux q = reg.toFlux(f);
    long nf = 2L * nx * ny + nx + ny;
    for (int lev = 0; lev < ngrid; ++lev)
        for (long k = 0; k < nf; ++k) q_out[lev * nf + k] = q(lev, (int)k);
    Flux qi(grid);
    for (int lev = 0; lev < ngrid; ++lev)
        for (long k = 0; k < nf; ++k) qi(lev, (int)k) = q_in[lev * nf + k];
    BoundaryVector fo = reg.toBoundary(qi);
    for (int k = 0; k < 2 * n; ++k) f_out[k] = fo(k);
    return n;
}

// Flux::UniformFlow (Flux.cc:79-99) for magnitude 1 and angle alpha (radians).
void ref_uniform_flow(int nx, int ny, int ngrid, double length, double xoff, double yoff,
                      double alpha, double* q_out) {
    Grid grid(nx, ny, ngrid, length, xoff, yoff);
    Flux q = Flux::UniformFlow(grid, 1.0, alpha);
    long nf = 2L * nx * ny + nx + ny;
    for (int lev = 0; lev < ngrid; ++lev)
        for (long k = 0; k < nf; ++k) q_out[lev * nf + k] = q(lev, (int)k);
}

// Unsteady base flow of ibpm.cc:166-178 for the one-body geometry in `path`:
// BaseFlow(grid, 1, alpha) takes the body's motion and centre, then moveFlow
// at each time of a sequence; writes the base flux (all levels, nf per level)
// and (getAlpha, getMag) after each call.  Returns 0, or -1 when the geometry
// is not a single moving body.
int ref_baseflow_sequence(const char* path, int nx, int ny, int ngrid, double length, double xoff, double yoff,
                          double alpha, int ntimes, const double* times, double* q_out, double* am_out) {
    Grid grid(nx, ny, ngrid, length, xoff, yoff);
    Geometry geom;
    if (!geom.load(std::string(path))) return -1;
    if (geom.isStationary() || geom.getNumBodies() != 1) return -1;
    BaseFlow q(grid, 1.0, alpha);
    Motion* m = geom.transferMotion();
    double xC = 0, yC = 0;
    geom.transferCenter(xC, yC);
    q.setMotion(*m);
    q.setCenter(xC, yC);
    const long nf = 2L * nx * ny + nx + ny;
    for (int c = 0; c < ntimes; ++c) {
        q.moveFlow(times[c]);
        Flux f = q.getFlux();
        for (int lev = 0; lev < ngrid; ++lev)
            for (long k = 0; k < nf; ++k) q_out[(long)c * ngrid * nf + lev * nf + k] = f(lev, (int)k);
        am_out[2 * c] = q.getAlpha();
        am_out[2 * c + 1] = q.getMag();
    }
    delete m;
    return 0;
}

// State::computeNetForce (State.cc:53-63).
void ref_net_force(int nx, int ny, int ngrid, double length, double xoff, double yoff, int npts,
                   const double* f, double* xf, double* yf) {
    Grid grid(nx, ny, ngrid, length, xoff, yoff);
    State x(grid, npts);
    for (int k = 0; k < 2 * npts; ++k) x.f(k) = f[k];
    x.computeNetForce(*xf, *yf);
}

// State::save of a state built from flat arrays (State.cc:158-223).
int ref_state_save(const char* path, int nx, int ny, int ngrid, double length, double xoff, double yoff, int npts,
                   const double* q, const double* omega, const double* f, int timestep, double time) {
    Grid grid(nx, ny, ngrid, length, xoff, yoff);
    State x(grid, npts);
    long nf = 2L * nx * ny + nx + ny;
    for (int lev = 0; lev < ngrid; ++lev)
        for (long k = 0; k < nf; ++k) x.q(lev, (int)k) = q[lev * nf + k];
    load_scalar(x.omega, omega);
    for (int k = 0; k < 2 * npts; ++k) x.f(k) = f[k];
    x.timestep = timestep;
    x.time = time;
    return x.save(path) ? 0 : -1;
}

// State::load (State.cc:65-156) into flat arrays sized by the caller.
int ref_state_load(const char* path, int nx, int ny, int ngrid, double length, double xoff, double yoff, int npts,
                   double* q, double* omega, double* f, int* timestep, double* time) {
    Grid grid(nx, ny, ngrid, length, xoff, yoff);
    State x(grid, npts);
    if (!x.load(path)) return -1;
    long nf = 2L * nx * ny + nx + ny;
    for (int lev = 0; lev < ngrid; ++lev)
        for (long k = 0; k < nf; ++k) q[lev * nf + k] = x.q(lev, (int)k);
    store_scalar(x.omega, omega);
    for (int k = 0; k < 2 * npts; ++k) f[k] = x.f(k);
    *timestep = x.timestep;
    *time = x.time;
    return 0;
}

}  // extern "C"
