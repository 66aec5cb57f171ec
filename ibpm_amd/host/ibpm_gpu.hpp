// ibpm_gpu.hpp -- C++ host classes with the reference's interfaces over the
// C ABI of include/ibpm_gpu.h, so a driver written against the reference
// (reference src/ibpm.cc:140-372) compiles against the GPU path:
//
//   Grid grid(nx, ny, ngrid, length, xoffset, yoffset);     // Grid.h:34-41
//   Geometry geom;  geom.load("cylinder.geom");               // Geometry.h
//   BaseFlow q_potential(grid, 1.0, alpha);                    // BaseFlow.h
//   NavierStokesModel model(grid, geom, Re, q_potential);      // NavierStokesModel.h:45-56
//   NonlinearIBSolver solver(grid, model, dt, Scheme::RK3);    // IBSolver.h:69-90
//   State x(grid, geom.getNumPoints());                        // State.h
//   model.init(); solver.init();
//   for (...) { solver.advance(x); x.computeNetForce(xF, yF); }
//
// Fields live on the GPU; State keeps host views that are refreshed after
// every advance (f, time, timestep) or on request (omega, q).
#pragma once
#include <cmath>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ibpm_gpu.h"

namespace ibpm {

inline void gpu_check(int rc) {
    if (rc < 0) throw std::runtime_error(std::string("ibpm_gpu: ") + ibpm_last_error());
}

class Grid {   // reference src/Grid.h:32-176 (zero shift)
public:
    Grid(int nx, int ny, int ngrid, double length, double xOffset, double yOffset)
        : _nx(nx), _ny(ny), _ngrid(ngrid), _length(length), _x0(xOffset), _y0(yOffset), _dx(length / nx) {}
    int Nx() const { return _nx; }
    int Ny() const { return _ny; }
    int Ngrid() const { return _ngrid; }
    double Dx() const { return _dx; }
    double Dx(int lev) const { return _dx * (1 << lev); }
    double length() const { return _length; }
    double getXEdge(int lev, int i) const { return _x0 - 0.5 * ((1 << lev) - 1) * (_nx * _dx) + i * Dx(lev); }
    double getYEdge(int lev, int j) const { return _y0 - 0.5 * ((1 << lev) - 1) * (_ny * _dx) + j * Dx(lev); }
    double xOffset() const { return _x0; }
    double yOffset() const { return _y0; }

private:
    int _nx, _ny, _ngrid;
    double _length, _x0, _y0, _dx;
};

class Motion {};   // reference src/Motion.h: a token here; the library holds the motion

// The context of the operator-level classes below (one per device, kept for
// the life of the process).
inline ibpm_ctx* default_context(int device = 0) {
    static ibpm_ctx* ctx[64] = {};
    if (device < 0 || device >= 64) throw std::runtime_error("ibpm_gpu: bad device");
    if (!ctx[device]) gpu_check(ibpm_ctx_create(device, &ctx[device]));
    return ctx[device];
}

// Device copy of a host array for one call (ibpm_malloc / ibpm_memcpy_*).
class DeviceBuffer {
public:
    explicit DeviceBuffer(size_t n, ibpm_ctx* ctx = default_context()) : _ctx(ctx), _n(n) {
        gpu_check(ibpm_malloc(_ctx, sizeof(double) * (n ? n : 1), &_p));
    }
    DeviceBuffer(const std::vector<double>& h, ibpm_ctx* ctx = default_context()) : DeviceBuffer(h.size(), ctx) {
        if (_n) gpu_check(ibpm_memcpy_h2d(_ctx, _p, h.data(), sizeof(double) * _n));
    }
    ~DeviceBuffer() { ibpm_free(_ctx, _p); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    double* get() const { return (double*)_p; }
    void to(std::vector<double>& h) const {
        if (_n) gpu_check(ibpm_memcpy_d2h(_ctx, h.data(), _p, sizeof(double) * _n));
    }

private:
    ibpm_ctx* _ctx;
    size_t _n;
    void* _p = nullptr;
};

// 2-D array of a single-level solve, indices (1..n1, 1..n2), the second
// contiguous: the reference's Array2<double>(nx-1, ny-1, 1, 1).
class Array2d : public std::vector<double> {
public:
    Array2d(int n1, int n2) : std::vector<double>((size_t)n1 * n2, 0.0), _n1(n1), _n2(n2) {}
    double& operator()(int i, int j) { return (*this)[(size_t)(i - 1) * _n2 + (j - 1)]; }
    double operator()(int i, int j) const { return (*this)[(size_t)(i - 1) * _n2 + (j - 1)]; }
    Array2d& operator=(double v) { assign(size(), v); return *this; }
    int n1() const { return _n1; }
    int n2() const { return _n2; }

private:
    int _n1, _n2;
};

// Boundary values of one level, reference src/BC.h:22-79: the ring of
// 2(nx+ny) values, left(j) j = 0..ny, top(i) i = 1..nx, right(j) from
// j = ny-1 down, bottom(i) from i = nx-1 down (bottom(0) = left(0)).
class BC : public std::vector<double> {
public:
    BC(int nx, int ny) : std::vector<double>(2 * (size_t)(nx + ny), 0.0), _nx(nx), _ny(ny) {}
    BC& operator=(double v) { assign(size(), v); return *this; }
    double& left(int j) { return (*this)[j]; }
    double& top(int i) { return (*this)[_ny + i]; }
    double& right(int j) { return (*this)[2 * _ny + _nx - j]; }
    double& bottom(int i) { return i == 0 ? (*this)[0] : (*this)[2 * (_nx + _ny) - i]; }

private:
    int _nx, _ny;
};

// Nodal field on every level, reference src/Scalar.h (layout Scalar.cc:81-89).
class Scalar : public std::vector<double> {
public:
    explicit Scalar(const Grid& g)
        : std::vector<double>((size_t)g.Ngrid() * (g.Nx() - 1) * (g.Ny() - 1), 0.0), _grid(g) {}
    double& operator()(int lev, int i, int j) { return (*this)[index(lev, i, j)]; }
    double operator()(int lev, int i, int j) const { return (*this)[index(lev, i, j)]; }
    Scalar& operator=(double v) { assign(size(), v); return *this; }
    Scalar& operator+=(const Scalar& o) { for (size_t k = 0; k < size(); ++k) (*this)[k] += o[k]; return *this; }
    Scalar& operator-=(const Scalar& o) { for (size_t k = 0; k < size(); ++k) (*this)[k] -= o[k]; return *this; }
    Scalar& operator*=(double a) { for (double& v : *this) v *= a; return *this; }
    const Grid& getGrid() const { return _grid; }
    size_t index(int lev, int i, int j) const {
        return ((size_t)lev * (_grid.Nx() - 1) + (i - 1)) * (_grid.Ny() - 1) + (j - 1);
    }

private:
    Grid _grid;
};
inline Scalar operator-(const Scalar& a, const Scalar& b) { Scalar c = a; c -= b; return c; }
inline Scalar operator*(double s, const Scalar& a) { Scalar c = a; c *= s; return c; }

// Edge fluxes on every level, reference src/Flux.h (layout Flux.cc:42-50):
// per level the X fluxes (i = 0..nx, j = 0..ny-1), then the Y fluxes
// (i = 0..nx-1, j = 0..ny).
class Flux : public std::vector<double> {
public:
    enum Direction { X = 0, Y = 1 };
    explicit Flux(const Grid& g)
        : std::vector<double>((size_t)g.Ngrid() * (2 * g.Nx() * g.Ny() + g.Nx() + g.Ny()), 0.0), _grid(g) {}
    double& operator()(int dir, int lev, int i, int j) {
        const int nx = _grid.Nx(), ny = _grid.Ny();
        const size_t base = (size_t)lev * (2 * nx * ny + nx + ny);
        return dir == X ? (*this)[base + (size_t)i * ny + j] : (*this)[base + (size_t)(nx + 1) * ny + (size_t)i * (ny + 1) + j];
    }
    Flux& operator=(double v) { assign(size(), v); return *this; }
    Flux& operator-=(const Flux& o) { for (size_t k = 0; k < size(); ++k) (*this)[k] -= o[k]; return *this; }
    const Grid& getGrid() const { return _grid; }

private:
    Grid _grid;
};

// Per-point 2-vectors, reference src/BoundaryVector.h: [X_0.., Y_0..].
class BoundaryVector : public std::vector<double> {
public:
    enum Direction { X = 0, Y = 1 };
    explicit BoundaryVector(int numPoints = 0) : std::vector<double>(2 * (size_t)numPoints, 0.0) {}
    double& operator()(int dir, int i) { return (*this)[(size_t)dir * (size() / 2) + i]; }
    double operator()(int dir, int i) const { return (*this)[(size_t)dir * (size() / 2) + i]; }
    BoundaryVector& operator=(double v) { assign(size(), v); return *this; }
    BoundaryVector& operator-=(const BoundaryVector& o) { for (size_t k = 0; k < size(); ++k) (*this)[k] -= o[k]; return *this; }
    int getNumPoints() const { return (int)(size() / 2); }
};

class Geometry {   // reference src/Geometry.h (.geom text is parsed by the library)
public:
    Geometry() = default;
    explicit Geometry(const std::string& filename) { load(filename); }
    bool load(const std::string& filename) {
        std::ifstream in(filename);
        if (!in.good()) return false;
        std::stringstream ss;
        ss << in.rdbuf();
        _text = ss.str();
        _n = ibpm_geom_num_points(_text.c_str());
        return _n >= 0;
    }
    bool loadText(const std::string& text) {
        _text = text;
        _n = ibpm_geom_num_points(_text.c_str());
        return _n >= 0;
    }
    int getNumPoints() const { return _n; }
    int getNumBodies() const {
        int nb = 0;
        ibpm_geom_info(_text.c_str(), nullptr, &nb, nullptr, nullptr, nullptr);
        return nb;
    }
    // Geometry::isStationary: no body with a time-dependent motion, or the
    // motion has been transferred to an unsteady base flow
    bool isStationary() const {
        int st = 1;
        ibpm_geom_info(_text.c_str(), nullptr, nullptr, &st, nullptr, nullptr);
        return st == 1 || _transferred;
    }
    // Geometry::transferMotion / transferCenter (Geometry.cc:46-77): the
    // library moves the first body's motion to the base flow itself when the
    // solver is created with an unsteady base flow; these record the request.
    Motion* transferMotion() {
        _transferred = true;
        return new Motion();
    }
    void transferCenter(double& x, double& y) const {
        ibpm_geom_info(_text.c_str(), nullptr, nullptr, nullptr, &x, &y);
    }
    bool motionTransferred() const { return _transferred; }
    const std::string& text() const { return _text; }
    // Geometry::moveBodies / getPoints / getVelocities (Geometry.cc:115-128):
    // the points and velocities at the time of the last moveBodies call
    // (motions evaluated at that time by the library).
    void moveBodies(double time) { _time = time; }
    double time() const { return _time; }
    BoundaryVector getPoints() const { return pointsAt(false); }
    BoundaryVector getVelocities() const { return pointsAt(true); }

private:
    BoundaryVector pointsAt(bool vel) const {
        BoundaryVector xy(_n), uv(_n);
        if (_n && ibpm_geom_points(_text.c_str(), 1, &_time, xy.data(), uv.data(), 2L * _n) != _n)
            throw std::runtime_error("ibpm_gpu: the geometry does not parse");
        return vel ? uv : xy;
    }
    std::string _text;
    int _n = 0;
    bool _transferred = false;
    double _time = 0.0;
};

// Multi-domain-free solves of one level, reference src/EllipticSolver2d.h:26-99:
// solve(f, u) with zero boundary values, solve(f, bc, u) with the given ones.
class EllipticSolver2d {
public:
    virtual ~EllipticSolver2d() = default;
    void solve(const Array2d& f, Array2d& u) const { run(f, nullptr, u); }
    void solve(const Array2d& f, const BC& bc, Array2d& u) const { run(f, &bc, u); }

protected:
    EllipticSolver2d(int nx, int ny, double dx, int kind, double alpha)
        : _nx(nx), _ny(ny), _dx(dx), _kind(kind), _alpha(alpha) {}
    int _nx, _ny;
    double _dx;

private:
    void run(const Array2d& f, const BC* bc, Array2d& u) const {
        ibpm_ctx* c = default_context();
        DeviceBuffer df(f, c), du(u.size(), c);
        std::unique_ptr<DeviceBuffer> db(bc ? new DeviceBuffer(*bc, c) : nullptr);
        gpu_check(ibpm_elliptic2d_solve(c, _kind, _nx, _ny, _dx, _alpha, df.get(), db ? db->get() : nullptr,
                                        du.get()));
        du.to(u);
    }
    int _kind;
    double _alpha;
};

// L u = f (reference PoissonSolver2d, EllipticSolver2d.h:66-77)
class PoissonSolver2d : public EllipticSolver2d {
public:
    PoissonSolver2d(int nx, int ny, double dx) : EllipticSolver2d(nx, ny, dx, 0, 0.0) {}
};

// (1 + alpha L) u = f (reference HelmholtzSolver2d, EllipticSolver2d.h:85-97)
class HelmholtzSolver2d : public EllipticSolver2d {
public:
    HelmholtzSolver2d(int nx, int ny, double dx, double alpha) : EllipticSolver2d(nx, ny, dx, 1, alpha) {}
};

// Multi-domain solve with zero boundary values on the outer domain,
// reference src/EllipticSolver.h:26-76 (coarsify, coarsest level, finer
// levels with boundary values from the coarser solution).
class EllipticSolver {
public:
    virtual ~EllipticSolver() = default;
    void solve(const Scalar& f, Scalar& u) const {
        ibpm_ctx* c = default_context();
        DeviceBuffer df(f, c), du(u.size(), c);
        gpu_check(ibpm_elliptic_solve(c, _kind, _grid.Nx(), _grid.Ny(), _grid.Ngrid(), _grid.Dx(), _alpha,
                                      df.get(), du.get()));
        du.to(u);
    }
    Scalar solve(const Scalar& f) const {
        Scalar u(_grid);
        solve(f, u);
        return u;
    }

protected:
    EllipticSolver(const Grid& grid, int kind, double alpha) : _grid(grid), _kind(kind), _alpha(alpha) {}

private:
    Grid _grid;
    int _kind;
    double _alpha;
};

class PoissonSolver : public EllipticSolver {   // reference EllipticSolver.h:54-62
public:
    explicit PoissonSolver(const Grid& grid) : EllipticSolver(grid, 0, 0.0) {}
};

class HelmholtzSolver : public EllipticSolver {   // reference EllipticSolver.h:66-76
public:
    HelmholtzSolver(const Grid& grid, double alpha) : EllipticSolver(grid, 1, alpha) {}
};

// Regularizer of a geometry's points on the finest level, reference
// src/Regularizer.h:50-91: update() for the geometry's current points,
// toFlux = E^T (smear, times dx), toBoundary = E (interpolate, over dx).
class Regularizer {
public:
    Regularizer(const Grid& grid, const Geometry& geometry) : _grid(grid), _geom(geometry) {}
    ~Regularizer() {
        if (_reg) ibpm_reg_destroy(_reg);
    }
    Regularizer(const Regularizer&) = delete;
    Regularizer& operator=(const Regularizer&) = delete;
    void update() {
        const int n = _geom.getNumPoints();
        if (n == 0) return;
        BoundaryVector xy = _geom.getPoints();
        if (!_reg)
            gpu_check(ibpm_reg_create(default_context(), _grid.Nx(), _grid.Ny(), _grid.Ngrid(), _grid.length(),
                                      _grid.xOffset(), _grid.yOffset(), n, xy.data(), &_reg));
        else
            gpu_check(ibpm_reg_update(_reg, xy.data()));
    }
    Flux toFlux(const BoundaryVector& u) const {
        Flux q(_grid);
        if (!_reg) return q;
        DeviceBuffer du(u), dq(q.size());
        gpu_check(ibpm_reg_to_flux(_reg, du.get(), dq.get()));
        dq.to(q);
        return q;
    }
    BoundaryVector toBoundary(const Flux& q) const {
        BoundaryVector f(_geom.getNumPoints());
        if (!_reg) return f;
        DeviceBuffer dq(q), df(f.size());
        gpu_check(ibpm_reg_to_boundary(_reg, dq.get(), df.get()));
        df.to(f);
        return f;
    }

private:
    Grid _grid;
    const Geometry& _geom;
    ibpm_reg* _reg = nullptr;
};

class BaseFlow {   // reference src/BaseFlow.h: uniform flow of magnitude mag at angle alpha
public:
    BaseFlow(const Grid&, double mag, double alpha) : _mag(mag), _alpha(alpha) {}
    double getAlpha() const { return _alpha; }
    double getMag() const { return _mag; }
    // BaseFlow::setMotion / setCenter (BaseFlow.cc:100-109): the flow becomes
    // unsteady; its angle for the force rotation then comes from the solver
    // (NonlinearIBSolver::baseFlowAngle)
    void setMotion(const Motion&) { _unsteady = true; }
    void setCenter(double x, double y) { _xc = x; _yc = y; }
    bool isStationary() const { return !_unsteady; }

private:
    double _mag, _alpha;
    bool _unsteady = false;
    double _xc = 0, _yc = 0;
};

struct Scheme {   // reference src/Scheme.h:12
    enum SchemeType { EULER, AB2, RK3, RK3b };
};

class State {      // reference src/State.h
public:
    State(const Grid& grid, int numPoints)
        : omega(grid), q(grid), f(numPoints), _dx(grid.Dx()), _grid(grid) {}
    // Restart file in the reference layout, src/State.cc:65-223.
    bool save(const std::string& filename) const {
        return ibpm_state_save(filename.c_str(), _grid.Nx(), _grid.Ny(), _grid.Ngrid(), _grid.Dx(),
                               _grid.getXEdge(0, 0), _grid.getYEdge(0, 0), (int)(f.size() / 2), q.data(),
                               omega.data(), f.data(), timestep, time) == 0;
    }
    bool load(const std::string& filename) {
        int nx, ny, ng, np;
        double dx, x0, y0;
        if (ibpm_state_load(filename.c_str(), &nx, &ny, &ng, &dx, &x0, &y0, &np, nullptr, nullptr, nullptr,
                            nullptr, nullptr) != 0)
            return false;
        if (nx != _grid.Nx() || ny != _grid.Ny() || ng != _grid.Ngrid() || dx != _grid.Dx() ||
            x0 != _grid.getXEdge(0, 0) || y0 != _grid.getYEdge(0, 0) || 2 * (size_t)np != f.size())
            return false;   // the reference resizes and reports failure (State.cc:94-117)
        return ibpm_state_load(filename.c_str(), &nx, &ny, &ng, &dx, &x0, &y0, &np, q.data(), omega.data(),
                               f.data(), &timestep, &time) == 0;
    }
    Scalar omega;
    Flux q;
    BoundaryVector f;
    int timestep = 0;
    double time = 0.;
    // State::computeNetForce, reference src/State.cc:53-63
    void computeNetForce(double& xforce, double& yforce) const {
        xforce = 0;
        yforce = 0;
        const size_t n = f.size() / 2;
        for (size_t i = 0; i < n; ++i) { xforce += f[i]; yforce += f[n + i]; }
        xforce *= _dx * _dx;
        yforce *= _dx * _dx;
    }

private:
    double _dx;
    Grid _grid;
};

// reference src/NavierStokesModel.h:40-106 (nonlinear model).  init() builds
// the operators (regulariser of the geometry's points at the geometry's
// time) over ibpm_model_*; the whole-step solver below does not need it.
class NavierStokesModel {
public:
    NavierStokesModel(const Grid& grid, const Geometry& geom, double Reynolds, const BaseFlow& q0)
        : _grid(grid), _geom(geom), _Re(Reynolds), _q0(q0) {}
    NavierStokesModel(const Grid& grid, const Geometry& geom, double Reynolds)
        : _grid(grid), _geom(geom), _Re(Reynolds), _q0(grid, 0.0, 0.0) {}
    ~NavierStokesModel() {
        if (_m) ibpm_model_destroy(_m);
    }
    NavierStokesModel(const NavierStokesModel&) = delete;
    NavierStokesModel& operator=(const NavierStokesModel&) = delete;
    void init() {
        if (_m) return;
        ibpm_flow_params p{};
        p.nx = _grid.Nx(); p.ny = _grid.Ny(); p.ngrid = _grid.Ngrid();
        p.length = _grid.length(); p.xoffset = _grid.xOffset(); p.yoffset = _grid.yOffset();
        p.reynolds = _Re;
        p.alpha_deg = _q0.getAlpha() * 180. / (4. * std::atan(1.));
        p.geom = _geom.text().c_str();
        gpu_check(ibpm_model_create(default_context(), &p, _geom.time(), &_m));
        gpu_check(ibpm_model_set_base_flow(_m, _q0.getMag(), _q0.getAlpha()));
    }
    int getNumPoints() const { return _geom.getNumPoints(); }
    // NavierStokesModel::updateOperators (NavierStokesModel.cc:93-99)
    void updateOperators(double time) {
        init();
        if (geTimeDependent()) gpu_check(ibpm_model_update_operators(_m, time));
    }
    // omega = B(f) = Curl(E^T f) (NavierStokesModel.cc:101-105)
    void B(const BoundaryVector& f, Scalar& omega) const {
        DeviceBuffer df(f), dw(omega.size());
        gpu_check(ibpm_model_B(handle(), df.get(), dw.get()));
        dw.to(omega);
    }
    // f = C(omega) = E Curl(psi), L psi = -omega (NavierStokesModel.cc:107-112)
    void C(const Scalar& omega, BoundaryVector& f) const {
        DeviceBuffer dw(omega), df(f.size());
        gpu_check(ibpm_model_C(handle(), dw.get(), df.get()));
        df.to(f);
    }
    // q = Curl(psi) + q0 (NavierStokesModel.cc:121-125)
    void computeFlux(const Scalar& omega, Flux& q) const {
        DeviceBuffer dw(omega), dq(q.size());
        gpu_check(ibpm_model_compute_flux(handle(), dw.get(), dq.get()));
        dq.to(q);
    }
    void refreshState(State& x) const { computeFlux(x.omega, x.q); }   // NavierStokesModel.cc:127-129
    // L psi = -omega with zero boundary values (NavierStokesModel.cc:133-140)
    Scalar vorticityToStreamfunction(const Scalar& omega) const {
        Scalar psi = -1. * omega;
        PoissonSolver(_grid).solve(psi, psi);
        return psi;
    }
    // body velocities minus the base flow at the points (NavierStokesModel.cc:86-91)
    BoundaryVector getConstraints() const {
        BoundaryVector b(getNumPoints());
        gpu_check(ibpm_model_constraints(handle(), b.data()));
        return b;
    }
    double getAlpha() const { return 1. / _Re; }
    bool geTimeDependent() const { return !_geom.isStationary(); }          // NavierStokesModel.cc:65-69
    bool bfTimeDependent() const { return !_q0.isStationary(); }            // NavierStokesModel.cc:71-75
    bool isTimeDependent() const { return geTimeDependent() || bfTimeDependent(); }
    const Grid& getGrid() const { return _grid; }
    const Geometry& geometry() const { return _geom; }
    const BaseFlow& getBaseFlow() const { return _q0; }
    double Reynolds() const { return _Re; }

    ibpm_model* handle() const {
        if (!_m) throw std::runtime_error("ibpm_gpu: NavierStokesModel::init() not called");
        return _m;
    }

private:
    const Grid& _grid;
    const Geometry& _geom;
    double _Re;
    BaseFlow _q0;
    ibpm_model* _m = nullptr;
};

// Projection step of the fractional-step method, reference
// src/ProjectionSolver.h:54-126 and ProjectionSolver.cc:57-101:
//   A omega* = a,  M f = C omega* - b,  omega = omega* - A^-1 B f,
// with A = 1 - beta/(2 Re) L, B = beta model.B, C = model.C and
// M = C A^-1 B.  Minv is the subclass's (Cholesky or CG).
class ProjectionSolver {
public:
    ProjectionSolver(const Grid& grid, const NavierStokesModel& model, double beta)
        : _beta(beta), _grid(grid), _model(model), _helmholtz(grid, -beta / 2. * model.getAlpha()) {}
    virtual ~ProjectionSolver() = default;
    virtual void init() {}
    virtual bool save(const std::string&) { return false; }
    virtual bool load(const std::string&) { return false; }
    void solve(const Scalar& a, const BoundaryVector& b, Scalar& omega, BoundaryVector& f) {
        Scalar omegaStar(a.getGrid());
        Ainv(a, omegaStar);
        BoundaryVector rhs(f.getNumPoints());
        C(omegaStar, rhs);
        rhs -= b;
        Minv(rhs, f);
        Scalar c(a.getGrid());
        B(f, c);
        Ainv(c, c);
        omega = omegaStar - c;
    }

protected:
    void Ainv(const Scalar& x, Scalar& y) { _helmholtz.solve(x, y); }
    void B(const BoundaryVector& f, Scalar& y) {
        _model.B(f, y);
        y *= _beta;
    }
    void C(const Scalar& x, BoundaryVector& y) { _model.C(x, y); }
    // y = M f on the GPU in one call (ibpm_model_M: the same B, A^-1, C)
    void M(const BoundaryVector& f, BoundaryVector& y) {
        if (f.empty()) return;
        DeviceBuffer df(f), dy(y.size());
        gpu_check(ibpm_model_M(_model.handle(), _beta, df.get(), dy.get()));
        dy.to(y);
    }
    BoundaryVector M(const BoundaryVector& f) {
        BoundaryVector y(f.getNumPoints());
        M(f, y);
        return y;
    }
    virtual void Minv(const BoundaryVector& b, BoundaryVector& x) = 0;
    double beta() const { return _beta; }
    const NavierStokesModel& model() const { return _model; }

private:
    double _beta;
    Grid _grid;
    const NavierStokesModel& _model;
    HelmholtzSolver _helmholtz;
};

// M^-1 by the FP64 Cholesky factor of M, built column by column on the GPU
// (reference src/CholeskySolver.h:36-78, CholeskySolver.cc:49-224); the
// factor is saved to and loaded from <basename>.cholesky in the reference's
// layout.
class CholeskySolver : public ProjectionSolver {
public:
    CholeskySolver(const Grid& grid, const NavierStokesModel& model, double beta)
        : ProjectionSolver(grid, model, beta), _size(2 * model.getNumPoints()),
          _alphaBeta(model.getAlpha() * beta), _lower((size_t)_size * _size), _diag(_size) {}
    void init() override {
        if (_init) return;
        std::vector<double> Mmat((size_t)_size * _size);
        if (_size) gpu_check(ibpm_cholesky_build(model().handle(), beta(), Mmat.data()));
        gpu_check(ibpm_cholesky_factor(_size, Mmat.data(), _lower.data(), _diag.data()));
        _init = true;
    }
    bool load(const std::string& basename) override {
        if (ibpm_cholesky_load((basename + ".cholesky").c_str(), _size, _alphaBeta, _lower.data(), _diag.data()) != 0)
            return false;
        _init = true;
        return true;
    }
    bool save(const std::string& basename) override {
        return _init && ibpm_cholesky_save((basename + ".cholesky").c_str(), _size, _alphaBeta, _lower.data(),
                                           _diag.data()) == 0;
    }

protected:
    void Minv(const BoundaryVector& b, BoundaryVector& x) override {
        if (!_init) throw std::runtime_error("ibpm_gpu: CholeskySolver::init() not called");
        gpu_check(ibpm_cholesky_solve(_size, _lower.data(), _diag.data(), b.data(), x.data()));
    }

private:
    int _size;
    double _alphaBeta;
    std::vector<double> _lower, _diag;
    bool _init = false;
};

// M^-1 by conjugate gradients from the previous x, absolute tolerance on
// the residual (reference src/ConjugateGradientSolver.h:28-57,
// ConjugateGradientSolver.cc:35-65), each M application on the GPU.
class ConjugateGradientSolver : public ProjectionSolver {
public:
    ConjugateGradientSolver(const Grid& grid, const NavierStokesModel& model, double beta, double tolerance)
        : ProjectionSolver(grid, model, beta), _tol(tolerance) {}
    void setTolerance(double tolerance) { _tol = tolerance; }
    double getTolerance() const { return _tol; }
    int lastIterations() const { return _its; }

protected:
    void Minv(const BoundaryVector& b, BoundaryVector& x) override {
        int conv = 0;
        if (b.empty()) { _its = 0; return; }
        BoundaryVector x0 = x;
        gpu_check(ibpm_cg_solve(model().handle(), beta(), _tol, 0, b.data(), x0.data(), x.data(), &_its, &conv));
    }

private:
    double _tol;
    int _its = 0;
};

// IBSolver (reference src/IBSolver.h:20-63): the timestepper of one flow on
// the GPU (a batch of one).  Cholesky for stationary bodies and CG for moving
// ones, chosen as in IBSolver.cc:122-135.  The model's N term is the
// subclass's: NonlinearIBSolver, LinearizedIBSolver, AdjointIBSolver,
// LinearizedPeriodicIBSolver, SFDSolver (IBSolver.h:69-222).
class IBSolver {
public:
    virtual ~IBSolver() {
        if (_batch) ibpm_batch_destroy(_batch);
        if (_ctx) ibpm_ctx_destroy(_ctx);
    }
    IBSolver(const IBSolver&) = delete;
    IBSolver& operator=(const IBSolver&) = delete;

    std::string getName() const {
        static const char* names[] = {"Explicit Euler", "Adams Bashforth", "3rd-order Runge Kutta (3-step)",
                                      "3rd-order Runge Kutta (4-step)"};
        return names[_scheme];
    }
    double getTimestep() const { return _dt; }
    void setTol(double tol) { _tol = tol; }   // before init(); CG absolute residual tolerance
    // Force solve of moving bodies (no counterpart in the reference): mode 0
    // plain CG, 1 CG preconditioned by a stale FP64 inverse of M with `pairs`
    // low-rank pairs (ibpm_batch_set_precond).  Before init().
    void setPrecond(int mode, int pairs) { _pc_mode = mode; _pc_pairs = pairs; }
    // P built from M (levels 0), from M's finest-level form (1), from the
    // kernel-table M (2) or from that table refreshed before every solve (3),
    // with a build's cost in M applications per unknown for the rebuild rule
    // (ibpm_batch_set_precond_build).  Before init().
    void setPrecondBuild(int levels, double rebuild_cost) {
        _pc_levels = levels;
        _pc_cost = rebuild_cost;
    }
    // A^-1 inside the moving-body force solve by local sweeps where they equal
    // the DST solve to rounding (true, the default) or always by the DST
    // solve (ibpm_batch_set_local_helmholtz).  Before init().
    void setLocalHelmholtz(bool on) { _local_helm = on; }
    // IBSolver::load/save (src/IBSolver.cc:77-108): the Cholesky factors live
    // in <basename>_0i.cholesky.  load() only names the cache and returns false;
    // init() then reads matching files or builds the factors and writes them,
    // which is what the reference's load-else-init-and-save sequence does.
    bool load(const std::string& basename) {
        _cache = basename;
        return false;
    }
    bool save(const std::string&) { return true; }   // written by init()
    int factorsLoaded() const { return _batch ? ibpm_batch_cholesky_loaded(_batch, 0) : 0; }
    void init() {
        if (_batch) return;
        gpu_check(ibpm_ctx_create(_device, &_ctx));
        ibpm_flow_params p{};
        p.nx = _grid.Nx(); p.ny = _grid.Ny(); p.ngrid = _grid.Ngrid();
        p.length = _grid.length(); p.xoffset = _grid.xOffset(); p.yoffset = _grid.yOffset();
        p.reynolds = _model.Reynolds(); p.dt = _dt; p.scheme = (int)_scheme;
        p.alpha_deg = _model.getBaseFlow().getAlpha() * 180. / (4. * std::atan(1.));
        p.cg_tol = _tol; p.force_cg = 0;
        p.geom = _model.geometry().text().c_str();
        p.ubf = _model.bfTimeDependent() ? 1 : 0;
        p.model = _kind;
        p.period = (int)std::max<size_t>(1, _base.size());
        p.sfd_chi = _chi;
        p.sfd_delta = _Delta;
        gpu_check(ibpm_batch_create(_ctx, 1, &p, &_batch));
        for (size_t k = 0; k < _base.size(); ++k)
            gpu_check(ibpm_batch_set_base_state(_batch, 0, (int)k, _base[k].omega.data(), _base[k].q.data()));
        if (!_cache.empty() && !_model.geTimeDependent())
            gpu_check(ibpm_batch_cholesky_cache(_batch, 0, _cache.c_str()));
        gpu_check(ibpm_batch_init(_batch));
        if (!_model.geTimeDependent()) gpu_check(ibpm_batch_use_graph(_batch, 1));
        else if (_pc_mode) gpu_check(ibpm_batch_set_precond(_batch, _pc_mode, _pc_pairs));
        if (!_local_helm) gpu_check(ibpm_batch_set_local_helmholtz(_batch, 0));
        if (_pc_mode && _model.geTimeDependent() && (_local_helm || _pc_levels != 1))
            gpu_check(ibpm_batch_set_precond_build(_batch, _pc_levels, _pc_cost));
    }
    // IBSolver::advance, reference src/IBSolver.cc:142-150
    void advance(State& x) {
        init();
        gpu_check(ibpm_batch_advance(_batch, 1));
        x.time += _dt;
        ++x.timestep;
        gpu_check(ibpm_batch_get_f(_batch, 0, x.f.data()));
    }
    void advance(State& x, int nsteps) {
        init();
        gpu_check(ibpm_batch_advance(_batch, nsteps));
        x.time += _dt * nsteps;
        x.timestep += nsteps;
        gpu_check(ibpm_batch_get_f(_batch, 0, x.f.data()));
    }
    // initial condition or restart: vorticity, force, time and step, then
    // refreshState (reference src/ibpm.cc:250-305)
    void setState(const State& x) {
        init();
        gpu_check(ibpm_batch_set_state(_batch, 0, x.omega.data(), x.f.data(), x.time, x.timestep));
    }
    void getState(State& x) {
        gpu_check(ibpm_batch_get_omega(_batch, 0, x.omega.data()));
        gpu_check(ibpm_batch_get_q(_batch, 0, x.q.data()));
        gpu_check(ibpm_batch_get_f(_batch, 0, x.f.data()));
    }
    // The angle (radians) that rotates the net force into drag and lift after
    // the last step: BaseFlow::setAlphaMag(x.time); getAlpha() of ibpm.cc:347-351
    // for an unsteady base flow, else the free-stream angle.
    double baseFlowAngle() const {
        double al = _model.getBaseFlow().getAlpha();
        if (_batch) ibpm_batch_base_flow(_batch, 0, &al, nullptr);
        return al;
    }
    int cgIterations(int substep) const {
        int it[4] = {-1, -1, -1, -1};
        ibpm_batch_cg_iters(_batch, 0, it, 4);
        return it[substep & 3];
    }

protected:
    IBSolver(const Grid& grid, NavierStokesModel& model, double dt, Scheme::SchemeType scheme, double tol,
             int device, int kind)
        : _grid(grid), _model(model), _dt(dt), _scheme(scheme), _tol(tol), _device(device), _kind(kind) {}
    const Grid& _grid;
    NavierStokesModel& _model;
    double _dt;
    Scheme::SchemeType _scheme;
    double _tol;
    int _device;
    int _kind;                      // ibpm_flow_params::model
    std::vector<State> _base;       // base states of the linear models
    double _chi = 0.02, _Delta = 15.0;
    int _pc_mode = 0, _pc_pairs = 0, _pc_levels = 0;
    double _pc_cost = 1.0;
    bool _local_helm = true;
    std::string _cache;
    ibpm_ctx* _ctx = nullptr;
    ibpm_batch* _batch = nullptr;
};

// N = Curl(q x omega) (IBSolver.cc:203-207)
class NonlinearIBSolver : public IBSolver {
public:
    NonlinearIBSolver(const Grid& grid, NavierStokesModel& model, double dt, Scheme::SchemeType scheme,
                      double tol = 1e-7, int device = 0)
        : IBSolver(grid, model, dt, scheme, tol, device, 0) {}
};

// N = Curl(q0 x omega + q x omega0) about the base state x0 (IBSolver.cc:209-214)
class LinearizedIBSolver : public IBSolver {
public:
    LinearizedIBSolver(const Grid& grid, NavierStokesModel& model, double dt, Scheme::SchemeType scheme,
                       const State& x0, double tol = 1e-7, int device = 0)
        : IBSolver(grid, model, dt, scheme, tol, device, 1) {
        _base.push_back(x0);
    }
};

// N = Laplacian(q0 x q) - Curl(q x omega0) (IBSolver.cc:216-220)
class AdjointIBSolver : public IBSolver {
public:
    AdjointIBSolver(const Grid& grid, NavierStokesModel& model, double dt, Scheme::SchemeType scheme,
                    const State& x0, double tol = 1e-7, int device = 0)
        : IBSolver(grid, model, dt, scheme, tol, device, 2) {
        _base.push_back(x0);
    }
};

// The linearized N about base state x0[timestep % period] (IBSolver.cc:222-229)
class LinearizedPeriodicIBSolver : public IBSolver {
public:
    LinearizedPeriodicIBSolver(const Grid& grid, NavierStokesModel& model, double dt, Scheme::SchemeType scheme,
                               const std::vector<State>& x0, int period, double tol = 1e-7, int device = 0)
        : IBSolver(grid, model, dt, scheme, tol, device, 3) {
        _base.assign(x0.begin(), x0.begin() + period);
    }
};

// Selective frequency damping: N = Curl(q x omega) - chi (omega - omega_hat),
// with the filtered state omega_hat advanced by the same scheme
// (IBSolver.cc:236-311, IBSolver.h:189-222)
class SFDSolver : public IBSolver {
public:
    SFDSolver(const Grid& grid, NavierStokesModel& model, double dt, Scheme::SchemeType scheme, double Delta,
              double chi, double tol = 1e-7, int device = 0)
        : IBSolver(grid, model, dt, scheme, tol, device, 4) {
        _Delta = Delta;
        _chi = chi;
    }
    // the filtered vorticity in and out (loadFilteredState / saveFilteredState)
    void getFilteredState(Scalar& omega_hat) { gpu_check(ibpm_batch_get_filtered(_batch, 0, omega_hat.data())); }
    void setFilteredState(const Scalar& omega_hat) {
        init();
        gpu_check(ibpm_batch_set_filtered(_batch, 0, omega_hat.data()));
    }
};

}  // namespace ibpm
