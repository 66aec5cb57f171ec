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

private:
    std::string _text;
    int _n = 0;
    bool _transferred = false;
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
        : omega((size_t)grid.Ngrid() * (grid.Nx() - 1) * (grid.Ny() - 1)),
          q((size_t)grid.Ngrid() * (2 * grid.Nx() * grid.Ny() + grid.Nx() + grid.Ny())),
          f(2 * (size_t)numPoints), _dx(grid.Dx()), _grid(grid) {}
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
    std::vector<double> omega, q, f;
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

class NavierStokesModel {   // reference src/NavierStokesModel.h:40-106 (nonlinear model)
public:
    NavierStokesModel(const Grid& grid, const Geometry& geom, double Reynolds, const BaseFlow& q0)
        : _grid(grid), _geom(geom), _Re(Reynolds), _q0(q0) {}
    void init() {}
    double getAlpha() const { return 1. / _Re; }
    bool geTimeDependent() const { return !_geom.isStationary(); }          // NavierStokesModel.cc:65-69
    bool bfTimeDependent() const { return !_q0.isStationary(); }            // NavierStokesModel.cc:71-75
    bool isTimeDependent() const { return geTimeDependent() || bfTimeDependent(); }
    const Grid& getGrid() const { return _grid; }
    const Geometry& geometry() const { return _geom; }
    const BaseFlow& getBaseFlow() const { return _q0; }
    double Reynolds() const { return _Re; }

private:
    const Grid& _grid;
    const Geometry& _geom;
    double _Re;
    BaseFlow _q0;
};

// IBSolver with the nonlinear term (reference src/IBSolver.h:20-90).  Cholesky
// for stationary bodies and CG for moving ones, chosen as in IBSolver.cc:122-135.
class NonlinearIBSolver {
public:
    NonlinearIBSolver(const Grid& grid, NavierStokesModel& model, double dt, Scheme::SchemeType scheme,
                      double tol = 1e-7, int device = 0)
        : _grid(grid), _model(model), _dt(dt), _scheme(scheme), _tol(tol), _device(device) {}
    ~NonlinearIBSolver() {
        if (_batch) ibpm_batch_destroy(_batch);
        if (_ctx) ibpm_ctx_destroy(_ctx);
    }
    NonlinearIBSolver(const NonlinearIBSolver&) = delete;
    NonlinearIBSolver& operator=(const NonlinearIBSolver&) = delete;

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
        gpu_check(ibpm_batch_create(_ctx, 1, &p, &_batch));
        if (!_cache.empty() && !_model.geTimeDependent())
            gpu_check(ibpm_batch_cholesky_cache(_batch, 0, _cache.c_str()));
        gpu_check(ibpm_batch_init(_batch));
        if (!_model.geTimeDependent()) gpu_check(ibpm_batch_use_graph(_batch, 1));
        else if (_pc_mode) gpu_check(ibpm_batch_set_precond(_batch, _pc_mode, _pc_pairs));
        if (!_local_helm) gpu_check(ibpm_batch_set_local_helmholtz(_batch, 0));
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

private:
    const Grid& _grid;
    NavierStokesModel& _model;
    double _dt;
    Scheme::SchemeType _scheme;
    double _tol;
    int _device;
    int _pc_mode = 0, _pc_pairs = 0;
    bool _local_helm = true;
    std::string _cache;
    ibpm_ctx* _ctx = nullptr;
    ibpm_batch* _batch = nullptr;
};

}  // namespace ibpm
