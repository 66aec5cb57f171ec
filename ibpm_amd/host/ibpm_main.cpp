// ibpm_main.cpp -- the `ibpm` command-line driver on the GPU path.
//
// Accepts the reference's flags and defaults (reference src/ibpm.cc:41-88)
// for the nonlinear model and writes the force file in the reference format
// "%5d %.5e %.5e %.5e" = step, time, Cd, Cl (src/OutputForce.cc:46-64), plus
// <name>.force17 with full precision.  -restart n writes the reference's
// binary restart file <name>%05d.bin every n steps and -ic loads one
// (src/State.cc:65-223, src/ibpm.cc:250-274).  Added flag: -tol (CG tolerance,
// which the reference sets only from code, src/ibpm.cc:292-294) and -precond
// (force solve of moving bodies: none = the reference's plain CG, dense (P
// alone), stale (its former name), or lowrank; all end on the same residual
// test; -pcbuild finest builds P from M's finest-level form, table from the
// kernel-table M, tracked from that table refreshed before every solve).  -ubf 1 (unsteady base flow,
// src/ibpm.cc:166-178) solves a single moving body in its own frame and
// rotates the forces by the base-flow angle after each step (:347-351).
// -model linear|adjoint|linearperiodic|sfd with -baseflow, -period,
// -periodstart, -pbaseflowname, -chi, -Delta, -subbaseflow and -resettime
// selects the other timesteppers as the reference does (src/ibpm.cc:100-235,
// 250-280); SFD writes its filtered vorticity beside each restart file
// (<restart>.bin_xhat, IBSolver.cc:281-287).
// Tecplot and energy outputs are not produced by this driver.
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "ibpm_gpu.hpp"

using namespace ibpm;

namespace {
struct Args {   // ParmParser-style "-name value" pairs (reference src/ParmParser.cc:22-200)
    std::map<std::string, std::string> kv;
    std::set<std::string> used;
    bool help = false;
    Args(int argc, char** argv) {
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            if (a.size() < 2 || a[0] != '-') { kv["?" + a] = ""; continue; }
            a = a.substr(1);
            if (a == "h") { help = true; continue; }
            if (i + 1 < argc) kv[a] = argv[++i];
            else kv[a] = "";
        }
    }
    std::string str(const std::string& k, const std::string& d) {
        used.insert(k);
        auto it = kv.find(k);
        return it == kv.end() ? d : it->second;
    }
    int i(const std::string& k, int d) { return std::atoi(str(k, std::to_string(d)).c_str()); }
    double f(const std::string& k, double d) {
        used.insert(k);
        auto it = kv.find(k);
        return it == kv.end() ? d : std::atof(it->second.c_str());
    }
    bool valid() const {
        for (auto& p : kv)
            if (!used.count(p.first)) {
                std::cerr << "unrecognised argument: -" << p.first << "\n";
                return false;
            }
        return true;
    }
};
}  // namespace

int main(int argc, char** argv) {
    std::cout << "Immersed Boundary Projection Method (IBPM), MI355X projection-step path\n" << std::endl;
    Args P(argc, argv);
    const std::string name = P.str("name", "ibpm");
    std::string outdir = P.str("outdir", ".");
    P.i("tecplot", 100);
    const int iRestart = P.i("restart", 100);
    const std::string icFile = P.str("ic", "");
    const std::string numDigit = P.str("numdigfilename", "%05d");
    const int iForce = P.i("force", 1);
    const int nx = P.i("nx", 200), ny = P.i("ny", 200), ngrid = P.i("ngrid", 1);
    const double length = P.f("length", 4.0), xOffset = P.f("xoffset", -2.), yOffset = P.f("yoffset", -2.);
    const double alpha_deg = P.f("alpha", 0.);
    const std::string geomFile = P.str("geom", name + ".geom");
    const double Re = P.f("Re", 100.);
    const std::string model = P.str("model", "nonlinear");
    const double dt = P.f("dt", 0.02);
    const int numSteps = P.i("nsteps", 250);
    const std::string scheme = P.str("scheme", "rk3");
    const double tol = P.f("tol", 1e-7);
    const int device = P.i("device", 0);
    const std::string precond = P.str("precond", "none");
    const std::string helmholtz = P.str("helmholtz", "local");
    const std::string pcbuild = P.str("pcbuild", "all");
    const bool ubf = P.i("ubf", 0) != 0;
    const std::string baseFlow = P.str("baseflow", "");
    const bool resetTime = P.i("resettime", 0) != 0;
    const bool subtractBaseflow = P.i("subbaseflow", 0) != 0;
    const int period = P.i("period", 1);
    const int periodStart = P.i("periodstart", 0);
    const std::string periodBaseFlowName = P.str("pbaseflowname", "");
    const double chi = P.f("chi", 0.02);
    const double Delta = P.f("Delta", 15.);
    std::set<std::string> models = {"nonlinear", "linear", "adjoint", "linearperiodic", "sfd"};
    std::map<std::string, Scheme::SchemeType> schemes = {
        {"euler", Scheme::EULER}, {"ab2", Scheme::AB2}, {"rk3", Scheme::RK3}, {"rk3b", Scheme::RK3b}};
    std::map<std::string, std::pair<int, int>> preconds = {
        {"none", {0, 0}}, {"dense", {1, 0}}, {"stale", {1, 0}}, {"lowrank", {1, 8}}};
    const std::map<std::string, std::pair<int, double>> pcbuilds = {
        {"all", {0, 1.0}}, {"finest", {1, 0.45}}, {"table", {2, 0.1}}, {"tracked", {3, 0.1}}};
    if (!P.valid() || P.help || !schemes.count(scheme) || !models.count(model) || !preconds.count(precond) ||
        (helmholtz != "local" && helmholtz != "dst") || !pcbuilds.count(pcbuild)) {
        std::cerr << "usage: ibpm [-name ibpm] [-outdir .] [-force 1] [-nx 200] [-ny 200] [-ngrid 1]\n"
                     "            [-length 4] [-xoffset -2] [-yoffset -2] [-alpha 0] [-geom name.geom]\n"
                     "            [-Re 100] [-dt 0.02] [-nsteps 250] [-scheme euler|ab2|rk3|rk3b]\n"
                     "            [-model nonlinear|linear|adjoint|linearperiodic|sfd] [-baseflow file.bin]\n"
                     "            [-period 1] [-periodstart 0] [-pbaseflowname flow%05d.bin] [-chi 0.02]\n"
                     "            [-Delta 15] [-subbaseflow 0|1] [-resettime 0|1] [-ic file.bin]\n"
                     "            [-tol 1e-7] [-device 0] [-precond none|dense|stale|lowrank]\n"
                     "            [-ubf 0|1] [-helmholtz local|dst] [-pcbuild all|finest|table|tracked]\n";
        return 1;
    }
    const bool linearModel = model == "linear" || model == "adjoint" || model == "linearperiodic";
    if (linearModel && model != "linearperiodic" && baseFlow.empty()) {   // ibpm.cc:100-121
        std::cout << "ERROR: for linear or adjoint models, must specify a base flow" << std::endl;
        return 1;
    }
    if (model == "linearperiodic" && periodBaseFlowName.empty()) {
        std::cout << "ERROR: for linear periodic model, must specify a periodic base flow" << std::endl;
        return 1;
    }
    if (outdir.empty() || outdir.back() != '/') outdir += "/";
    mkdir(outdir.c_str(), 0777);

    Grid grid(nx, ny, ngrid, length, xOffset, yOffset);
    Geometry geom;
    std::cout << "Reading geometry from file " << geomFile << std::endl;
    if (!geom.load(geomFile)) {
        std::cerr << "could not load geometry " << geomFile << ": " << ibpm_last_error() << std::endl;
        return -1;
    }
    std::cout << "    " << geom.getNumPoints() << " points on the boundary\n" << std::endl;
    const double pi = 4. * std::atan(1.);
    BaseFlow q_potential(grid, 1.0, alpha_deg * pi / 180.);
    // unsteady base flow: only for a single moving body (ibpm.cc:166-178)
    if (!geom.isStationary() && geom.getNumBodies() == 1 && ubf) {
        Motion* m = geom.transferMotion();
        double xC = 0, yC = 0;
        geom.transferCenter(xC, yC);
        q_potential.setMotion(*m);
        q_potential.setCenter(xC, yC);
        delete m;
    }
    if (ubf && geom.getNumBodies() != 1) {
        std::cout << "Unsteady base flow is only supported for a single moving body.  Exiting program." << std::endl;
        return 1;
    }
    // the model and timestepper of ibpm.cc:185-235 (the linear models have no potential flow)
    const int npts = geom.getNumPoints();
    std::unique_ptr<NavierStokesModel> nsm(linearModel ? new NavierStokesModel(grid, geom, Re)
                                                       : new NavierStokesModel(grid, geom, Re, q_potential));
    State x00(grid, npts);
    std::unique_ptr<IBSolver> solver;
    SFDSolver* sfd = nullptr;
    if (model == "linear" || model == "adjoint") {
        if (!x00.load(baseFlow)) {
            std::cout << "baseflow failed to load.  Exiting program." << std::endl;
            return 1;
        }
        if (model == "linear") solver.reset(new LinearizedIBSolver(grid, *nsm, dt, schemes[scheme], x00, tol, device));
        else solver.reset(new AdjointIBSolver(grid, *nsm, dt, schemes[scheme], x00, tol, device));
    } else if (model == "linearperiodic") {
        std::vector<State> x0(period, x00);
        for (int i = 0; i < period; ++i) {
            char fn[512];
            std::snprintf(fn, sizeof fn, periodBaseFlowName.c_str(), i + periodStart);
            x0[i].load(fn);
        }
        x00 = x0[0];
        solver.reset(new LinearizedPeriodicIBSolver(grid, *nsm, dt, schemes[scheme], x0, period, tol, device));
    } else if (model == "sfd") {
        std::cout << "SFD parameters:\n    chi =   " << chi << "\n    Delta = " << Delta << "\n" << std::endl;
        sfd = new SFDSolver(grid, *nsm, dt, schemes[scheme], Delta, chi, tol, device);
        solver.reset(sfd);
    } else {
        solver.reset(new NonlinearIBSolver(grid, *nsm, dt, schemes[scheme], tol, device));
    }
    solver->setPrecond(preconds[precond].first, preconds[precond].second);
    solver->setLocalHelmholtz(helmholtz == "local");
    if (pcbuild != "all") solver->setPrecondBuild(pcbuilds.at(pcbuild).first, pcbuilds.at(pcbuild).second);
    State x(grid, npts);
    nsm->init();
    try {
        // load the factors or compute and save them (reference ibpm.cc:291-296)
        if (!solver->load(outdir + name)) solver->init();
        if (solver->factorsLoaded() > 0)
            std::cout << "Loaded Cholesky factors from " << outdir << name << "_0*.cholesky" << std::endl;
        if (!icFile.empty()) {   // ibpm.cc:250-280
            std::cout << "Loading initial condition from file: " << icFile << std::endl;
            bool loaded = x.load(icFile);
            if (!loaded) std::cout << "    (failed: using zero initial condition)" << std::endl;
            if (subtractBaseflow) {
                if (model == "nonlinear") {
                    std::cout << "Flag subbaseflow should be true only for linear cases" << std::endl;
                    return 1;
                }
                x.q -= x00.q;
                x.omega -= x00.omega;
                x.f = 0.0;
            }
            if (resetTime) {
                x.timestep = 0;
                x.time = 0.;
            }
            if (loaded || subtractBaseflow) solver->setState(x);
            if (sfd) {   // SFDSolver::loadFilteredState (IBSolver.cc:289-311): zero unless the file loads
                State xh(grid, npts);
                xh.omega = 0.0;
                if (!xh.load(icFile + "_xhat")) std::cout << "  (failed: filtered state set to zero)" << std::endl;
                sfd->setFilteredState(xh.omega);
            }
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return -1;
    }
    IBSolver& solverRef = *solver;
    std::cout << "using " << solverRef.getName() << " timestepper\n    dt = " << dt << "\n" << std::endl;
    FILE* ff = iForce > 0 ? std::fopen((outdir + name + ".force").c_str(), "w") : nullptr;
    FILE* ff17 = iForce > 0 ? std::fopen((outdir + name + ".force17").c_str(), "w") : nullptr;
    double al = q_potential.getAlpha();
    std::cout << "Integrating for " << numSteps << " steps" << std::endl;
    for (int i = 1; i <= numSteps; ++i) {
        solverRef.advance(x);
        double xF, yF;
        x.computeNetForce(xF, yF);
        if (!q_potential.isStationary()) al = solverRef.baseFlowAngle();   // ibpm.cc:347-351
        double sal, cal;   // one sincos, as the reference's gcc build takes the pair
        ::sincos(al, &sal, &cal);
        const double drag = 2 * (xF * cal + yF * sal);
        const double lift = 2 * (xF * -1. * sal + yF * cal);
        std::printf("step %d    x force: %16.9g, y force: %16.9g\n", i, drag, lift);
        if (ff && x.timestep % iForce == 0) {
            std::fprintf(ff, "%5d %.5e %.5e %.5e\n", x.timestep, x.time, drag, lift);
            std::fprintf(ff17, "%5d %.17e %.17e %.17e\n", x.timestep, x.time, drag, lift);
        }
        if (iRestart > 0 && x.timestep % iRestart == 0) {
            solverRef.getState(x);
            char num[64];
            std::snprintf(num, sizeof num, numDigit.c_str(), x.timestep);
            if (!x.save(outdir + name + num + ".bin"))
                std::cerr << "could not write restart file: " << ibpm_last_error() << std::endl;
            if (sfd && chi != 0.0) {   // SFDSolver::saveFilteredState: the filtered vorticity
                State xh(grid, npts);
                sfd->getFilteredState(xh.omega);
                xh.timestep = x.timestep;
                xh.time = x.time;
                if (!xh.save(outdir + name + num + ".bin_xhat"))
                    std::cerr << "could not write filtered state: " << ibpm_last_error() << std::endl;
            }
        }
    }
    if (ff) std::fclose(ff);
    if (ff17) std::fclose(ff17);
    return 0;
}
