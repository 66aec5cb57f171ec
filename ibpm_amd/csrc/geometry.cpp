// geometry.cpp -- .geom parsing, rigid-body motion and regulariser lists.
#include "geometry.h"

#include <fstream>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <map>
#include <sstream>

namespace ibpmg {

static std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

// The command set of RigidBody::load (RigidBody.cc:248-529) that the
// projection-step configurations use: circle, circle_n, line, line_n,
// line_aoa, point, center, name and every motion of RigidBody.cc:351-492.
bool Geometry::parse(const std::string& text, std::string& err) {
    bodies.clear();
    std::istringstream in(text);
    std::string line;
    Body* cur = nullptr;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string cmd;
        if (!(ls >> cmd)) continue;
        cmd = lower(cmd);
        if (cmd[0] == '#') continue;
        if (!cur) {
            if (cmd == "body") {
                bodies.emplace_back();
                cur = &bodies.back();
            } else if (cmd == "end") {
                break;
            } else {
                err = "geometry: could not parse line: " + line;
                return false;
            }
            continue;
        }
        if (cmd == "end") {
            cur = nullptr;
        } else if (cmd == "center") {
            if (!(ls >> cur->xc >> cur->yc)) { err = "geometry: bad center: " + line; return false; }
        } else if (cmd == "circle" || cmd == "circle_n") {   // RigidBody.cc:92-118
            double xc, yc, r, d;
            int n;
            if (cmd == "circle") {
                if (!(ls >> xc >> yc >> r >> d)) { err = "geometry: bad circle: " + line; return false; }
                const double dTheta = d / r;
                const double twopi = 8. * std::atan(1.);
                n = (int)std::floor(twopi / dTheta + 1 + 0.5);
            } else if (!(ls >> xc >> yc >> r >> n)) {
                err = "geometry: bad circle_n: " + line;
                return false;
            }
            const double twopi = 8. * std::atan(1.);
            const double dTheta = twopi / n;
            for (int i = 0; i < n; ++i) {
                double s, c;
                sin_cos(i * dTheta, s, c);
                cur->rx.push_back(xc + r * c);
                cur->ry.push_back(yc + r * s);
            }
        } else if (cmd == "line" || cmd == "line_n") {       // RigidBody.cc:120-146
            double x1, y1, x2, y2, d;
            int n;
            if (cmd == "line") {
                if (!(ls >> x1 >> y1 >> x2 >> y2 >> d)) { err = "geometry: bad line: " + line; return false; }
                const double len = std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
                n = (int)std::floor(len / d + 1 + 0.5);
            } else if (!(ls >> x1 >> y1 >> x2 >> y2 >> n)) {
                err = "geometry: bad line_n: " + line;
                return false;
            }
            const double ddx = (x2 - x1) / (n - 1), ddy = (y2 - y1) / (n - 1);
            for (int i = 0; i < n; ++i) {
                cur->rx.push_back(x1 + i * ddx);
                cur->ry.push_back(y1 + i * ddy);
            }
        } else if (cmd == "line_aoa") {                       // RigidBody.cc:148-171
            double l, xC, yC, aoa;
            int n;
            if (!(ls >> l >> xC >> yC >> aoa >> n)) { err = "geometry: bad line_aoa: " + line; return false; }
            const double pi = 3.141592653589793238462643383279502884197169399375;
            const double a = aoa / 180 * pi;
            double sa, ca;
            sin_cos(a, sa, ca);
            const double del = l / (n - 1);
            for (int i = 0; i < n; ++i) {
                const double x0r = i * del - xC, y0r = 0. - yC;
                cur->rx.push_back(x0r * ca - y0r * sa + xC);
                cur->ry.push_back(x0r * sa + y0r * ca + yC);
            }
            cur->xc = xC;
            cur->yc = yC;
        } else if (cmd == "point") {
            double x, y;
            if (!(ls >> x >> y)) { err = "geometry: bad point: " + line; return false; }
            cur->rx.push_back(x);
            cur->ry.push_back(y);
        } else if (cmd == "name") {
        } else if (cmd == "motion") {
            std::string mt;
            ls >> mt;
            mt = lower(mt);
            auto need = [&](int k, const char* what) {
                for (int q = 0; q < k; ++q)
                    if (!(ls >> cur->mp[q])) {
                        err = std::string("geometry: ") + what;
                        return false;
                    }
                return true;
            };
            if (mt == "fixed") {                               // FixedPosition.h, RigidBody.cc:353-361
                if (!need(3, "bad fixed motion")) return false;
                cur->motion = M_FIXED;
            } else if (mt == "fixedvel") {                     // FixedVelocity.h:26-50, RigidBody.cc:363-372
                if (!need(3, "fixedvel needs xdot ydot thetadot")) return false;
                cur->motion = M_FIXEDVEL;
            } else if (mt == "pitchplunge") {                  // PitchPlunge.h:30-58, RigidBody.cc:373-385
                double a[6];
                if (!(ls >> a[0] >> a[1] >> a[2] >> a[3] >> a[4] >> a[5])) {
                    err = "geometry: pitchplunge needs 6 values (RigidBody.cc:373-385)";
                    return false;
                }
                const double twopi = 8. * std::atan(1.);
                cur->motion = M_PITCHPLUNGE;
                cur->mp[0] = a[0]; cur->mp[1] = a[1] * twopi; cur->mp[2] = a[2];
                cur->mp[3] = a[3]; cur->mp[4] = a[4] * twopi; cur->mp[5] = a[5];
            } else if (mt == "sigmoidalstep") {                // SigmoidalStep.h:29-52, RigidBody.cc:386-395
                if (!need(3, "sigmoidalstep needs AMP DUR t0")) return false;
                cur->motion = M_SIGMOIDALSTEP;
            } else if (mt == "lagstep1" || mt == "lagstep2") { // LagStep1.h, LagStep2.h:27-63, RigidBody.cc:396-417
                if (!need(4, "lagstep needs AMP PW TAU T0")) return false;
                cur->motion = mt == "lagstep1" ? M_LAGSTEP1 : M_LAGSTEP2;
            } else if (mt == "eldredge" || mt == "eldredge1" || mt == "eldredge2") {
                // EldredgeManeuver.h:27-62, Eldredge1.h:26-57, Eldredge2.h:27-72, RigidBody.cc:418-475
                if (!need(6, "eldredge needs AMP a t1 t2 t3 t4")) return false;
                cur->motion = mt == "eldredge" ? M_ELDREDGE : mt == "eldredge1" ? M_ELDREDGE1 : M_ELDREDGE2;
                const double a = cur->mp[1], t1 = cur->mp[2], t2 = cur->mp[3], t3 = cur->mp[4], t4 = cur->mp[5];
                const double tt = (t2 + t3) / 2.;
                cur->mp[6] = std::log((std::cosh(a * (tt - t1)) * std::cosh(a * (tt - t4))) /
                                      (std::cosh(a * (tt - t2)) * std::cosh(a * (tt - t3))));   // maxG
            } else if (mt == "eldredgecombined2") {            // EldredgeCombined2.h:27-88, RigidBody.cc:431-449
                if (!need(12, "eldredgecombined2 needs 12 values")) return false;
                cur->motion = M_ELDREDGECOMBINED2;
                const double* m = cur->mp;
                const double a = m[1], b = m[7];
                const double tta = (m[3] + m[4]) / 2., ttb = (m[9] + m[10]) / 2.;
                const double maxa = std::log((std::cosh(a * (tta - m[2])) * std::cosh(a * (tta - m[5]))) /
                                             (std::cosh(a * (tta - m[3])) * std::cosh(a * (tta - m[4]))));
                // the reference's second factor uses a, not b (EldredgeCombined2.h:59)
                const double maxb = std::log((std::cosh(b * (ttb - m[8])) * std::cosh(a * (ttb - m[11]))) /
                                             (std::cosh(b * (ttb - m[9])) * std::cosh(b * (ttb - m[10]))));
                cur->mp[12] = maxa;
                cur->mp[13] = maxb;
            } else if (mt == "motionfile" || mt == "motionfileperiodic") {
                // MotionFile.h:57-116, MotionFilePeriodic.h:56-131, RigidBody.cc:476-492
                std::string fname;
                if (!(ls >> fname)) { err = "geometry: motion file name missing"; return false; }
                if (mt == "motionfileperiodic" && !(ls >> cur->mp[0])) {
                    err = "geometry: motionfileperiodic needs a period";
                    return false;
                }
                std::ifstream mf(fname);
                int nrows = 0;
                if (!(mf >> nrows) || nrows < 0) { err = "geometry: cannot read motion file " + fname; return false; }
                double tlast = -1.e+5;
                for (int r = 0; r < nrows; ++r) {
                    double row[7];
                    for (double& v : row)
                        if (!(mf >> v)) { err = "geometry: motion file " + fname + " formatted incorrectly"; return false; }
                    if (row[0] < tlast) { err = "geometry: motion file times must increase"; return false; }
                    tlast = row[0];
                    cur->table.insert(cur->table.end(), row, row + 7);
                }
                cur->motion = mt == "motionfile" ? M_MOTIONFILE : M_MOTIONFILEPERIODIC;
            } else {
                err = "geometry: unsupported motion '" + mt + "'";
                return false;
            }
        } else {
            err = "geometry: unsupported command '" + cmd + "'";
            return false;
        }
    }
    npoints = 0;
    stationary = true;
    for (auto& b : bodies) {
        npoints += (int)b.rx.size();
        if (b.moving()) stationary = false;
    }
    return true;
}

void body_transformation(const Body& B, double t, double& X, double& Y, double& th, double& xd, double& yd,
                         double& thd) {
    X = Y = th = xd = yd = thd = 0;
    const double* m = B.mp;
    const double pi = 4. * std::atan(1.);
    switch (B.motion) {
    case M_FIXED: X = m[0]; Y = m[1]; th = m[2]; break;
    case M_PITCHPLUNGE: {   // PitchPlunge.h:52-58
        double sp, cp, sq, cq;
        sin_cos(m[4] * t + m[5], sp, cp);
        sin_cos(m[1] * t + m[2], sq, cq);
        Y = m[3] * sp;
        yd = m[3] * m[4] * cp;
        th = m[0] * sq;
        thd = m[0] * m[1] * cq;
        break;
    }
    case M_FIXEDVEL: {    // FixedVelocity.h:42-50: positions integrated over successive calls
        double dtt = t - B.oldtime;
        if (dtt > 1) dtt = 0.;
        B.acc[0] = B.acc[0] + m[0] * dtt;
        B.acc[1] = B.acc[1] + m[1] * dtt;
        B.acc[2] = B.acc[2] + m[2] * dtt;
        B.oldtime = t;
        X = B.acc[0]; Y = B.acc[1]; th = B.acc[2]; xd = m[0]; yd = m[1]; thd = m[2];
        break;
    }
    case M_SIGMOIDALSTEP: {   // SigmoidalStep.h:41-52 (AMP, DUR, t0)
        const double tdiff = t - m[2];
        const double arg = ((12. * tdiff / m[1]) - 6.);
        const double coeff = 12. / (m[1] * std::sqrt(pi));
        const double nerf = std::erf(arg);
        th = (1. / 2) * m[0] * (pi / 180) * (1. + nerf);
        thd = m[0] * (pi / 180) * coeff * std::exp(-arg * arg);
        break;
    }
    case M_LAGSTEP1:
    case M_LAGSTEP2: {        // LagStep1.h:41-63, LagStep2.h:41-63 (AMP, PW, TAU, T0)
        const double AMPrad = m[0] * (pi / 180), PW = m[1], TAU = m[2];
        const double tdiff = t - m[3];
        double a = 0., ad = 0.;
        if (tdiff <= 0) {
        } else if (B.motion == M_LAGSTEP1) {
            if (tdiff <= PW) {
                a = tdiff + TAU * std::exp(-tdiff / TAU) - TAU;
                ad = 1 - std::exp(-tdiff / TAU);
            } else {
                a = PW + TAU * std::exp(-tdiff / TAU) - TAU * std::exp(-(tdiff - PW) / TAU);
                ad = -1. * std::exp(-tdiff / TAU) + std::exp(-(tdiff - PW) / TAU);
            }
        } else {
            if (tdiff <= PW) {
                a = tdiff + (2 * TAU + tdiff) * std::exp(-tdiff / TAU) - 2 * TAU;
                ad = 1. - (1. + tdiff / TAU) * std::exp(-tdiff / TAU);
            } else {
                a = PW + (2 * TAU + tdiff) * std::exp(-tdiff / TAU) + (PW - 2 * TAU) * std::exp(-(tdiff - PW) / TAU) -
                    tdiff * std::exp(-(tdiff - PW) / TAU);
                ad = -(1. + tdiff / TAU) * std::exp(-tdiff / TAU) + (1. + (tdiff - PW) / TAU) * std::exp(-(tdiff - PW) / TAU);
            }
        }
        th = a * AMPrad / PW;
        thd = ad * AMPrad / PW;
        break;
    }
    case M_ELDREDGE:
    case M_ELDREDGE1:
    case M_ELDREDGE2: {       // EldredgeManeuver.h:48-62, Eldredge1.h:46-57, Eldredge2.h:52-72
        const double a = m[1], maxG = m[6];
        const double t1d = t - m[2], t2d = t - m[3], t3d = t - m[4], t4d = t - m[5];
        const double Garg = ((std::cosh(a * t1d) * std::cosh(a * t4d)) / (std::cosh(a * t2d) * std::cosh(a * t3d)));
        const double G = std::log(Garg);
        const double dGdt = a * (std::tanh(a * t1d) - std::tanh(a * t2d) - std::tanh(a * t3d) + std::tanh(a * t4d));
        if (B.motion == M_ELDREDGE) {
            const double AMPrad = m[0] * (pi / 180);
            th = G * AMPrad / maxG;
            thd = dGdt * AMPrad / maxG;
        } else if (B.motion == M_ELDREDGE1) {
            Y = G * m[0] / maxG;
            yd = dGdt * m[0] / maxG;
        } else {   // plunge from the running integral of G over successive calls
            const double dtt = t - B.oldtime;
            B.acc[0] = B.acc[0] + G * dtt;
            B.oldtime = t;
            Y = B.acc[0] * m[0] / maxG;
            yd = G * m[0] / maxG;
        }
        break;
    }
    case M_ELDREDGECOMBINED2: {   // EldredgeCombined2.h:64-88
        const double AMPrad = m[6] * (pi / 180);
        const double a = m[1], b = m[7];
        const double a1d = t - m[2], a2d = t - m[3], a3d = t - m[4], a4d = t - m[5];
        const double b1d = t - m[8], b2d = t - m[9], b3d = t - m[10], b4d = t - m[11];
        const double Garga = ((std::cosh(a * a1d) * std::cosh(a * a4d)) / (std::cosh(a * a2d) * std::cosh(a * a3d)));
        const double Gargb = ((std::cosh(b * b1d) * std::cosh(b * b4d)) / (std::cosh(b * b2d) * std::cosh(b * b3d)));
        const double Ga = std::log(Garga), Gb = std::log(Gargb);
        const double dGdtb = b * (std::tanh(b * b1d) - std::tanh(b * b2d) - std::tanh(b * b3d) + std::tanh(b * b4d));
        const double dtt = t - B.oldtime;
        B.acc[0] = B.acc[0] + Ga * dtt;
        B.oldtime = t;
        Y = B.acc[0] * m[0] / m[12];
        th = Gb * AMPrad / m[13];
        yd = Ga * m[0] / m[12];
        thd = dGdtb * AMPrad / m[13];
        break;
    }
    case M_MOTIONFILE:
    case M_MOTIONFILEPERIODIC: {  // MotionFile.h:90-116, MotionFilePeriodic.h:91-131
        const double tt = (B.motion == M_MOTIONFILEPERIODIC) ? std::fmod(t, m[0]) : t;
        const double* d = B.table.data();
        const int n = (int)(B.table.size() / 7);
        int index = -1;
        for (int i = 0; i < n - 1; i++)
            if ((tt >= d[7 * i]) && (tt < d[7 * (i + 1)])) index = i;
        if (index != -1) {
            const double* p = d + 7 * index;
            const double* q = p + 7;
            const double tdiff = q[0] - p[0];
            const double alpha = (tt - p[0]) / tdiff;
            X = alpha * q[1] + (1 - alpha) * p[1];
            Y = alpha * q[2] + (1 - alpha) * p[2];
            th = alpha * q[3] + (1 - alpha) * p[3];
            xd = alpha * q[4] + (1 - alpha) * p[4];
            yd = alpha * q[5] + (1 - alpha) * p[5];
            thd = alpha * q[6] + (1 - alpha) * p[6];
        }
        break;
    }
    default: break;
    }
}

void Geometry::move(double t, std::vector<double>& x, std::vector<double>& y,
                    std::vector<double>& u, std::vector<double>& v) const {
    x.clear(); y.clear(); u.clear(); v.clear();
    for (const auto& B : bodies) {
        if (B.motion == M_NONE) {   // no motion: RigidBody::moveBody returns early
            for (size_t k = 0; k < B.rx.size(); ++k) {
                x.push_back(B.rx[k]); y.push_back(B.ry[k]);
                u.push_back(0.0); v.push_back(0.0);
            }
            continue;
        }
        double X, Y, th, xd, yd, thd;
        body_transformation(B, t, X, Y, th, xd, yd, thd);
        double st, ct;   // TangentSE2.h:83-107
        sin_cos(th, st, ct);
        for (size_t k = 0; k < B.rx.size(); ++k) {
            const double a = B.rx[k] - B.xc, b = B.ry[k] - B.yc;
            x.push_back((X + a * ct - b * st) + B.xc);
            y.push_back((Y + a * st + b * ct) + B.yc);
            u.push_back(-a * st * thd - b * ct * thd + xd);
            v.push_back(a * ct * thd - b * st * thd + yd);
        }
    }
}

void BaseFlowMotion::set_alpha_mag(double t) {   // BaseFlow.cc:111-125
    double x, y, th, xd, yd, thd;
    body_transformation(motion, t, x, y, th, xd, yd, thd);
    double sb, cb;
    sin_cos(alpha_bf, sb, cb);
    const double xdot_bf = mag_bf * cb, ydot_bf = mag_bf * sb;
    const double xdot_t = xd - xdot_bf, ydot_t = yd - ydot_bf;
    const double gamma = std::atan2(ydot_t, -1. * xdot_t);
    alpha = -1. * th - gamma;
    mag = std::sqrt(xdot_t * xdot_t + ydot_t * ydot_t);
}

void BaseFlowMotion::move(double t) {   // BaseFlow.cc:127-150
    double x, y, th, xd, yd, thd;
    body_transformation(motion, t, x, y, th, xd, yd, thd);
    set_alpha_mag(t);
    double sa, ca;
    sin_cos(alpha, sa, ca);
    u = mag * ca;
    v = mag * sa;
    w = -1. * thd;
}

// Roma-Peskin-Berger discrete delta, Regularizer.cc:37-49
static double delta_fn(double r) {
    if (r > 1.5) return 0;
    if (r <= 0.5) return (1 + std::sqrt(1 - 3 * r * r)) / 3;
    return (5 - 3 * r - std::sqrt(1 - 3 * (1 - r) * (1 - r))) / 6;
}

void build_reg_lists(int nx, int ny, double h, double xoff, double yoff,
                     const std::vector<double>& px, const std::vector<double>& py, RegLists& out) {
    const int nb = (int)px.size();
    const long nfx = (long)(nx + 1) * ny;
    out.e_ptr.assign(2 * nb + 1, 0);
    out.e_f.clear();
    out.e_w.clear();
    out.e_sep.assign((size_t)2 * nb * kESep, 0.0);
    // (flux, component, weight) in the order the reference visits them
    struct Hit {
        int f, k;
        double w;
    };
    std::vector<Hit> hits;
    hits.reserve((size_t)nb * 2 * 16);
    for (int dir = 0; dir <= 1; ++dir) {
        for (int p = 0; p < nb; ++p) {
            const int k = dir * nb + p;
            // Window of +-3 cells around the point, scanned in ascending flux
            // index with the reference's own inclusion test, so the members
            // and their order equal the brute-force scan of every flux.
            const int ci = (int)std::floor((px[p] - xoff) / h);
            const int cj = (int)std::floor((py[p] - yoff) / h);
            const int imax = dir == 0 ? nx : nx - 1, jmax = dir == 0 ? ny - 1 : ny;
            const int i0 = std::max(0, ci - 3), i1 = std::min(imax, ci + 3);
            const int j0 = std::max(0, cj - 3), j1 = std::min(jmax, cj + 3);
            for (int i = i0; i <= i1; ++i) {
                for (int j = j0; j <= j1; ++j) {
                    double fx, fy;
                    int ind;
                    if (dir == 0) { fx = xoff + i * h; fy = yoff + (j + 0.5) * h; ind = i * ny + j; }
                    else { fx = xoff + (i + 0.5) * h; fy = yoff + j * h; ind = (int)(nfx + (long)i * (ny + 1) + j); }
                    const double ddx = std::fabs(fx - px[p]) / h;
                    const double ddy = std::fabs(fy - py[p]) / h;
                    if (ddx < 1.5 && ddy < 1.5) {
                        const double w = delta_fn(ddx) * delta_fn(ddy);
                        out.e_f.push_back(ind);
                        out.e_w.push_back(w);
                        hits.push_back({ind, k, w});
                    }
                }
            }
            out.e_ptr[k + 1] = (int)out.e_f.size();
            // the same members as a rectangle of rows and columns (the
            // inclusion test is separable) with the row and column factors of
            // the weights: at most 3 of each within 1.5 cells
            double* sep = &out.e_sep[(size_t)k * kESep];
            int ni = 0, nj = 0;
            for (int i = i0; i <= i1; ++i) {
                const double fx = dir == 0 ? xoff + i * h : xoff + (i + 0.5) * h;
                const double ddx = std::fabs(fx - px[p]) / h;
                if (ddx < 1.5 && ni < 3) {
                    if (ni == 0) sep[0] = i;
                    sep[4 + ni++] = delta_fn(ddx);
                }
            }
            for (int j = j0; j <= j1; ++j) {
                const double fy = dir == 0 ? yoff + (j + 0.5) * h : yoff + j * h;
                const double ddy = std::fabs(fy - py[p]) / h;
                if (ddy < 1.5 && nj < 3) {
                    if (nj == 0) sep[1] = j;
                    sep[7 + nj++] = delta_fn(ddy);
                }
            }
            sep[2] = ni;
            sep[3] = nj;
        }
    }
    // E^T: per flux in ascending order, components in visiting (= ascending) order
    std::stable_sort(hits.begin(), hits.end(), [](const Hit& a, const Hit& b) { return a.f < b.f; });
    out.flux_of_slot.clear();
    out.t_ptr.assign(1, 0);
    out.t_k.clear();
    out.t_w.clear();
    for (size_t q = 0; q < hits.size();) {
        const int f = hits[q].f;
        out.flux_of_slot.push_back(f);
        for (; q < hits.size() && hits[q].f == f; ++q) {
            out.t_k.push_back(hits[q].k);
            out.t_w.push_back(hits[q].w);
        }
        out.t_ptr.push_back((int)out.t_k.size());
    }
    // nodes adjacent to a touched flux, ascending; curl slots Y(i,j), Y(i-1,j), X(i,j-1), X(i,j)
    const std::vector<int>& fs = out.flux_of_slot;
    auto slot = [&](long f) {
        auto it = std::lower_bound(fs.begin(), fs.end(), (int)f);
        return (it != fs.end() && *it == f) ? (int)(it - fs.begin()) : -1;
    };
    std::vector<int> keys;
    keys.reserve(fs.size() * 2);
    auto touch = [&](int i, int j) {
        if (i < 1 || i > nx - 1 || j < 1 || j > ny - 1) return;
        keys.push_back((i - 1) * (ny - 1) + (j - 1));
    };
    for (int f : fs) {
        if (f < nfx) {
            const int i = f / ny, j = f % ny;
            touch(i, j);
            touch(i, j + 1);
        } else {
            const int u = (int)(f - nfx);
            const int i = u / (ny + 1), j = u % (ny + 1);
            touch(i, j);
            touch(i + 1, j);
        }
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    out.node_idx.assign(keys.begin(), keys.end());
    out.node_nbr.clear();
    out.node_nbr.reserve(4 * keys.size());
    for (int key : keys) {
        const int i = key / (ny - 1) + 1, j = key % (ny - 1) + 1;
        out.node_nbr.push_back(slot(nfx + (long)i * (ny + 1) + j));
        out.node_nbr.push_back(slot(nfx + (long)(i - 1) * (ny + 1) + j));
        out.node_nbr.push_back(slot((long)i * ny + j - 1));
        out.node_nbr.push_back(slot((long)i * ny + j));
    }
}

}  // namespace ibpmg
