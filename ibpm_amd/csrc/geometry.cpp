// geometry.cpp -- .geom parsing, rigid-body motion and regulariser lists.
#include "geometry.h"

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
// line_aoa, point, center, name and the motions fixed / pitchplunge.
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
                cur->rx.push_back(xc + r * std::cos(i * dTheta));
                cur->ry.push_back(yc + r * std::sin(i * dTheta));
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
            const double a = aoa / 180 * pi, ca = std::cos(a), sa = std::sin(a);
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
            if (mt == "fixed") {                               // FixedPosition.h
                if (!(ls >> cur->mp[0] >> cur->mp[1] >> cur->mp[2])) { err = "geometry: bad fixed motion"; return false; }
                cur->motion = 1;
            } else if (mt == "pitchplunge") {                  // PitchPlunge.h:30-58
                double a[6];
                if (!(ls >> a[0] >> a[1] >> a[2] >> a[3] >> a[4] >> a[5])) {
                    err = "geometry: pitchplunge needs 6 values (RigidBody.cc:373-385)";
                    return false;
                }
                const double twopi = 8. * std::atan(1.);
                cur->motion = 2;
                cur->mp[0] = a[0]; cur->mp[1] = a[1] * twopi; cur->mp[2] = a[2];
                cur->mp[3] = a[3]; cur->mp[4] = a[4] * twopi; cur->mp[5] = a[5];
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
        if (b.motion == 2) stationary = false;
    }
    return true;
}

void Geometry::move(double t, std::vector<double>& x, std::vector<double>& y,
                    std::vector<double>& u, std::vector<double>& v) const {
    x.clear(); y.clear(); u.clear(); v.clear();
    for (const auto& B : bodies) {
        double X = 0, Y = 0, th = 0, xd = 0, yd = 0, thd = 0;
        if (B.motion == 1) {
            X = B.mp[0]; Y = B.mp[1]; th = B.mp[2];
        } else if (B.motion == 2) {   // PitchPlunge::getTransformation, PitchPlunge.h:52-58
            Y = B.mp[3] * std::sin(B.mp[4] * t + B.mp[5]);
            yd = B.mp[3] * B.mp[4] * std::cos(B.mp[4] * t + B.mp[5]);
            th = B.mp[0] * std::sin(B.mp[1] * t + B.mp[2]);
            thd = B.mp[0] * B.mp[1] * std::cos(B.mp[1] * t + B.mp[2]);
        }
        if (B.motion == 0) {   // no motion: RigidBody::moveBody returns early
            for (size_t k = 0; k < B.rx.size(); ++k) {
                x.push_back(B.rx[k]); y.push_back(B.ry[k]);
                u.push_back(0.0); v.push_back(0.0);
            }
            continue;
        }
        const double ct = std::cos(th), st = std::sin(th);   // TangentSE2.h:83-107
        for (size_t k = 0; k < B.rx.size(); ++k) {
            const double a = B.rx[k] - B.xc, b = B.ry[k] - B.yc;
            x.push_back((X + a * ct - b * st) + B.xc);
            y.push_back((Y + a * st + b * ct) + B.yc);
            u.push_back(-a * st * thd - b * ct * thd + xd);
            v.push_back(a * ct * thd - b * st * thd + yd);
        }
    }
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
    // per flux: (component, weight) in the order the reference visits them
    std::map<int, std::vector<std::pair<int, double>>> per_flux;
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
                        per_flux[ind].push_back({k, w});
                    }
                }
            }
            out.e_ptr[k + 1] = (int)out.e_f.size();
        }
    }
    out.flux_of_slot.clear();
    out.t_ptr.assign(1, 0);
    out.t_k.clear();
    out.t_w.clear();
    std::map<int, int> slot_of_flux;
    for (auto& kv : per_flux) {   // ascending flux index; components already ascending
        slot_of_flux[kv.first] = (int)out.flux_of_slot.size();
        out.flux_of_slot.push_back(kv.first);
        for (auto& cw : kv.second) { out.t_k.push_back(cw.first); out.t_w.push_back(cw.second); }
        out.t_ptr.push_back((int)out.t_k.size());
    }
    // nodes adjacent to a touched flux; curl slots Y(i,j), Y(i-1,j), X(i,j-1), X(i,j)
    std::map<int, std::vector<int>> nodes;
    auto slot = [&](int f) { auto it = slot_of_flux.find(f); return it == slot_of_flux.end() ? -1 : it->second; };
    auto touch = [&](int i, int j) {
        if (i < 1 || i > nx - 1 || j < 1 || j > ny - 1) return;
        const int key = (i - 1) * (ny - 1) + (j - 1);
        if (nodes.count(key)) return;
        nodes[key] = {slot((int)(nfx + (long)i * (ny + 1) + j)), slot((int)(nfx + (long)(i - 1) * (ny + 1) + j)),
                      slot(i * ny + j - 1), slot(i * ny + j)};
    };
    for (int f : out.flux_of_slot) {
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
    out.node_idx.clear();
    out.node_nbr.clear();
    for (auto& kv : nodes) {
        out.node_idx.push_back(kv.first);
        for (int s : kv.second) out.node_nbr.push_back(s);
    }
}

}  // namespace ibpmg
