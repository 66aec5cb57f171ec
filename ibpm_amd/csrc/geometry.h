// geometry.h -- host-side bodies, motion and regulariser association lists.
#pragma once
#include <string>
#include <vector>

namespace ibpmg {

struct Body {
    std::vector<double> rx, ry;   // reference points
    double xc = 0, yc = 0;        // centre of rotation
    int motion = 0;               // 0 none, 1 fixed, 2 pitch/plunge
    double mp[6] = {0, 0, 0, 0, 0, 0};
};

// Geometry / RigidBody (reference src/Geometry.cc:143-191, src/RigidBody.cc:248-529).
struct Geometry {
    std::vector<Body> bodies;
    int npoints = 0;
    bool stationary = true;       // no body with a time-dependent motion
    bool parse(const std::string& text, std::string& err);
    // Geometry::moveBodies(t) -> points (x, y) and velocities (u, v), in body
    // order then point order (Geometry.cc:79-93, RigidBody.cc:586-609).
    void move(double t, std::vector<double>& x, std::vector<double>& y,
              std::vector<double>& u, std::vector<double>& v) const;
};

// Regulariser association lists for one set of points (Regularizer.cc:54-90)
// rearranged for gathers:
//   E   : per boundary component k (0..2nb-1), level-0 flux indices ascending
//   E^T : per touched flux, components ascending
//   nodes touched by Curl(E^T f) with the flux-list slots of their 4 fluxes
struct RegLists {
    std::vector<int> e_ptr, e_f;
    std::vector<double> e_w;
    std::vector<int> flux_of_slot;            // flux index of each flux-list slot
    std::vector<int> t_ptr, t_k;
    std::vector<double> t_w;
    std::vector<int> node_idx;                // level-0 scalar index (0-based)
    std::vector<int> node_nbr;                // 4 slots per node (-1: untouched)
};

void build_reg_lists(int nx, int ny, double h, double xoff, double yoff,
                     const std::vector<double>& px, const std::vector<double>& py, RegLists& out);

}  // namespace ibpmg
