// geometry.h -- host-side bodies, motion and regulariser association lists.
#pragma once
#include <cmath>
#include <string>
#include <vector>

namespace ibpmg {

// sin and cos of one argument from glibc's sincos.  The reference, compiled
// with gcc, gets the pair from one sincos call wherever both are taken of the
// same argument (gcc merges them; clang does not), and sincos differs from
// separate sin and cos by one ulp for about one argument in 700: taking the
// pair the same way keeps positions, velocities and the base flow bit-equal.
inline void sin_cos(double x, double& s, double& c) { ::sincos(x, &s, &c); }

// Motions of the .geom "motion" command (RigidBody.cc:351-492).
enum MotionKind {
    M_NONE = 0, M_FIXED, M_PITCHPLUNGE, M_FIXEDVEL, M_SIGMOIDALSTEP, M_LAGSTEP1, M_LAGSTEP2, M_ELDREDGE,
    M_ELDREDGE1, M_ELDREDGE2, M_ELDREDGECOMBINED2, M_MOTIONFILE, M_MOTIONFILEPERIODIC
};

struct Body {
    std::vector<double> rx, ry;   // reference points
    double xc = 0, yc = 0;        // centre of rotation
    int motion = M_NONE;
    double mp[14] = {};           // motion parameters (and derived constants)
    std::vector<double> table;    // motion files: rows of t, x, y, theta, dx, dy, dtheta
    // state of the motions that integrate over successive calls
    // (FixedVelocity, Eldredge2, EldredgeCombined2): running sums, last time
    mutable double acc[3] = {0, 0, 0};
    mutable double oldtime = 0;
    bool moving() const { return motion != M_NONE && motion != M_FIXED; }
};

// Motion::getTransformation(t) of one body's motion: position (X, Y, th) and
// velocity (xd, yd, thd) of the body frame.  Each formula is the reference's,
// term for term, so results agree bit for bit; motions that integrate update
// the body's state (FixedVelocity.h:42-50, Eldredge2.h:62-64).
void body_transformation(const Body& B, double t, double& X, double& Y, double& th, double& xd, double& yd,
                         double& thd);

// Geometry / RigidBody (reference src/Geometry.cc:143-191, src/RigidBody.cc:248-529).
struct Geometry {
    std::vector<Body> bodies;
    int npoints = 0;
    bool stationary = true;       // no body with a time-dependent motion
    bool parse(const std::string& text, std::string& err);
    // Geometry::moveBodies(t) -> points (x, y) and velocities (u, v), in body
    // order then point order (Geometry.cc:79-93, RigidBody.cc:586-609).  Like
    // the reference, motions that integrate keep state between calls, so the
    // sequence of times matters.
    void move(double t, std::vector<double>& x, std::vector<double>& y,
              std::vector<double>& u, std::vector<double>& v) const;
};

// Unsteady base flow (ibpm -ubf, ibpm.cc:166-178): the motion of the only
// body moves to the base flow, the body stays at its reference points, and
// the flow is solved in the body frame.  The base flux is the free stream
// seen by the body plus a rotation about its centre (BaseFlow.cc:111-150,
// Flux.cc:101-129): X fluxes (u - w (y - yc)) dx, Y fluxes (v + w (x - xc)) dx
// on every level.
struct BaseFlowMotion {
    Body motion;                 // the transferred motion (its points are unused)
    double xc = 0, yc = 0;       // Geometry::transferCenter
    double mag_bf = 1, alpha_bf = 0;   // free stream magnitude and angle (rad)
    double alpha = 0, mag = 1;   // BaseFlow::_alpha, _mag after the last update
    double u = 0, v = 0, w = 0;  // velocities of the base flux after moveFlow
    // BaseFlow::setAlphaMag(t): angle and magnitude of the relative free stream
    void set_alpha_mag(double t);
    // BaseFlow::moveFlow(t): set_alpha_mag, then u, v, w
    void move(double t);
    // Base flux through an edge at (x, y) of a level with spacing d; dir 0 is X
    double flux(int dir, double x, double y, double d) const {
        return dir == 0 ? (u - w * (y - yc)) * d : (v + w * (x - xc)) * d;
    }
};

// Regulariser association lists for one set of points (Regularizer.cc:54-90)
// rearranged for gathers:
//   E   : per boundary component k (0..2nb-1), level-0 flux indices ascending
//   E^T : per touched flux, components ascending
//   nodes touched by Curl(E^T f) with the flux-list slots of their 4 fluxes
// Per component of E, its row in separable form (kESep doubles): first row
// and column index of its nodes, their counts ni and nj (at most 3), the
// row factors u[ni] and the column factors v[nj]; the node (i0 + a, j0 + b)
// has weight u[a] v[b].
constexpr int kESep = 10;

struct RegLists {
    std::vector<int> e_ptr, e_f;
    std::vector<double> e_w;
    std::vector<double> e_sep;                // kESep per component: i0, j0, ni, nj, u[3], v[3]
    std::vector<int> flux_of_slot;            // flux index of each flux-list slot
    std::vector<int> t_ptr, t_k;
    std::vector<double> t_w;
    std::vector<int> node_idx;                // level-0 scalar index (0-based)
    std::vector<int> node_nbr;                // 4 slots per node (-1: untouched)
};

void build_reg_lists(int nx, int ny, double h, double xoff, double yoff,
                     const std::vector<double>& px, const std::vector<double>& py, RegLists& out);

}  // namespace ibpmg
