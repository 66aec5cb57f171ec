"""Body motions of the .geom "motion" command (RigidBody.cc:351-492): every
type, moved through the call sequence of an ibpm run (moveBodies at t0 twice,
ibpm.cc:285 and 304, then at t + c_n dt for each RK3 substep, IBSolver.cc:165).

  * the library's host geometry (ibpm_geom_points) equals the oracle bit for
    bit, and the oracle equals the reference's own Motion classes compiled
    from its sources (oracle/_ref) bit for bit, for whole sequences -- which
    also pins the state that FixedVelocity, Eldredge2 and EldredgeCombined2
    carry between calls and the restart guard of FixedVelocity (dtt > 1);
  * independently of the reference build, the oracle is pinned by what the
    formulas fix: velocities are the time derivatives of positions (central
    differences; one-sided for the motions that integrate), the maneuvers
    reach their amplitude where the closed form says, the lag steps settle
    at the amplitude, motion-file knots are reproduced, and a periodic file
    repeats with its period.
No GPU call is made here.
"""
import ctypes
import math
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ibpm_amd", "libibpm_gpu.so")
REFPART = os.path.join(ROOT, "oracle", "_ref", "libibpm_refpart.so")
D2R = math.pi / 180
XC = 0.1           # centre of rotation of the test plate
BODY = "body P\n line_n -0.5 0 0.5 0 21\n center 0.1 0.0\n motion {}\nend\n"


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def motion_table(tmp):
    """A motion file (MotionFile.h:57-81): n, then rows t x y theta dx dy dtheta."""
    t = np.linspace(0.0, 2.5, 11)
    rows = np.stack([t, 0.2 * np.sin(t), 0.1 * t * t, 0.3 * np.sin(2 * t),
                     0.2 * np.cos(t), 0.2 * t, 0.6 * np.cos(2 * t)], 1)
    path = os.path.join(str(tmp), "plate.motion")
    with open(path, "w") as f:
        f.write("%d\n" % len(rows))
        for r in rows:
            f.write(" ".join("%.17g" % v for v in r) + "\n")
    return path, rows


def specs(tmp):
    path, _ = motion_table(tmp)
    return {
        "none": None,
        "fixed": "fixed 0.1 0.05 0.3",
        "fixedvel": "fixedvel 0.3 -0.1 0.2",
        "pitchplunge": "pitchplunge 0.3 0.25 0.1 0.2 0.4 0.5",
        "sigmoidalstep": "sigmoidalstep 30 2.0 1.0",
        "lagstep1": "lagstep1 20 2.0 0.3 0.5",
        "lagstep2": "lagstep2 20 2.0 0.3 0.5",
        "eldredge": "eldredge 45 11 1 3 4 6",
        "eldredge1": "eldredge1 0.5 11 1 3 4 6",
        "eldredge2": "eldredge2 0.5 11 1 3 4 6",
        "eldredgecombined2": "eldredgecombined2 0.5 11 1 3 4 6 30 9 1.5 3 4 6.5",
        "motionfile": "motionfile " + path,
        "motionfileperiodic": "MotionFilePeriodic %s 2.5" % path,
    }


def geom_of(spec):
    if spec is None:
        return "body P\n line_n -0.5 0 0.5 0 21\n center 0.1 0.0\nend\n"
    return BODY.format(spec)


def run_sequence(t0, dt, nsteps):
    """Times at which an RK3 ibpm run calls moveBodies (time advanced by
    repeated addition of dt, as State::time is)."""
    cn = (8.0 / 15.0, 2.0 / 3.0, 1.0)
    ts, t = [t0, t0], t0
    for _ in range(nsteps):
        ts += [t + c * dt for c in cn]
        t = t + dt
    return np.array(ts)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libibpm_gpu.so is not built (python -m ibpm_amd.build)")
    L = ctypes.CDLL(LIB)
    L.ibpm_geom_points.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_long]
    L.ibpm_last_error.restype = ctypes.c_char_p
    return L


def lib_sequence(L, geom, times):
    times = np.ascontiguousarray(times, dtype=np.float64)
    cap = 1 << 21
    xy = np.zeros(cap)
    uv = np.zeros(cap)
    n = L.ibpm_geom_points(geom.encode(), len(times), times.ctypes.data, xy.ctypes.data, uv.ctypes.data, cap)
    assert n >= 0, L.ibpm_last_error()
    k = 2 * n * len(times)
    return xy[:k].reshape(len(times), 2 * n), uv[:k].reshape(len(times), 2 * n)


def ref_sequence(geom, times, tmp):
    R = ctypes.CDLL(REFPART)
    path = os.path.join(str(tmp), "seq.geom")
    with open(path, "w") as f:
        f.write(geom)
    times = np.ascontiguousarray(times, dtype=np.float64)
    cap = 1 << 21
    xy = np.zeros(cap)
    uv = np.zeros(cap)
    n = R.ref_geometry_sequence(path.encode(), len(times), _p(times), _p(xy), _p(uv), cap)
    assert n >= 0
    k = 2 * n * len(times)
    return xy[:k].reshape(len(times), 2 * n), uv[:k].reshape(len(times), 2 * n)


def pose(xy, uv):
    """(X, Y, theta, xdot, ydot, thetadot) of the test plate from its first and
    last points (reference offsets a = -0.6 and 0.4 from the centre, b = 0)."""
    n = xy.shape[-1] // 2
    x, y, u, v = xy[..., :n], xy[..., n:], uv[..., :n], uv[..., n:]
    a0, a1 = -0.5 - XC, 0.5 - XC
    th = np.arctan2(y[..., -1] - y[..., 0], x[..., -1] - x[..., 0])
    c, s = np.cos(th), np.sin(th)
    X = x[..., 0] - XC - a0 * c
    Y = y[..., 0] - a0 * s
    thd = ((v[..., -1] - v[..., 0]) * c - (u[..., -1] - u[..., 0]) * s) / (a1 - a0)
    xd = u[..., 0] + a0 * s * thd
    yd = v[..., 0] - a0 * c * thd
    return X, Y, th, xd, yd, thd


NAMES = list(specs("/tmp").keys())


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("t0", [0.0, 3.2])
def test_library_matches_oracle_and_reference_bitexact(lib, oracle, tmp_path, name, t0):
    geom = geom_of(specs(tmp_path)[name])
    times = run_sequence(t0, 0.02, 300 if t0 == 0.0 else 60)
    oxy, ouv = oracle.geom_sequence(geom, times)
    gxy, guv = lib_sequence(lib, geom, times)
    assert np.array_equal(gxy, oxy) and np.array_equal(guv, ouv)
    if not os.path.exists(REFPART):
        pytest.skip("oracle/_ref not built (reference tree absent): library == oracle checked")
    rxy, ruv = ref_sequence(geom, times, tmp_path)
    assert np.array_equal(rxy, oxy) and np.array_equal(ruv, ouv)


def test_several_moving_bodies_keep_their_own_state(lib, oracle, tmp_path):
    sp = specs(tmp_path)
    geom = (BODY.format(sp["fixedvel"]) + BODY.format(sp["eldredge2"]).replace("body P", "body Q")
            + BODY.format(sp["eldredgecombined2"]).replace("body P", "body R"))
    times = run_sequence(0.0, 0.05, 150)
    oxy, ouv = oracle.geom_sequence(geom, times)
    gxy, guv = lib_sequence(lib, geom, times)
    assert np.array_equal(gxy, oxy) and np.array_equal(guv, ouv)
    one, _ = oracle.geom_sequence(geom_of(sp["eldredge2"]), times)
    n = one.shape[1] // 2
    assert np.array_equal(oxy[:, [*range(n, 2 * n), *range(4 * n, 5 * n)]], one)   # body Q alone
    if os.path.exists(REFPART):
        rxy, ruv = ref_sequence(geom, times, tmp_path)
        assert np.array_equal(rxy, oxy) and np.array_equal(ruv, ouv)


SMOOTH = ["pitchplunge", "sigmoidalstep", "lagstep1", "lagstep2", "eldredge", "eldredge1"]
INTEGRATING = ["fixedvel", "eldredge2", "eldredgecombined2"]


@pytest.mark.parametrize("name", SMOOTH + INTEGRATING)
def test_velocities_are_time_derivatives(oracle, tmp_path, name):
    """A dropped term, a wrong sign or a wrong scale in any velocity formula
    breaks this: positions are differenced and compared with the velocities
    each call reports."""
    h = 1e-3
    times = np.arange(0.0, 8.0, h)
    xy, uv = oracle.geom_sequence(geom_of(specs(tmp_path)[name]), times)
    X, Y, th, xd, yd, thd = pose(xy, uv)
    for pos, vel in ((X, xd), (Y, yd), (th, thd)):
        scale = max(np.abs(vel).max(), 1e-3)
        if name in INTEGRATING:   # position += velocity(t) * dt: one-sided difference
            fd = (pos[1:] - pos[:-1]) / h
            err = np.abs(fd - vel[1:])
            tol = 2e-2
        else:
            fd = (pos[2:] - pos[:-2]) / (2 * h)
            err = np.abs(fd - vel[1:-1])
            tol = 1e-3
        assert err.max() <= tol * scale, (name, err.max(), scale)


def test_maneuvers_reach_their_amplitudes(oracle, tmp_path):
    sp = specs(tmp_path)
    tt = 3.5                       # (t2 + t3) / 2: G = maxG
    _, _, th, *_ = pose(*oracle.geom_sequence(geom_of(sp["eldredge"]), [tt]))
    assert abs(th[0] - 45 * D2R) <= 1e-14
    _, Y, *_ = pose(*oracle.geom_sequence(geom_of(sp["eldredge1"]), [tt]))
    assert abs(Y[0] - 0.5) <= 1e-14
    *_, yd, _ = pose(*oracle.geom_sequence(geom_of(sp["eldredge2"]), [0.0, tt]))
    assert abs(yd[1] - 0.5) <= 1e-14
    *_, yd, _ = pose(*oracle.geom_sequence(geom_of(sp["eldredgecombined2"]), [0.0, tt]))
    assert abs(yd[1] - 0.5) <= 1e-14
    # SigmoidalStep: half the amplitude at t0 + DUR/2, all of it by t0 + 2 DUR
    _, _, th, *_ = pose(*oracle.geom_sequence(geom_of(sp["sigmoidalstep"]), [2.0, 5.0]))
    assert abs(th[0] - 15 * D2R) <= 1e-14 and abs(th[1] - 30 * D2R) <= 1e-14
    # lag steps: zero before T0, settled at the amplitude long after the pulse
    for name in ("lagstep1", "lagstep2"):
        _, _, th, _, _, thd = pose(*oracle.geom_sequence(geom_of(sp[name]), [0.4, 0.5, 2.5, 30.0]))
        assert th[0] == 0 and th[1] == 0 and thd[0] == 0
        assert abs(th[3] - 20 * D2R) <= 1e-12 and abs(thd[3]) <= 1e-12
        assert 0 < th[2] < 20 * D2R
    # FixedVelocity: position = velocity * t from t = 0; a first call more than
    # 1 after the last one (a restart) integrates nothing (FixedVelocity.h:44)
    X, Y, th, *_ = pose(*oracle.geom_sequence(geom_of(sp["fixedvel"]), [0.0, 0.25, 0.5, 0.75]))
    assert np.allclose(X, 0.3 * np.array([0, 0.25, 0.5, 0.75]), rtol=0, atol=1e-14)
    assert np.allclose(th, 0.2 * np.array([0, 0.25, 0.5, 0.75]), rtol=0, atol=1e-14)
    X, *_ = pose(*oracle.geom_sequence(geom_of(sp["fixedvel"]), [3.2, 3.2, 3.7]))
    assert abs(X[0]) <= 1e-15 and abs(X[2] - 0.3 * 0.5) <= 1e-14


def test_motion_files_interpolate_and_repeat(oracle, tmp_path):
    sp = specs(tmp_path)
    _, rows = motion_table(tmp_path)
    knots = rows[:-1, 0]          # the last knot closes the last interval
    X, Y, th, xd, yd, thd = pose(*oracle.geom_sequence(geom_of(sp["motionfile"]), knots))
    got = np.stack([X, Y, th, xd, yd, thd], 1)
    assert np.allclose(got, rows[:-1, 1:], rtol=0, atol=1e-13)
    mid = 0.5 * (rows[3, 0] + rows[4, 0])
    X, *_ = pose(*oracle.geom_sequence(geom_of(sp["motionfile"]), [mid, -1.0, 2.6]))
    assert abs(X[0] - 0.5 * (rows[3, 1] + rows[4, 1])) <= 1e-13
    assert abs(X[1]) <= 1e-15 and abs(X[2]) <= 1e-15          # outside the table: at rest
    ts = np.array([0.3, 1.7, 2.2])
    a = np.stack(pose(*oracle.geom_sequence(geom_of(sp["motionfileperiodic"]), ts)), 1)
    b = np.stack(pose(*oracle.geom_sequence(geom_of(sp["motionfileperiodic"]), ts + 2.5)), 1)
    assert np.allclose(a, b, rtol=0, atol=1e-12)


def test_stationary_only_for_no_motion_and_fixed(lib, oracle, tmp_path):
    sp = specs(tmp_path)
    for name, spec in sp.items():
        s = oracle.Sim(16, 16, 1, 2.0, -1.0, -1.0, 100.0, 0.01, geom_of(spec))
        assert s.uses_cg == (name not in ("none", "fixed")), name
    if os.path.exists(REFPART):
        R = ctypes.CDLL(REFPART)
        for name, spec in sp.items():
            path = os.path.join(str(tmp_path), "st.geom")
            with open(path, "w") as f:
                f.write(geom_of(spec))
            assert R.ref_geometry_stationary(path.encode()) == (1 if name in ("none", "fixed") else 0), name


@pytest.mark.parametrize("bad", [
    "eldredge 45 11 1 3 4",
    "lagstep1 20 2.0 0.3",
    "motionfile /nonexistent/plate.motion",
    "motionfileperiodic",
    "wobble 1 2 3",
])
def test_malformed_motions_are_errors(lib, bad):
    cap = 1 << 10
    xy = np.zeros(cap)
    n = lib.ibpm_geom_points(BODY.format(bad).encode(), 1, np.zeros(1).ctypes.data, xy.ctypes.data,
                             xy.ctypes.data, cap)
    assert n == -1 and lib.ibpm_last_error()
