"""Unsteady base flow (ibpm -ubf, ibpm.cc:166-178; BaseFlow.cc:111-150,
Flux.cc:101-129) on the CPU: the motion of the only body moves to the base
flow and the flow is solved in the body frame.

  * the oracle's base flux on every level and its (alpha, mag) equal the
    reference's BaseFlow compiled from its sources (oracle/_ref) bit for bit
    over a sequence of moveFlow calls, stateful motions included;
  * independently of that build, the mechanics fix the flow seen by the body:
    pure pitch gives the free stream turned by -theta plus the rotation
    -thetadot about the centre, pure plunge gives (1, -ydot);
  * a body whose transferred motion is at rest reproduces the steady run
    (same forces to rounding), and more than one body is refused.
No GPU call is made here.
"""
import ctypes
import math
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFPART = os.path.join(ROOT, "oracle", "_ref", "libibpm_refpart.so")
GRID = (16, 12, 3, 2.0, -1.0, -0.75)      # nx, ny, ngrid, length, xoffset, yoffset
PLATE = "body P\n line_n -0.5 0 0.5 0 21\n center 0.1 0.05\n motion {}\nend\n"


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def edge_coordinates(nx, ny, ngrid, length, xoff, yoff):
    """(x, y, d) of every flux of every level in the Flux layout (X fluxes at
    x edges and y centres, then Y fluxes), Grid.cc:104-139 with zero shift."""
    dx = length / nx
    xs, ys, ds = [], [], []
    for lev in range(ngrid):
        d = dx * (1 << lev)
        x0 = xoff - 0.5 * ((1 << lev) - 1) * nx * dx
        y0 = yoff - 0.5 * ((1 << lev) - 1) * ny * dx
        i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny), indexing="ij")
        xs.append((x0 + i * d).ravel()); ys.append((y0 + (j + 0.5) * d).ravel())
        i, j = np.meshgrid(np.arange(nx), np.arange(ny + 1), indexing="ij")
        xs.append((x0 + (i + 0.5) * d).ravel()); ys.append((y0 + j * d).ravel())
        ds.append(np.full(2 * nx * ny + nx + ny, d))
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(ds)


def is_x(nx, ny, ngrid):
    one = np.r_[np.ones((nx + 1) * ny, bool), np.zeros(nx * (ny + 1), bool)]
    return np.tile(one, ngrid)


@pytest.mark.parametrize("motion", [
    "pitchplunge 0.3 0.25 0.1 0.2 0.4 0.5",
    "fixedvel 0.3 -0.1 0.2",
    "eldredge2 0.5 11 1 3 4 6",
    "eldredgecombined2 0.5 11 1 3 4 6 30 9 1.5 3 4 6.5",
    "lagstep1 20 2.0 0.3 0.5",
])
@pytest.mark.parametrize("alpha_deg", [0.0, 12.0])
def test_oracle_base_flow_equals_reference_bitexact(oracle, tmp_path, motion, alpha_deg):
    if not os.path.exists(REFPART):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    geom = PLATE.format(motion)
    path = tmp_path / "u.geom"
    path.write_text(geom)
    nx, ny, ng, L, xo, yo = GRID
    times = np.array([0.0, 0.0, 0.1, 0.2, 0.35, 1.5, 3.5, 3.5, 4.25])
    q, am = oracle.baseflow_sequence(geom, nx, ny, ng, L, xo, yo, alpha_deg, times)
    R = ctypes.CDLL(REFPART)
    qr = np.zeros(q.size)
    amr = np.zeros(am.size)
    D = ctypes.c_double
    alpha = alpha_deg * (4 * math.atan(1)) / 180.
    assert R.ref_baseflow_sequence(str(path).encode(), nx, ny, ng, D(L), D(xo), D(yo), D(alpha), len(times),
                                   _p(times), _p(qr), _p(amr)) == 0
    assert np.array_equal(q.ravel(), qr) and np.array_equal(am.ravel(), amr)


def test_pitch_gives_turned_stream_plus_rotation(oracle):
    """theta = A sin(2 pi f t): in the body frame the free stream (angle 0)
    arrives at -theta, and the fluid turns at -thetadot about the centre."""
    A, f = 0.3, 0.25
    geom = PLATE.format("pitchplunge %r %r 0 0 0 0" % (A, f))
    nx, ny, ng, L, xo, yo = GRID
    times = np.array([0.0, 0.3, 0.9, 1.7])
    q, am = oracle.baseflow_sequence(geom, nx, ny, ng, L, xo, yo, 0.0, times)
    x, y, d = edge_coordinates(*GRID)
    X = is_x(nx, ny, ng)
    for c, t in enumerate(times):
        th = A * math.sin(2 * math.pi * f * t)
        thd = A * 2 * math.pi * f * math.cos(2 * math.pi * f * t)
        want = np.where(X, (math.cos(th) + thd * (y - 0.05)) * d, (-math.sin(th) - thd * (x - 0.1)) * d)
        assert np.abs(q[c] - want).max() <= 1e-14 * np.abs(want).max()
        assert abs(am[c, 0] + th) <= 1e-15 and abs(am[c, 1] - 1) <= 1e-15


def test_plunge_gives_stream_against_the_motion(oracle):
    """y = h sin(2 pi f t): the body frame sees (1, -ydot), magnitude
    sqrt(1 + ydot^2), and no rotation."""
    h, f = 0.2, 0.4
    geom = PLATE.format("pitchplunge 0 0 0 %r %r 0" % (h, f))
    nx, ny, ng, L, xo, yo = GRID
    times = np.array([0.1, 0.6, 1.3])
    q, am = oracle.baseflow_sequence(geom, nx, ny, ng, L, xo, yo, 0.0, times)
    _, _, d = edge_coordinates(*GRID)
    X = is_x(nx, ny, ng)
    for c, t in enumerate(times):
        yd = h * 2 * math.pi * f * math.cos(2 * math.pi * f * t)
        want = np.where(X, d, -yd * d)
        assert np.abs(q[c] - want).max() <= 1e-14
        assert abs(am[c, 1] - math.hypot(1, yd)) <= 1e-15


def test_transferred_motion_at_rest_reproduces_the_steady_run(oracle):
    """A pitchplunge of zero amplitude is a moving body to the reference
    (only FixedPosition is stationary), so -ubf transfers it: the base flow is
    then the uniform stream again, up to the rounding of atan2/sqrt."""
    plate = "body P\n line_n -0.45 0.05 0.45 -0.05 15\n center 0 0\n{}end\n"
    steady = oracle.Sim(64, 64, 2, 4.0, -2.0, -2.0, 100.0, 0.02, plate.format(""), alpha_deg=8.0)
    ubf = oracle.Sim(64, 64, 2, 4.0, -2.0, -2.0, 100.0, 0.02,
                     plate.format(" motion pitchplunge 0 0 0 0 0 0\n"), alpha_deg=8.0, ubf=True)
    assert ubf.ubf == 1 and not ubf.uses_cg
    steady.init()
    ubf.init()
    for _ in range(3):
        steady.advance(1)
        ubf.advance(1)
        (cd, cl), (cdu, clu) = steady.forces(), ubf.forces()
        assert abs(cdu - cd) <= 1e-12 * abs(cd) and abs(clu - cl) <= 1e-12 * abs(cd)
    assert np.abs(ubf.omega() - steady.omega()).max() <= 1e-12 * np.abs(steady.omega()).max()


def test_more_than_one_body_is_refused(oracle):
    two = PLATE.format("fixedvel 0.1 0 0") + "body Q\n point 0.5 0.5\nend\n"
    with pytest.raises(ValueError):
        oracle.Sim(16, 16, 1, 2.0, -1.0, -1.0, 100.0, 0.01, two, ubf=True)
    one_still = "body P\n line_n -0.5 0 0.5 0 11\nend\n"
    assert oracle.Sim(16, 16, 1, 2.0, -1.0, -1.0, 100.0, 0.01, one_still, ubf=True).ubf == 0
