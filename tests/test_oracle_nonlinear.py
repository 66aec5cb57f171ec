"""Pins of the oracle's nonlinear term and time stepping -- CPU only.

The cross products (VectorOperations.cc:438-486) carry the sign of the
nonlinear term N = Curl(q x omega) (IBSolver.cc:205-209).  They are pinned
here, independently of any other part of the oracle, by:
  - the reference's own cross-product tests (VectorOperationsTest.cc:593-693)
    on its 8 x 12, 3-level grid with all 78 of its test functions:
    q x q = 0, (u,0) x (0,v) = uv for constants, q1 x q2 = -(q2 x q1) and
    the triple product <a, q1 x q2> = <q1, q2 x a>.  The first three fix the
    Flux x Flux product and its sign; the triple product then fixes the
    sign of Flux x Scalar, the product N uses;
  - a vortex blob in the uniform base flow, which must travel downstream at
    the free-stream speed.  A dropped sign sends it upstream, and RK3
    coefficients that do not sum to one (sum a_n + b_n) move it at the wrong
    speed;
  - the time-convergence order of the RK3 scheme (Scheme.h:33-45) on the same
    blob: successive halvings of dt must shrink the difference between runs by
    about 2^3.  A wrong coefficient that keeps the sum at one drops the order.
"""
import math

import numpy as np
import pytest

from tests.helpers import G

VG = G(8, 12, 3, 2.0, -1.0, -3.0)          # VectorOperationsTest.cc:14-38
NSCALARS = (VG.nx - 1) * (VG.ny - 1) + 1   # :35-36


def _scalar(m):
    """getScalar(m), VectorOperationsTest.cc:140-150."""
    if m == 0:
        return np.zeros(VG.ngrid * VG.ns)
    kx = (m - 1) % (VG.nx - 1) + 1
    ky = (m - 1) // (VG.nx - 1) + 1
    return VG.single_wavenumber(kx, ky)


@pytest.fixture(scope="module")
def fields(oracle):
    sc = [_scalar(m) for m in range(NSCALARS)]
    fl = [oracle.curl_scalar(VG.nx, VG.ny, VG.ngrid, s) for s in sc]   # getFlux, :152-158
    return sc, fl


def _fudge_boundary(val, f):
    """SetScalarBoundary, VectorOperationsTest.cc:117-128 (coarsest level)."""
    g = VG
    lev = g.ngrid - 1
    f = f.copy()
    for i in range(1, g.nx):
        f[g.si(lev, i, 1)] = val
        f[g.si(lev, i, g.ny - 1)] = val
    for j in range(1, g.ny):
        f[g.si(lev, 1, j)] = val
        f[g.si(lev, g.nx - 1, j)] = val
    return f


def test_cross_product_with_self_is_zero(oracle, fields):
    _, fl = fields
    g = VG
    for q in fl:
        assert np.all(oracle.cross_flux_flux(g.nx, g.ny, g.ngrid, g.dx, q, q) == 0.0)


def test_cross_product_of_orthogonal_constants(oracle):
    g = VG
    u = np.full(g.ngrid * g.ns, 3.0)
    v = np.full(g.ngrid * g.ns, 5.0)
    q1 = oracle.xvel_to_flux(g.nx, g.ny, g.ngrid, g.dx, u)
    q2 = oracle.yvel_to_flux(g.nx, g.ny, g.ngrid, g.dx, v)
    cross = oracle.cross_flux_flux(g.nx, g.ny, g.ngrid, g.dx, q1, q2)
    err = _fudge_boundary(0.0, u * v - cross)
    assert oracle.ip_scalar(g.nx, g.ny, g.ngrid, g.dx, err, err) < 2e-28
    # the sign itself: +15 inside, so a flipped product (-15) is far off
    assert abs(cross[g.si(0, 4, 6)] - 15.0) < 1e-12


def test_cross_product_antisymmetric(oracle, fields):
    _, fl = fields
    g = VG
    for m in range(0, NSCALARS, 3):
        for n in range(NSCALARS):
            a = oracle.cross_flux_flux(g.nx, g.ny, g.ngrid, g.dx, fl[m], fl[n])
            b = oracle.cross_flux_flux(g.nx, g.ny, g.ngrid, g.dx, fl[n], fl[m])
            assert np.array_equal(a, -b)


def test_cross_product_triple_product(oracle, fields):
    # <a, q1 x q2> = <q1, q2 x a> for all test scalars a and fluxes q1, q2
    # (VectorOperationsTest.cc:657-693, tol 5e-14)
    sc, fl = fields
    g = VG
    worst = big = 0.0
    for m in range(NSCALARS):
        a = sc[m]
        for n in range(NSCALARS):
            q1, q2 = fl[n], fl[NSCALARS - n - 1]
            q1q2 = oracle.cross_flux_flux(g.nx, g.ny, g.ngrid, g.dx, q1, q2)
            q2a = oracle.cross_product(g.nx, g.ny, g.ngrid, g.dx, q2, a)
            ipf = oracle.ip_flux(g.nx, g.ny, g.ngrid, q1, q2a)
            ips = oracle.ip_scalar(g.nx, g.ny, g.ngrid, g.dx, a, q1q2)
            worst = max(worst, abs(ipf - ips))
            big = max(big, abs(ips))
    assert worst < 5e-14
    # the identity holds between sizable inner products, so a sign slip in
    # either product cannot hide in it
    assert big > 1e-2


# ----------------------------------------------------- blob in uniform flow
BLOB_NX, BLOB_LEN = 48, 1.5
BLOB_RE = 1000.0
BLOB_X0, BLOB_SIG = -0.25, 0.09


def _blob_run(oracle, dt, nsteps, amp=1e-4):
    g = G(BLOB_NX, BLOB_NX, 1, BLOB_LEN, -BLOB_LEN / 2, -BLOB_LEN / 2)
    s = oracle.Sim(g.nx, g.ny, 1, g.length, g.xoff, g.yoff, BLOB_RE, dt, "", alpha_deg=0.0)
    s.init()
    w0 = g.scalar(lambda x, y: amp * math.exp(-((x - BLOB_X0) ** 2 + y ** 2) / BLOB_SIG ** 2))
    s.set_omega(w0)
    s.advance(nsteps)
    return g, w0, s.omega()


def _centroid_x(g, w):
    xs = g.scalar(lambda x, y: x)
    return float(np.dot(xs, w) / np.sum(w))


def test_vortex_blob_moves_downstream_at_free_stream_speed(oracle):
    # Free stream U = 1 along +x (alpha 0).  A weak blob is carried by it:
    # after T its centroid has moved U T downstream.
    dt, nsteps = 0.01, 30
    g, w0, w = _blob_run(oracle, dt, nsteps)
    shift = _centroid_x(g, w) - _centroid_x(g, w0)
    T = dt * nsteps
    assert abs(shift - T) < 0.03 * T, shift


@pytest.fixture(scope="module")
def blob_series(oracle):
    T = 0.24
    out = {}
    for n in (6, 12, 24, 48):
        out[n] = _blob_run(oracle, T / n, n)[2]
    return out


def test_rk3_time_convergence_order(blob_series):
    # Differences between successive halvings of dt shrink by 2^p; RK3 with the
    # viscous term treated by the trapezoidal rule is third order while the
    # explicit (advective) error dominates, as here (nu dt / dx^2 < 0.02).
    w = blob_series
    d1 = np.linalg.norm(w[6] - w[12])
    d2 = np.linalg.norm(w[12] - w[24])
    d3 = np.linalg.norm(w[24] - w[48])
    p1, p2 = math.log2(d1 / d2), math.log2(d2 / d3)
    assert 2.6 < p1 < 3.4 and 2.6 < p2 < 3.4, (p1, p2)
