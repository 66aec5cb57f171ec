"""Pins of the CPU FP64 oracle (oracle/ibpm_oracle.c) -- CPU only.

How each oracle function is pinned (a plausible slip -- dropped term, wrong
sign or index, transposed operand -- fails at least one check):

  og_dst2d            direct summation (og_dst2d_direct, exact sine table),
                      scipy.fft.dstn(type=1) at every configured length, and
                      the involution DST(DST(x)) = 4 (n1+1)(n2+1) x.
  og_elliptic2d_solve reference EllipticSolver2dTest closed forms: every single
                      wavenumber satisfies L(L^-1 f) = f and (1+aL)u = f with the
                      5-point Laplacian (1e-10); constant and linear BC cases.
  og_elliptic_solve   reference EllipticSolverTest identity Lap(solve f) =
                      coarsify f on all levels, plus the Helmholtz analogue
                      (its reference assertion is commented out; it holds here).
  og_coarsify/getbc   bit-exact vs the reference's own Scalar.cc (oracle/_ref),
                      constants preserved, linear fields reproduced.
  operators           reference VectorOperationsTest identities: adjointness of
                      FluxToX/YVelocity and X/YVelocityToFlux under the weighted
                      multi-domain inner products (1e-13), curl of linear and
                      constant fields, <curl q, a> = <q, curl a>.
  og_reg_*            bit-exact vs the reference's own Regularizer.cc (oracle/_ref)
                      and the RegularizerTest properties (sum 1, zero moment, 6+6
                      nonzeros, support, interpolate constant, round trip 1/4).
  geometry/motion     bit-exact vs the reference's Geometry/RigidBody/PitchPlunge.
  projection          ProjectionSolverTest identities (1e-10) for Cholesky and CG;
                      Cholesky factor reproduces M; CG == Cholesky.
  time stepping       no-body Oseen-type decay check and zero-state fixed point;
                      whole-run Cd/Cl are pinned only through these pieces
                      (no golden end-to-end data exists in the reference).
"""
import ctypes
import math
import os
import tempfile

import numpy as np
import pytest
import scipy.fft as sf

from tests.helpers import G, bc_ring, single_wavenumber_2d, CYLINDER_157, plate_geom

REFPART = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "oracle", "_ref", "libibpm_refpart.so")


def refpart():
    if not os.path.exists(REFPART):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return ctypes.CDLL(REFPART)


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ---------------------------------------------------------------- DST-I
@pytest.mark.parametrize("shape", [(1, 1), (2, 5), (3, 7), (11, 23), (24, 9), (31, 17)])
def test_dst_matches_direct_sum(oracle, shape):
    x = np.random.default_rng(1).standard_normal(shape)
    a = oracle.dst2d(x)
    b = oracle.dst2d_direct(x)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


@pytest.mark.parametrize("n", [199, 255, 383, 399, 511, 1023])
def test_dst_matches_scipy_rodft00(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((7, n)) if n > 511 else rng.standard_normal((n, n - 16))
    a = oracle.dst2d(x)
    c = sf.dstn(x, type=1)
    assert np.abs(a - c).max() <= 1e-14 * np.abs(c).max() * math.log2(n)
    xt = np.ascontiguousarray(x.T)
    assert np.abs(oracle.dst2d(xt) - sf.dstn(xt, type=1)).max() <= 1e-14 * np.abs(c).max() * math.log2(n)


def test_dst_involution(oracle):
    x = np.random.default_rng(3).standard_normal((19, 12))
    y = oracle.dst2d(oracle.dst2d(x))
    assert np.allclose(y, 4 * 20 * 13 * x, rtol=0, atol=1e-11)


# ---------------------------------------------------------- 2-D elliptic
NX, NY, DX, ALPHA = 4, 8, 0.1, 0.2


def _zero_bc(nx, ny):
    return np.zeros(2 * (nx + ny))


@pytest.mark.parametrize("kind", [0, 1])
def test_elliptic2d_every_wavenumber(oracle, kind):
    bc = _zero_bc(NX, NY)
    for kx in range(NX):
        for ky in range(NY):
            f = single_wavenumber_2d(kx, ky, NX, NY).ravel()
            u = oracle.elliptic2d_solve(kind, NX, NY, DX, ALPHA, f)
            Lu = oracle.laplacian2d_bc(NX, NY, DX, u, bc)
            lhs = Lu if kind == 0 else u + ALPHA * Lu
            assert np.abs(lhs - f).max() < 1e-10


@pytest.mark.parametrize("kind", [0, 1])
def test_elliptic2d_constant_bc(oracle, kind):
    val = 8.0
    bc = np.full(2 * (NX + NY), val)
    f = np.zeros((NX - 1) * (NY - 1)) if kind == 0 else np.full((NX - 1) * (NY - 1), val)
    u = oracle.elliptic2d_solve(kind, NX, NY, DX, ALPHA, f, bc)
    assert np.abs(u - val).max() < 1e-10


@pytest.mark.parametrize("kind", [0, 1])
def test_elliptic2d_linear_bc(oracle, kind):
    ue = lambda i, j: i * DX + 2 * j * DX
    bc = bc_ring(NX, NY, left=lambda j: ue(0, j), top=lambda i: ue(i, NY),
                 right=lambda j: ue(NX, j), bottom=lambda i: ue(i, 0))
    exact = np.array([[ue(i, j) for j in range(1, NY)] for i in range(1, NX)]).ravel()
    f = np.zeros_like(exact) if kind == 0 else exact.copy()
    u = oracle.elliptic2d_solve(kind, NX, NY, DX, ALPHA, f, bc)
    assert np.abs(u - exact).max() < 1e-10


def test_elliptic2d_detects_wrong_sign(oracle):
    # Planted mistake: solving with the Helmholtz sign flipped must break the identity.
    bc = _zero_bc(NX, NY)
    f = single_wavenumber_2d(1, 2, NX, NY).ravel()
    u = oracle.elliptic2d_solve(1, NX, NY, DX, -ALPHA, f)
    Lu = oracle.laplacian2d_bc(NX, NY, DX, u, bc)
    assert np.abs(u + ALPHA * Lu - f).max() > 1e-3


# -------------------------------------------------------- multi-domain
@pytest.mark.parametrize("nx,ny,ngrid,length", [(4, 8, 1, 0.4), (4, 8, 2, 0.4), (16, 24, 3, 2.0)])
def test_multidomain_poisson_identity(oracle, nx, ny, ngrid, length):
    g = G(nx, ny, ngrid, length, -1.0, -1.0)
    for kx in range(nx):
        for ky in range(0, ny, max(1, ny // 6)):
            f = g.single_wavenumber(kx, ky)
            u = oracle.elliptic_solve(0, nx, ny, ngrid, g.dx, 0.0, f)
            Lu = oracle.laplacian(nx, ny, ngrid, g.dx, u)
            assert np.abs(Lu - oracle.coarsify(nx, ny, ngrid, f)).max() < 1e-10 * max(1, np.abs(f).max())


@pytest.mark.parametrize("nx,ny,ngrid", [(4, 8, 2), (16, 24, 3)])
def test_multidomain_helmholtz_identity(oracle, nx, ny, ngrid):
    g = G(nx, ny, ngrid, 0.4, -1.0, -1.0)
    alpha = -2.0 * g.dx ** 2   # the solver uses alpha = -beta nu / 2 < 0
    rng = np.random.default_rng(5)
    f = rng.standard_normal(ngrid * g.ns)
    u = oracle.elliptic_solve(1, nx, ny, ngrid, g.dx, alpha, f)
    Lu = oracle.laplacian(nx, ny, ngrid, g.dx, u)
    assert np.abs(u + alpha * Lu - oracle.coarsify(nx, ny, ngrid, f)).max() < 1e-10


def test_coarsify_constant_and_idempotent(oracle):
    g = G(16, 24, 3, 2.0, -1, -3)
    one = np.full(3 * g.ns, 1.7)
    assert np.allclose(oracle.coarsify(16, 24, 3, one), 1.7, rtol=0, atol=1e-15)
    x = np.random.default_rng(0).standard_normal(3 * g.ns)
    c1 = oracle.coarsify(16, 24, 3, x)
    assert np.array_equal(oracle.coarsify(16, 24, 3, c1), c1)


def test_getbc_linear_field(oracle):
    g = G(16, 24, 3, 2.0, -1, -3)
    a, b, c = 7.0, 11.0, 13.0
    f = g.scalar(lambda x, y: a * x + b * y + c)
    for lev in range(2):
        bc = oracle.getbc(16, 24, 3, f, lev)
        ref = bc_ring(16, 24, left=lambda j: a * g.xedge(lev, 0) + b * g.yedge(lev, j) + c,
                      top=lambda i: a * g.xedge(lev, i) + b * g.yedge(lev, 24) + c,
                      right=lambda j: a * g.xedge(lev, 16) + b * g.yedge(lev, j) + c,
                      bottom=lambda i: a * g.xedge(lev, i) + b * g.yedge(lev, 0) + c)
        assert np.abs(bc - ref).max() < 1e-12


# ------------------------------------------------------------ operators
VG = G(8, 12, 3, 2.0, -1.0, -3.0)


def _test_scalars(n=12):
    out = [np.zeros(VG.ngrid * VG.ns)]
    for m in range(1, n):
        kx = (m - 1) % (VG.nx - 1) + 1
        ky = (m - 1) // (VG.nx - 1) + 1
        out.append(VG.single_wavenumber(kx, ky))
    return out


def test_flux_to_velocity_adjoint(oracle):
    g = VG
    sc = _test_scalars()
    fl = [oracle.curl_scalar(g.nx, g.ny, g.ngrid, s) for s in sc]
    for u2 in sc:
        for q1 in fl:
            # X: <q1, X* u2> = <X q1, u2>
            q2 = oracle.xvel_to_flux(g.nx, g.ny, g.ngrid, g.dx, u2, np.zeros(g.ngrid * g.nf))
            u1 = oracle.flux_to_xvel(g.nx, g.ny, g.ngrid, g.dx, q1)
            a = oracle.ip_flux(g.nx, g.ny, g.ngrid, q1, q2)
            b = oracle.ip_scalar(g.nx, g.ny, g.ngrid, g.dx, u1, u2)
            assert abs(a - b) < 1e-13
            q2 = oracle.yvel_to_flux(g.nx, g.ny, g.ngrid, g.dx, u2, np.zeros(g.ngrid * g.nf))
            v1 = oracle.flux_to_yvel(g.nx, g.ny, g.ngrid, g.dx, q1)
            a = oracle.ip_flux(g.nx, g.ny, g.ngrid, q1, q2)
            b = oracle.ip_scalar(g.nx, g.ny, g.ngrid, g.dx, v1, u2)
            assert abs(a - b) < 1e-13


def test_constant_velocity_to_flux(oracle):
    g = VG
    q = oracle.xvel_to_flux(g.nx, g.ny, g.ngrid, g.dx, np.full(g.ngrid * g.ns, 3.0))
    q = oracle.yvel_to_flux(g.nx, g.ny, g.ngrid, g.dx, np.full(g.ngrid * g.ns, 5.0), q)
    for lev in range(g.ngrid - 1):   # coarsest level's outer ring uses zero BCs
        d = g.dxl(lev)
        for i in range(g.nx + 1):
            for j in range(g.ny):
                assert abs(q[g.fxi(lev, i, j)] - 3.0 * d) < 1e-13
        for i in range(g.nx):
            for j in range(g.ny + 1):
                assert abs(q[g.fyi(lev, i, j)] - 5.0 * d) < 1e-13


def test_curl_of_linear_scalar(oracle):
    g = VG
    a, b, c = 7.0, 11.0, 13.0
    f = g.scalar(lambda x, y: a * x + b * y + c)
    q = oracle.curl_scalar(g.nx, g.ny, g.ngrid, f)
    top = g.ngrid - 1
    for lev in range(g.ngrid):
        d = g.dxl(lev)
        for i in range(g.nx + 1):
            for j in range(g.ny):
                if lev == top and (i in (0, g.nx) or j in (0, g.ny - 1)):
                    continue   # zero BC on the coarsest level's boundary
                assert abs(q[g.fxi(lev, i, j)] - b * d) < 1e-12
        for i in range(g.nx):
            for j in range(g.ny + 1):
                if lev == top and (j in (0, g.ny) or i in (0, g.nx - 1)):
                    continue
                assert abs(q[g.fyi(lev, i, j)] + a * d) < 1e-12


def test_curl_of_constant_flux_is_zero(oracle):
    g = VG
    q = np.zeros(g.ngrid * g.nf)
    for lev in range(g.ngrid):
        q[lev * g.nf: lev * g.nf + g.nfx] = 8.0
        q[lev * g.nf + g.nfx: (lev + 1) * g.nf] = 5.0
    w = oracle.curl_flux(g.nx, g.ny, g.ngrid, g.dx, q)
    assert np.abs(w).max() == 0.0


def test_curl_inner_product_single_domain(oracle):
    g = G(8, 12, 1, 2.0, -1.0, -3.0)
    for m in (1, 5, 9, 17):
        kx, ky = (m - 1) % 7 + 1, (m - 1) // 7 + 1
        a = g.single_wavenumber(kx, ky)
        q = oracle.curl_scalar(g.nx, g.ny, 1, g.single_wavenumber(ky % 7 + 1, kx))
        lhs = oracle.ip_flux(g.nx, g.ny, 1, q, oracle.curl_scalar(g.nx, g.ny, 1, a))
        rhs = oracle.ip_scalar(g.nx, g.ny, 1, g.dx, oracle.curl_flux(g.nx, g.ny, 1, g.dx, q), a)
        assert abs(lhs - rhs) < 5e-13


def test_laplacian_is_five_point_with_coarse_bc(oracle):
    # Lap = -Curl(Curl(.)); away from the coarsest boundary it equals the
    # 5-point stencil with ghosts from getBC of the next coarser level.
    g = VG
    w = np.random.default_rng(9).standard_normal(g.ngrid * g.ns)
    L = oracle.laplacian(g.nx, g.ny, g.ngrid, g.dx, w)
    for lev in range(g.ngrid):
        bc = oracle.getbc(g.nx, g.ny, g.ngrid, w, lev) if lev < g.ngrid - 1 else np.zeros(2 * (g.nx + g.ny))
        f = w[lev * g.ns:(lev + 1) * g.ns]
        ref = oracle.laplacian2d_bc(g.nx, g.ny, g.dxl(lev), f, bc)
        assert np.abs(L[lev * g.ns:(lev + 1) * g.ns] - ref).max() < 1e-9 * np.abs(ref).max()


# ---------------------------------------------------------- regulariser
RG = G(10, 20, 1, 2.0, -1.0, -2.0)


def _reg_origin(oracle):
    assoc = oracle.reg_build(RG.nx, RG.ny, RG.dx, RG.xoff, RG.yoff, np.array([0.0]), np.array([0.0]))
    u2 = oracle.reg_toflux(RG.nx, RG.ny, 1, RG.dx, assoc, np.array([3.0, 5.0]))
    return assoc, u2


def test_regularizer_properties(oracle):
    assoc, u2 = _reg_origin(oracle)
    g = RG
    ones = np.ones(g.ns)
    zeros = np.zeros(g.ns)
    qx = oracle.yvel_to_flux(g.nx, g.ny, 1, g.dx, zeros, oracle.xvel_to_flux(g.nx, g.ny, 1, g.dx, ones))
    assert abs(oracle.ip_flux(g.nx, g.ny, 1, qx, u2) / g.dx ** 2 - 3.0) < 1e-14 * 3
    qy = oracle.yvel_to_flux(g.nx, g.ny, 1, g.dx, ones, oracle.xvel_to_flux(g.nx, g.ny, 1, g.dx, zeros))
    assert abs(oracle.ip_flux(g.nx, g.ny, 1, qy, u2) / g.dx ** 2 - 5.0) < 1e-14 * 5
    xs = np.array([g.xedge(0, i) for i in range(g.nx + 1) for j in range(g.ny)])
    assert abs(np.dot(xs, u2[:g.nfx])) < 1e-14
    assert np.count_nonzero(np.abs(u2[:g.nfx]) > 1e-14) == 6
    assert np.count_nonzero(np.abs(u2[g.nfx:]) > 1e-14) == 6
    f = oracle.reg_toboundary(1, g.dx, assoc, u2)
    assert abs(f[0] - 0.75) < 1e-14 and abs(f[1] - 1.25) < 1e-14
    q = oracle.yvel_to_flux(g.nx, g.ny, 1, g.dx, np.full(g.ns, 9.0),
                            oracle.xvel_to_flux(g.nx, g.ny, 1, g.dx, np.full(g.ns, 7.0)))
    f = oracle.reg_toboundary(1, g.dx, assoc, q)
    assert abs(f[0] - 7) < 1e-14 and abs(f[1] - 9) < 1e-14


def test_regularizer_support(oracle):
    _, u2 = _reg_origin(oracle)
    g = RG
    sup = 1.5 * g.dx
    for i in range(g.nx + 1):
        for j in range(g.ny):
            x, y = g.xedge(0, i), g.ycenter(0, j)
            v = u2[g.fxi(0, i, j)]
            if abs(x) > sup or abs(y) > sup:
                assert v == 0.0
            if abs(x) + 1e-14 < sup and abs(y) + 1e-14 < sup:
                assert abs(v) > 0


# ---------------------------------------------- reference (partial build)
def test_refpart_coarsify_getbc_bitexact(oracle):
    R = refpart()
    nx, ny, ng = 16, 24, 3
    s = np.random.default_rng(1).standard_normal(ng * (nx - 1) * (ny - 1))
    b = s.copy()
    R.ref_coarsify(nx, ny, ng, ctypes.c_double(2.0), ctypes.c_double(-1), ctypes.c_double(-3), _p(b))
    assert np.array_equal(oracle.coarsify(nx, ny, ng, s), b)
    for lev in range(ng - 1):
        br = np.zeros(2 * (nx + ny))
        R.ref_getbc(nx, ny, ng, ctypes.c_double(2.0), ctypes.c_double(-1), ctypes.c_double(-3), _p(s), lev, _p(br))
        assert np.array_equal(oracle.getbc(nx, ny, ng, s, lev), br)


GEOMS = [
    plate_geom(40, 0.9, 0.2, 0.25, 0.35, (0.1, 0.2)),
    "body C\n circle 0.3 -0.2 0.25 0.04\nend\nbody L\n line -0.5 0.5 0.5 0.7 0.05\n motion fixed 0.1 0.05 0.3\nend\n",
    "body A\n line_aoa 1.0 0.25 0 12 30\nend\nbody P\n point 0.2 0.2\n point 0.25 0.21\nend\n",
]


@pytest.mark.parametrize("gi", range(len(GEOMS)))
def test_refpart_geometry_and_regularizer_bitexact(oracle, gi):
    R = refpart()
    geom = GEOMS[gi]
    fd, fn = tempfile.mkstemp(suffix=".geom")
    os.write(fd, geom.encode())
    os.close(fd)
    try:
        nx, ny, ng, L, xo, yo = 64, 48, 2, 4.0, -2.0, -1.5
        for t in (0.0, 0.37, 1.9):
            cap = 4000
            xy = np.zeros(2 * cap)
            uv = np.zeros(2 * cap)
            n = R.ref_geometry(fn.encode(), ctypes.c_double(t), _p(xy), _p(uv), cap)
            oxy, ouv = oracle.geom_points(geom, t)
            assert n == len(oxy) // 2
            assert np.array_equal(np.concatenate([xy[:n], xy[n:2 * n]]), oxy)
            assert np.array_equal(np.concatenate([uv[:n], uv[n:2 * n]]), ouv)
            dx = L / nx
            assoc = oracle.reg_build(nx, ny, dx, xo, yo, oxy[:n], oxy[n:])
            rng = np.random.default_rng(int(t * 100) + gi)
            f = rng.standard_normal(2 * n)
            q = rng.standard_normal(ng * (2 * nx * ny + nx + ny))
            qr = np.zeros_like(q)
            fr = np.zeros_like(f)
            R.ref_regularizer(nx, ny, ng, ctypes.c_double(L), ctypes.c_double(xo), ctypes.c_double(yo),
                              fn.encode(), ctypes.c_double(t), _p(f), _p(qr), _p(q), _p(fr))
            assert np.array_equal(oracle.reg_toflux(nx, ny, ng, dx, assoc, f), qr)
            assert np.array_equal(oracle.reg_toboundary(n, dx, assoc, q), fr)
    finally:
        os.unlink(fn)


def test_refpart_uniform_flow_and_force(oracle):
    R = refpart()
    nx, ny, ng, L = 20, 16, 3, 4.0
    alpha = 0.3
    q = np.zeros(ng * (2 * nx * ny + nx + ny))
    R.ref_uniform_flow(nx, ny, ng, ctypes.c_double(L), ctypes.c_double(-2), ctypes.c_double(-2),
                       ctypes.c_double(alpha), _p(q))
    g = G(nx, ny, ng, L, -2, -2)
    for lev in range(ng):
        assert np.all(q[lev * g.nf: lev * g.nf + g.nfx] == math.cos(alpha) * g.dxl(lev))
        assert np.all(q[lev * g.nf + g.nfx:(lev + 1) * g.nf] == math.sin(alpha) * g.dxl(lev))
    f = np.random.default_rng(2).standard_normal(14)
    xf, yf = ctypes.c_double(), ctypes.c_double()
    R.ref_net_force(nx, ny, ng, ctypes.c_double(L), ctypes.c_double(-2), ctypes.c_double(-2), 7, _p(f),
                    ctypes.byref(xf), ctypes.byref(yf))
    dx2 = (L / nx) ** 2
    assert xf.value == sum(f[:7]) * dx2 and yf.value == sum(f[7:]) * dx2


# ------------------------------------------------------------ projection
PLATE4 = "body Plate\n line_n -0.75 0 0.75 0 4\nend\n"


@pytest.mark.parametrize("use_cg", [False, True])
@pytest.mark.parametrize("geom", ["", PLATE4])
def test_projection_identities(oracle, use_cg, geom):
    # ProjectionSolverTest.cc:70-112: 6x6, L=1, length 2, dt 0.123, Re 100
    dt, Re = 0.123, 100.0
    s = oracle.Sim(6, 6, 1, 2.0, -1.0, -1.0, Re, dt, geom, scheme="euler", tol=1e-12, force_cg=use_cg)
    s.init()
    g = G(6, 6, 1, 2.0, -1.0, -1.0)
    a = g.single_wavenumber(1, 1)
    b = np.full(2 * s.nb, 3.0)
    w, f = s.projection_solve(0, a, b, np.zeros(2 * s.nb))
    lhs = w + (-dt / 2.0) * (1.0 / Re) * oracle.laplacian(6, 6, 1, g.dx, w) + s.B(0, f)
    assert np.abs(lhs - a).max() < 1e-10
    if s.nb:
        assert np.abs(s.C(w) - b).max() < 1e-10


@pytest.mark.parametrize("ngrid", [1, 2])
def test_cholesky_factor_reproduces_M_and_matches_cg(oracle, ngrid):
    geom = "body C\n circle_n 0 0 0.3 24\nend\n"
    s = oracle.Sim(32, 32, ngrid, 2.0, -1.0, -1.0, 100.0, 0.01, geom, tol=1e-13)
    s.init()
    n = 2 * s.nb
    d, L = s.cholesky(0)
    Lf = np.tril(L, -1) + np.diag(d)
    M = np.stack([s.M(0, e) for e in np.eye(n)], axis=1)
    # The factor is built from M's upper triangle (CholeskySolver.cc:93-107).
    assert np.abs(Lf @ Lf.T - np.triu(M) - np.triu(M, 1).T).max() < 1e-10 * np.abs(M).max()
    assert np.all(np.linalg.eigvalsh(0.5 * (M + M.T)) > 0)
    asym = np.abs(M - M.T).max() / np.abs(M).max()
    # Single domain: M = C A^-1 B is symmetric to roundoff.  Multi-domain: the
    # coarse-level Curl is not the adjoint of Curl (VectorOperationsTest.cc:772),
    # so M is slightly non-symmetric and CG and Cholesky answers differ by O(asym).
    assert asym < (1e-14 if ngrid == 1 else 1e-3)
    c = oracle.Sim(32, 32, ngrid, 2.0, -1.0, -1.0, 100.0, 0.01, geom, tol=1e-13, force_cg=True)
    c.init()
    a = G(32, 32, ngrid, 2.0, -1, -1).single_wavenumber(2, 3)
    b = np.linspace(-1, 1, n)
    w1, f1 = s.projection_solve(0, a, b, np.zeros(n))
    w2, f2 = c.projection_solve(0, a, b, np.zeros(n))
    tol = 1e-10 if ngrid == 1 else 20 * asym
    assert np.abs(f1 - f2).max() < tol * np.abs(f1).max()
    assert np.abs(c.M(0, f2) - (c.C(c.Ainv(0, a)) - b)).max() < 1e-12


# -------------------------------------------------------- time stepping
def test_zero_state_is_fixed_point_without_body(oracle):
    s = oracle.Sim(16, 16, 2, 2.0, -1, -1, 100.0, 0.01, "")
    s.init()
    s.advance(3)
    assert np.abs(s.omega()).max() == 0.0


def test_no_body_vortex_decays_like_heat_equation(oracle):
    # A single sine mode of vorticity in a box with no body and no base flow
    # interaction at leading order: omega(t) ~ omega(0) exp(lambda nu t) for
    # small amplitude (the nonlinear term is quadratic in the amplitude).
    nx = 32
    g = G(nx, nx, 1, 1.0, -0.5, -0.5)
    s = oracle.Sim(nx, nx, 1, 1.0, -0.5, -0.5, 50.0, 0.002, "", alpha_deg=0.0)
    s.init()
    w0 = 1e-6 * g.single_wavenumber(1, 1)
    s.set_omega(w0)
    # with the uniform base flow the mode is advected; measure the L2 norm decay
    nsteps = 20
    s.advance(nsteps)
    lam = 2 * (math.cos(math.pi / nx) + math.cos(math.pi / nx) - 2) / g.dx ** 2
    ratio = np.linalg.norm(s.omega()) / np.linalg.norm(w0)
    expect = math.exp(lam / 50.0 * nsteps * 0.002)
    assert abs(ratio - expect) < 2e-3


def test_cylinder_short_run_is_symmetric(oracle):
    s = oracle.Sim(48, 48, 1, 4.0, -2.0, -2.0, 100.0, 0.02, "body C\n circle_n 0 0 0.5 40\nend\n")
    s.init()
    s.advance(3)
    cd, cl = s.forces()
    assert cd > 0.5 and abs(cl) < 1e-10 * cd
    w = s.omega().reshape(47, 47)
    assert np.abs(w + w[:, ::-1]).max() < 1e-10 * np.abs(w).max()   # antisymmetric in y
