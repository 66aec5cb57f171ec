"""GPU parity of the HIP path against the FP64 CPU oracle (same seeded inputs),
called through the C ABI.  Tolerances are stated per test: the DST-based
stages agree to a few ulps x log(n) (relative 1e-12); whole runs are held to
the accuracy gate of the task (vorticity relative L2 <= 1e-5, Cd relative
<= 1e-6, Cl relative to max|Cd| <= 1e-6 for the mirror-symmetric cylinder)."""
import math

import numpy as np
import pytest

from tests.helpers import G, bc_ring, single_wavenumber_2d, plate_geom

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu):
    return gpu.Context(0)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# ------------------------------------------------------------------ DST
@pytest.mark.parametrize("shape", [(3, 7), (11, 23), (199, 199), (399, 399), (255, 511), (1023, 15)])
def test_dst2d_matches_oracle(ctx, oracle, shape):
    print("dst2d", shape, flush=True)
    x = np.random.default_rng(sum(shape)).standard_normal(shape)
    g = ctx.dst2d(x)
    o = oracle.dst2d(x)
    assert np.abs(g - o).max() <= 1e-13 * np.abs(o).max() * math.log2(max(shape) + 2)


# -------------------------------------------------------- elliptic solves
@pytest.mark.parametrize("kind", [0, 1])
def test_elliptic2d_closed_forms(ctx, oracle, kind):
    nx, ny, dx, alpha = 4, 8, 0.1, 0.2
    bc0 = np.zeros(2 * (nx + ny))
    for kx in range(nx):
        for ky in range(ny):
            f = single_wavenumber_2d(kx, ky, nx, ny).ravel()
            u = ctx.elliptic2d_solve(kind, nx, ny, dx, alpha, f)
            Lu = oracle.laplacian2d_bc(nx, ny, dx, u, bc0)
            lhs = Lu if kind == 0 else u + alpha * Lu
            assert np.abs(lhs - f).max() < 1e-10
    ue = lambda i, j: i * dx + 2 * j * dx
    bc = bc_ring(nx, ny, left=lambda j: ue(0, j), top=lambda i: ue(i, ny),
                 right=lambda j: ue(nx, j), bottom=lambda i: ue(i, 0))
    exact = np.array([[ue(i, j) for j in range(1, ny)] for i in range(1, nx)]).ravel()
    f = np.zeros_like(exact) if kind == 0 else exact.copy()
    assert np.abs(ctx.elliptic2d_solve(kind, nx, ny, dx, alpha, f, bc) - exact).max() < 1e-10


@pytest.mark.parametrize("nx,ny,ngrid,kind", [(16, 24, 3, 0), (16, 24, 3, 1), (200, 200, 1, 1),
                                              (400, 400, 4, 0), (400, 400, 4, 1), (512, 512, 5, 1),
                                              (256, 384, 3, 0), (1024, 1024, 3, 0), (1024, 1024, 2, 1),
                                              (1024, 256, 2, 1), (256, 1024, 2, 0), (512, 256, 2, 0),
                                              (256, 512, 3, 1), (1024, 512, 2, 1)])
def test_multidomain_solve_matches_oracle(ctx, oracle, nx, ny, ngrid, kind):
    print("solve", nx, ny, ngrid, kind, flush=True)
    g = G(nx, ny, ngrid, 4.0, -2.0, -2.0)
    alpha = -0.5 * 0.01 / 200.0 if kind == 1 else 0.0
    f = np.random.default_rng(nx + ngrid).standard_normal(ngrid * g.ns)
    u = ctx.elliptic_solve(kind, nx, ny, ngrid, g.dx, alpha, f)
    o = oracle.elliptic_solve(kind, nx, ny, ngrid, g.dx, alpha, f)
    assert rel(u, o) < 1e-12


def test_multidomain_poisson_identity_on_gpu(ctx, oracle):
    g = G(16, 24, 3, 2.0, -1.0, -1.0)
    for kx, ky in [(1, 1), (3, 5), (7, 2)]:
        f = g.single_wavenumber(kx, ky)
        u = ctx.elliptic_solve(0, 16, 24, 3, g.dx, 0.0, f)
        Lu = ctx.laplacian(16, 24, 3, g.dx, u)
        assert np.abs(Lu - oracle.coarsify(16, 24, 3, f)).max() < 1e-10


# ------------------------------------------------------------ operators
@pytest.mark.parametrize("nx,ny,ngrid", [(8, 12, 3), (32, 32, 4), (200, 200, 1)])
def test_operators_match_oracle(ctx, oracle, nx, ny, ngrid):
    print("ops", nx, ny, ngrid, flush=True)
    g = G(nx, ny, ngrid, 2.0, -1.0, -3.0)
    rng = np.random.default_rng(nx * ngrid)
    w = rng.standard_normal(ngrid * g.ns)
    q = rng.standard_normal(ngrid * g.nf)
    assert rel(ctx.curl_flux(nx, ny, ngrid, g.dx, q), oracle.curl_flux(nx, ny, ngrid, g.dx, q)) < 1e-14
    assert rel(ctx.curl_scalar(nx, ny, ngrid, g.dx, w), oracle.curl_scalar(nx, ny, ngrid, w)) < 1e-14
    assert rel(ctx.laplacian(nx, ny, ngrid, g.dx, w), oracle.laplacian(nx, ny, ngrid, g.dx, w)) < 1e-14
    assert rel(ctx.nonlinear(nx, ny, ngrid, g.dx, q, w), oracle.nonlinear(nx, ny, ngrid, g.dx, q, w)) < 1e-13


# -------------------------------------------------------- whole solver
CYL40 = "body C\n circle_n 0 0 0.5 40\nend\n"


def _pair(gpu, oracle, nx, ngrid, Re, dt, geom, length=4.0, tol=1e-7, force_cg=False, scheme="rk3"):
    off = -length / 2
    ctx = gpu.Context(0)
    b = gpu.Batch(ctx, [gpu.Flow(nx, nx, ngrid, length, off, off, Re, dt, geom, scheme=scheme, tol=tol,
                                 force_cg=force_cg)])
    b.init()
    s = oracle.Sim(nx, nx, ngrid, length, off, off, Re, dt, geom, scheme=scheme, tol=tol, force_cg=force_cg)
    s.init()
    return ctx, b, s


def test_cylinder_small_matches_oracle(gpu, oracle):
    ctx, b, s = _pair(gpu, oracle, 64, 2, 100.0, 0.02, CYL40)
    for step in range(8):
        b.advance(1)
        s.advance(1)
        cd_g, cl_g = b.forces()
        cd_o, cl_o = s.forces()
        assert abs(cd_g - cd_o) <= 1e-9 * abs(cd_o), step
        assert abs(cl_g - cl_o) <= 1e-9 * abs(cd_o), step
    assert rel(b.omega(), s.omega()) < 1e-9
    assert rel(b.q(), s.q()) < 1e-9
    hist = b.force_history()
    assert hist.shape == (8, 2) and abs(hist[-1, 0] - s.forces()[0]) <= 1e-9 * abs(s.forces()[0])


def test_wake_crossing_the_fine_grid_edge_matches_oracle(gpu, oracle):
    """A cylinder wake that reaches the edge of the finest grid (t ~ 1):
    from then on the fine level's boundary values come from the coarser
    levels.  The solve omega = omega* - A^-1 c must take the finer level's
    BCs from A^-1 c (EllipticSolver.cc:62-66), not from omega* - A^-1 c;
    the latter drifts off the oracle by 1e-5 within 100 steps."""
    geom = "body C\n circle_n 0 0 0.5 78\nend\n"
    ctx, b, s = _pair(gpu, oracle, 100, 3, 200.0, 0.02, geom)
    worst = 0.0
    for step in range(100):
        b.advance(1)
        s.advance(1)
        cd_g, cl_g = b.forces()
        cd_o, cl_o = s.forces()
        worst = max(worst, abs(cd_g - cd_o) / abs(cd_o))
        assert abs(cd_g - cd_o) <= 1e-9 * abs(cd_o), step
        assert abs(cl_g - cl_o) <= 1e-9 * abs(cd_o), step
    print("max Cd rel err over 100 steps %.3e" % worst, flush=True)
    assert rel(b.omega(), s.omega()) < 1e-9


def test_graph_replay_matches_eager(gpu):
    ctx = gpu.Context(0)
    fl = gpu.Flow(64, 64, 2, 4.0, -2.0, -2.0, 100.0, 0.02, CYL40)
    a = gpu.Batch(ctx, [fl]).init()
    g = gpu.Batch(ctx, [fl]).init().use_graph(True)
    for _ in range(5):
        a.advance(1)
        g.advance(1)
        assert np.array_equal(a.omega(), g.omega())
        assert np.array_equal(a.f(), g.f())


def test_batch_of_uneven_flows_equals_single_flows(gpu):
    ctx = gpu.Context(0)
    f1 = gpu.Flow(64, 64, 2, 4.0, -2.0, -2.0, 100.0, 0.02, CYL40)
    f2 = gpu.Flow(48, 48, 3, 4.0, -2.0, -2.0, 150.0, 0.02, "body P\n line_n -0.4 0.1 0.4 -0.1 10\nend\n")
    both = gpu.Batch(ctx, [f1, f2]).init()
    s1 = gpu.Batch(ctx, [f1]).init()
    s2 = gpu.Batch(ctx, [f2]).init()
    both.advance(3)
    s1.advance(3)
    s2.advance(3)
    assert np.array_equal(both.omega(0), s1.omega())
    assert np.array_equal(both.omega(1), s2.omega())
    assert np.array_equal(both.f(1), s2.f())


def test_batch_of_512_grids_takes_register_transform(gpu):
    """Four 512^2 flows in one batch give the 1024-point passes 1024 row pairs
    per launch, enough for the register-held transform; one flow alone (256
    pairs) runs the LDS passes, which the multi-domain solve tests hold to the
    oracle.  The two transforms round differently, so the runs agree to
    1e-11 relative, not bit for bit."""
    ctx = gpu.Context(0)
    flows = [gpu.Flow(512, 512, 2, 4.0, -2.0, -2.0, Re, 0.01, CYL40) for Re in (100.0, 150.0, 200.0, 250.0)]
    print("batch of", len(flows), "512^2 flows", flush=True)
    both = gpu.Batch(ctx, flows).init()
    both.advance(2)
    for k, f in enumerate(flows):
        one = gpu.Batch(ctx, [f]).init()
        one.advance(2)
        assert rel(both.omega(k), one.omega()) < 1e-11, k
        assert rel(both.f(k), one.f()) < 1e-11, k


def test_local_helmholtz_equals_dst_solve(gpu):
    """Inside the moving-plate force solve, A^-1 by Jacobi sweeps over boxes
    around the plate (the default) and by the DST solve on every level give
    the same run: both solve the same system, to rounding, so the CG runs
    agree to within their tolerance's effect (forces 1e-8 relative to max|F|,
    vorticity 1e-8 relative L2 after 6 steps), and the local path was taken."""
    ctx = gpu.Context(0)
    geom = ("body Plate\n  line_n -0.5 0 0.5 0 33\n  center 0 0\n"
            "  motion pitchplunge 0.2617993877991494 0.4 0 0.2 0.4 0\nend\n")
    fl = gpu.Flow(128, 128, 3, 4.0, -2.0, -2.0, 300.0, 0.005, geom, tol=1e-10)
    loc = gpu.Batch(ctx, [fl]).init()
    dst = gpu.Batch(ctx, [fl]).init().set_local_helmholtz(False)
    for _ in range(6):
        loc.advance(1)
        dst.advance(1)
    hl, hd = loc.force_history(), dst.force_history()
    assert np.abs(hl - hd).max() <= 1e-8 * np.abs(hd).max()
    assert rel(loc.omega(), dst.omega()) < 1e-8
    sl, sd = loc.solver_stats(), dst.solver_stats()
    print("local", sl, "dst", sd, flush=True)
    assert sl["local_helmholtz_m_applications"] == sl["m_applications"] > 0
    assert sd["local_helmholtz_m_applications"] == 0


def test_local_helmholtz_equals_dst_solve_register_passes(gpu):
    """The same comparison at 1024^2 (2048-point register-held passes), where
    the L^-1 solve after the local A^-1 also transforms only the rows whose
    psi is read (level 0 near the plate, the coarser level's ring rows) and
    sums the ring values of the others directly."""
    ctx = gpu.Context(0)
    geom = ("body Plate\n  line_n -0.5 0 0.5 0 129\n  center 0 0\n"
            "  motion pitchplunge 0.2617993877991494 0.4 0 0.2 0.4 0\nend\n")
    fl = gpu.Flow(1024, 1024, 3, 8.0, -4.0, -4.0, 500.0, 0.001, geom, tol=1e-10)
    loc = gpu.Batch(ctx, [fl]).init()
    dst = gpu.Batch(ctx, [fl]).init().set_local_helmholtz(False)
    for step in range(3):
        print("step", step, flush=True)
        loc.advance(1)
        dst.advance(1)
    hl, hd = loc.force_history(), dst.force_history()
    assert np.abs(hl - hd).max() <= 1e-8 * np.abs(hd).max()
    assert rel(loc.omega(), dst.omega()) < 1e-8
    sl = loc.solver_stats()
    print("local", sl, "dst", dst.solver_stats(), flush=True)
    assert sl["local_helmholtz_m_applications"] == sl["m_applications"] > 0


def test_local_helmholtz_equals_dst_solve_uneven_batch(gpu):
    """Two moving flows of different grids, level counts and motions in one
    batch (their force solves converge at different iterations, so the local
    sweeps and the restricted L^-1 passes skip one flow while the other
    iterates): local and DST paths agree as for one flow."""
    ctx = gpu.Context(0)
    g1 = ("body Plate\n  line_n -0.5 0 0.5 0 33\n  center 0 0\n"
          "  motion pitchplunge 0.2617993877991494 0.4 0 0.2 0.4 0\nend\n")
    g2 = ("body Plate\n  line_n -0.4 0.1 0.4 -0.1 25\n  center 0 0\n"
          "  motion pitchplunge 0.1 0.7 0.2 0.1 0.5 0.3\nend\n")
    f1 = gpu.Flow(128, 128, 3, 4.0, -2.0, -2.0, 300.0, 0.005, g1, tol=1e-10)
    f2 = gpu.Flow(96, 128, 2, 3.0, -1.5, -2.0, 200.0, 0.005, g2, tol=1e-9)
    loc = gpu.Batch(ctx, [f1, f2]).init()
    dst = gpu.Batch(ctx, [f1, f2]).init().set_local_helmholtz(False)
    for _ in range(4):
        loc.advance(1)
        dst.advance(1)
    for k in range(2):
        hl, hd = loc.force_history(k), dst.force_history(k)
        assert np.abs(hl - hd).max() <= 1e-7 * np.abs(hd).max(), k
        assert rel(loc.omega(k), dst.omega(k)) < 1e-7, k
    assert loc.solver_stats()["local_helmholtz_m_applications"] > 0


def test_local_helmholtz_equals_dst_solve_tandem_plates(gpu):
    """One case of the 16-case tandem-plate batch (three moving plates,
    1024^2, 6 levels): boxes that span several bodies and all six levels of
    the register-held restricted passes, local against DST for 2 steps."""
    from ibpm_amd import workloads as W
    ctx = gpu.Context(0)
    fl = W.tandem_16x1024x6(ncases=1)[0]
    loc = gpu.Batch(ctx, [fl]).init()
    dst = gpu.Batch(ctx, [fl]).init().set_local_helmholtz(False)
    for step in range(2):
        print("tandem step", step, flush=True)
        loc.advance(1)
        dst.advance(1)
    hl, hd = loc.force_history(), dst.force_history()
    assert np.abs(hl - hd).max() <= 1e-8 * np.abs(hd).max()
    assert rel(loc.omega(), dst.omega()) < 1e-8
    sl = loc.solver_stats()
    assert sl["local_helmholtz_m_applications"] == sl["m_applications"] > 0


def test_local_helmholtz_falls_back_near_the_boundary(gpu):
    """A plate whose grown box would reach the fine grid's first rows (it sits
    4 cells from the edge): the planner turns the local path off for every
    substep, so no M application takes it, and the run equals the DST run bit
    for bit (128^2 grids run the LDS passes, which the pass restrictions leave
    whole)."""
    ctx = gpu.Context(0)
    geom = ("body Plate\n  line_n -1.6 -1.85 -0.6 -1.85 33\n  center -1.1 -1.85\n"
            "  motion pitchplunge 0.1 0.4 0 0.02 0.4 0\nend\n")
    fl = gpu.Flow(128, 128, 2, 4.0, -2.0, -2.0, 300.0, 0.005, geom, tol=1e-10)
    loc = gpu.Batch(ctx, [fl]).init()
    dst = gpu.Batch(ctx, [fl]).init().set_local_helmholtz(False)
    for _ in range(3):
        loc.advance(1)
        dst.advance(1)
    assert loc.solver_stats()["local_helmholtz_m_applications"] == 0
    assert np.array_equal(loc.force_history(), dst.force_history())
    assert np.array_equal(loc.omega(), dst.omega())


def test_moving_plate_cg_matches_oracle(gpu, oracle):
    geom = plate_geom(16, 1.0, 0.2, 0.2, 0.4)   # spacing 1/15 ~ dx = 1/16
    ctx, b, s = _pair(gpu, oracle, 64, 3, 300.0, 0.01, geom, tol=1e-10)
    for step in range(4):
        b.advance(1)
        s.advance(1)
        cd_g, cl_g = b.forces()
        cd_o, cl_o = s.forces()
        assert abs(cd_g - cd_o) <= 1e-7 * max(abs(cd_o), abs(cl_o)), step
        assert abs(cl_g - cl_o) <= 1e-7 * max(abs(cd_o), abs(cl_o)), step
        it_g, it_o = b.cg_iters(), s.cg_iters()
        assert np.all(np.abs(it_g - it_o) <= 1), (it_g, it_o)
    assert rel(b.omega(), s.omega()) < 1e-7


@pytest.mark.parametrize("motion", [
    "fixedvel 0.3 -0.1 0.5",
    "eldredge 40 11 0 0.02 0.03 0.05",
    "eldredge2 0.3 11 0 0.02 0.03 0.05",
    "eldredgecombined2 0.3 11 0 0.02 0.03 0.05 30 9 0 0.02 0.03 0.05",
    "lagstep2 20 0.02 0.01 0.005",
    "sigmoidalstep 30 0.04 0.0",
    "motionfile",
])
def test_other_motions_match_oracle(gpu, oracle, tmp_path, motion):
    """A plate under each remaining Motion type for 4 steps: the library's
    host kinematics (stateful ones included) feed the GPU path exactly as the
    oracle's feed its solver, so the forces agree to the CG tolerance."""
    if motion == "motionfile":
        path = tmp_path / "plate.motion"
        t = np.linspace(0.0, 0.05, 6)
        rows = np.stack([t, 0.5 * t, -0.2 * t, 3.0 * t, 0.5 + 0 * t, -0.2 + 0 * t, 3.0 + 0 * t], 1)
        path.write_text("%d\n" % len(rows) + "".join(" ".join("%.17g" % v for v in r) + "\n" for r in rows))
        motion = "motionfile %s" % path
    geom = "body P\n line_n -0.5 0 0.5 0 16\n center 0.25 0\n motion %s\nend\n" % motion
    ctx, b, s = _pair(gpu, oracle, 64, 3, 300.0, 0.01, geom, tol=1e-10)
    for step in range(4):
        b.advance(1)
        s.advance(1)
        cd_g, cl_g = b.forces()
        cd_o, cl_o = s.forces()
        scale = max(abs(cd_o), abs(cl_o))
        assert abs(cd_g - cd_o) <= 1e-7 * scale, (motion, step)
        assert abs(cl_g - cl_o) <= 1e-7 * scale, (motion, step)
    assert rel(b.omega(), s.omega()) < 1e-7


@pytest.mark.parametrize("motion", [
    "pitchplunge 0.3 0.5 0 0.1 0.5 0.2",
    "fixedvel -0.2 0.1 0.3",
    "eldredgecombined2 0.3 11 0 0.02 0.03 0.05 30 9 0 0.02 0.03 0.05",
])
def test_unsteady_base_flow_matches_oracle(gpu, oracle, motion):
    """-ubf: the body stays put (Cholesky path) and the per-substep base flux
    carries its motion; per-step forces are rotated by the base-flow angle.
    Graph replay is not used for such a batch, so asking for it changes
    nothing."""
    geom = "body P\n line_n -0.5 0 0.5 0 16\n center 0.25 0\n motion %s\nend\n" % motion
    ctx = gpu.Context(0)
    fl = gpu.Flow(64, 64, 3, 4.0, -2.0, -2.0, 300.0, 0.01, geom, alpha=5.0, ubf=True)
    b = gpu.Batch(ctx, [fl]).init()
    g = gpu.Batch(ctx, [fl]).init().use_graph(True)
    s = oracle.Sim(64, 64, 3, 4.0, -2.0, -2.0, 300.0, 0.01, geom, alpha_deg=5.0, ubf=True)
    assert s.ubf == 1
    s.init()
    want = []
    for step in range(6):
        b.advance(1)
        g.advance(1)
        s.advance(1)
        cd_o, cl_o = s.forces()
        want.append((cd_o, cl_o))
        cd_g, cl_g = b.forces()
        scale = max(abs(cd_o), abs(cl_o))
        assert abs(cd_g - cd_o) <= 1e-9 * scale, (motion, step)
        assert abs(cl_g - cl_o) <= 1e-9 * scale, (motion, step)
        assert np.array_equal(b.omega(), g.omega())
    assert np.abs(b.force_history() - np.array(want)).max() <= 1e-9 * np.abs(want).max()
    assert rel(b.omega(), s.omega()) < 1e-9
    assert b.base_flow()[2] is True


def test_unsteady_base_flow_needs_a_single_body(gpu):
    ctx = gpu.Context(0)
    two = "body P\n line_n -0.5 0 0.5 0 16\n motion fixedvel 0.1 0 0\nend\nbody Q\n point 0.5 0.5\nend\n"
    with pytest.raises(gpu.IbpmError, match="single moving body"):
        gpu.Batch(ctx, [gpu.Flow(64, 64, 2, 4.0, -2.0, -2.0, 100.0, 0.01, two, ubf=True)])


def test_restart_resume_is_bitwise_identical(gpu):
    """State after 3 steps loaded into a fresh batch (restart), 3 more steps:
    identical to 6 straight steps (Cholesky/RK3: no hidden multistep state)."""
    ctx = gpu.Context(0)
    fl = gpu.Flow(64, 64, 2, 4.0, -2.0, -2.0, 100.0, 0.02, CYL40)
    a = gpu.Batch(ctx, [fl]).init()
    a.advance(3)
    w, f = a.omega(), a.f()
    b = gpu.Batch(ctx, [fl]).init()
    b.set_state(w, f, 3 * 0.02, 3)
    a.advance(3)
    b.advance(3)
    assert np.array_equal(a.omega(), b.omega())
    assert np.array_equal(a.f(), b.f())


def test_rhs_matches_oracle_composition(ctx, oracle):
    """a = omega + dt (0.5 nu (an + bn) Lap omega + an N + bn N_prev)
    (IBSolver.cc:167-183) from the oracle's pinned Laplacian and N, with and
    without a previous N (without one the reference uses N itself)."""
    nx, ny, ngrid, dx = 32, 24, 3, 0.125
    rng = np.random.default_rng(11)
    q = rng.standard_normal(ngrid * (2 * nx * ny + nx + ny))
    w = rng.standard_normal(ngrid * (nx - 1) * (ny - 1))
    npv = rng.standard_normal(w.size)
    dt, nu, an, bn = 0.01, 1 / 200.0, 5 / 12, -17 / 60
    lap = oracle.laplacian(nx, ny, ngrid, dx, w)
    n = oracle.nonlinear(nx, ny, ngrid, dx, q, w)
    for prev in (npv, None):
        a, n_g = ctx.rhs(nx, ny, ngrid, dx, dt, nu, an, bn, q, w, prev)
        p = npv if prev is not None else n
        a_o = w + dt * (0.5 * nu * (an + bn) * lap + an * n + bn * p)
        assert rel(n_g, n) < 1e-13
        assert rel(a, a_o) < 1e-13


def test_flux_from_vorticity_matches_oracle_refresh(ctx, oracle):
    """q = Curl(Poisson(-coarsify omega)) + q0, NavierStokesModel::computeFlux,
    against the oracle's refreshState on the same vorticity."""
    nx, ngrid = 48, 3
    s = oracle.Sim(nx, nx, ngrid, 4.0, -2.0, -2.0, 100.0, 0.02, CYL40, alpha_deg=20.0)
    s.init()
    w = np.random.default_rng(12).standard_normal(ngrid * (nx - 1) ** 2)
    s.set_omega(w)
    al = math.radians(20.0)
    q = ctx.flux_from_vorticity(nx, nx, ngrid, 4.0 / nx, w, math.cos(al), math.sin(al))
    assert rel(q, s.q()) < 1e-12


@pytest.mark.parametrize("geom", [CYL40, plate_geom(21, 1.0, 0.2, 0.2, 0.4)])
def test_regularizer_matches_oracle(gpu, ctx, oracle, geom):
    """toFlux and toBoundary (Regularizer.cc:92-128) against the oracle's
    association lists (pinned bit-exactly to the reference's), before and
    after the points move (update)."""
    nx, ny, ngrid, length, x0, y0 = 64, 64, 2, 4.0, -2.0, -2.0
    dx = length / nx
    rng = np.random.default_rng(13)
    xy, _ = oracle.geom_points(geom, 0.0)
    reg = gpu.Regularizer(ctx, nx, ny, ngrid, length, x0, y0, xy)
    for t in (0.0, 0.37):
        xy, _ = oracle.geom_points(geom, t)
        if t:
            reg.update(xy)
        nb = len(xy) // 2
        assoc = oracle.reg_build(nx, ny, dx, x0, y0, xy[:nb], xy[nb:])
        f = rng.standard_normal(2 * nb)
        q = rng.standard_normal(ngrid * (2 * nx * ny + nx + ny))
        assert rel(reg.to_flux(f), oracle.reg_toflux(nx, ny, ngrid, dx, assoc, f)) < 1e-14
        assert rel(reg.to_boundary(q), oracle.reg_toboundary(nb, dx, assoc, q)) < 1e-14


@pytest.mark.parametrize("motion", ["eldredge2 0.3 11 0 0.02 0.03 0.05", "pitchplunge 0.3 0.5 0 0.1 0.5 0.2"])
def test_set_state_after_advancing_undoes_the_move_made_ahead(gpu, motion):
    """Each substep moves the bodies of the next one ahead of time; set_state
    at the current timestep undoes that move, so that a batch given its own
    state back continues exactly like one that was not (for Eldredge2, whose
    motion integrates over successive calls, a kept move would shift it)."""
    geom = "body P\n line_n -0.5 0 0.5 0 16\n center 0.25 0\n motion %s\nend\n" % motion
    ctx = gpu.Context(0)
    fl = gpu.Flow(64, 64, 3, 4.0, -2.0, -2.0, 300.0, 0.01, geom, tol=1e-10)
    a = gpu.Batch(ctx, [fl]).init()
    b = gpu.Batch(ctx, [fl]).init()
    a.advance(2)
    b.advance(2)
    a.set_state(a.omega(), a.f(), a.time, 2)
    a.advance(2)
    b.advance(2)
    assert np.array_equal(a.omega(), b.omega())
    assert np.array_equal(a.force_history(), b.force_history())
