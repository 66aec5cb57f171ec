"""The reduced-precision elliptic path (FP32 transforms and transposed
intermediates; right-hand sides, boundary terms, eigenvalue scaling and
fields in FP64) against the FP64 path and the oracle's fixtures.

Per solve, FP32 transforms perturb the multi-domain solution by about 4e-7
(Helmholtz) to 7e-6 (Poisson of a random right-hand side) relative L2.  Over
whole runs these perturbations reach the forces: they exceed the 1e-6 Cd/Cl
gate (cylinder_200: about 7e-5 after 250 steps), which is why the FP64 path
is the default.  The tests pin the error sizes so that a change in the FP32
kernels that breaks them (a wrong index or twiddle) fails by orders of
magnitude, and check that switching back to FP64 restores the FP64 results
bit for bit."""
import os

import numpy as np
import pytest

from ibpm_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("nx,ngrid", [(200, 1), (256, 3), (400, 4), (512, 5), (1024, 6)])
def test_fp32_solve_error_per_solve(gpu, nx, ngrid):
    ctx = gpu.Context(0)
    rng = np.random.default_rng(nx)
    dx = 4.0 / nx
    f = rng.standard_normal(ngrid * (nx - 1) ** 2)
    for kind, alpha, bound in ((0, 0.0, 2e-5), (1, -1.3e-5, 2e-6)):
        print("fp32 solve nx=%d ngrid=%d kind=%d" % (nx, ngrid, kind), flush=True)
        u64 = ctx.elliptic_solve_precision(kind, nx, nx, ngrid, dx, alpha, f, 0)
        u32 = ctx.elliptic_solve_precision(kind, nx, nx, ngrid, dx, alpha, f, 1)
        assert np.array_equal(u64, ctx.elliptic_solve(kind, nx, nx, ngrid, dx, alpha, f))
        err = np.linalg.norm(u32 - u64) / np.linalg.norm(u64)
        print("  rel L2 %.2e" % err, flush=True)
        assert 1e-9 < err <= bound


def test_fp32_step_solves_whole_run_error(gpu):
    """cylinder_200 for 250 steps with the step's solves in FP32: the error
    against the oracle stays within 1e-3 (Cd) and 1e-4 (vorticity), and the
    same batch switched back to FP64 reproduces the FP64 run bitwise."""
    g = np.load(os.path.join(GOLD, "cylinder_200.npz"))
    flow = W.cylinder_200()[0]
    ctx = gpu.Context(0)
    b = gpu.Batch(ctx, [flow]).init().use_graph(True).set_precision("fp32_step")
    b.advance(int(g["nsteps"]))
    h = b.force_history()
    cd = np.max(np.abs(h[:, 0] - g["forces"][:, 0]) / np.abs(g["forces"][:, 0]))
    w0 = b.omega()[:199 * 199].reshape(199, 199)
    w_err = np.linalg.norm(w0 - g["omega0"]) / np.linalg.norm(g["omega0"])
    print("fp32_step cylinder_200: Cd %.2e, omega %.2e" % (cd, w_err), flush=True)
    assert 1e-8 < cd <= 1e-3
    assert w_err <= 1e-4
    b.close()
    # precision switched after the graph was captured: FP64 again, bitwise
    ref = gpu.Batch(ctx, [flow]).init().use_graph(True)
    ref.advance(20)
    sw2 = gpu.Batch(ctx, [flow]).init().use_graph(True)
    sw2.advance(2)
    sw2.set_precision("fp32_step").set_precision("fp64")
    sw2.advance(18)
    assert np.array_equal(sw2.omega(), ref.omega())
    assert np.array_equal(sw2.force_history(), ref.force_history())


def test_fp32_force_solve_moving_plate(gpu):
    """plate_512x5 for 12 steps with every solve in FP32 (mode 2, including the
    force solve's M applications): CG still converges every substep, and the
    forces stay within 1e-3 of the oracle's."""
    g = np.load(os.path.join(GOLD, "plate_512x5_short.npz"))
    flow = W.plate_512x5()[0]
    ctx = gpu.Context(0)
    b = gpu.Batch(ctx, [flow]).init().set_precond("lowrank", 8).set_precision("fp32")
    b.advance(int(g["nsteps"]))
    h = b.force_history()
    cd = np.max(np.abs(h[:, 0] - g["forces"][:, 0]) / np.abs(g["forces"][:, 0]))
    cl = np.max(np.abs(h[:, 1] - g["forces"][:, 1])) / np.max(np.abs(g["forces"][:, 1]))
    print("fp32 plate_512x5: Cd %.2e Cl %.2e" % (cd, cl), flush=True)
    assert cd <= 1e-3 and cl <= 1e-3
    assert b.solver_stats()["fallbacks"] == 0


@pytest.mark.parametrize("nx,ngrid", [(200, 1), (256, 3), (400, 4), (512, 5)])
def test_gemm_dst_error_per_solve(gpu, nx, ngrid):
    """The split-BF16 GEMM-DST on MFMA (precision 2) against the FP64 solve:
    split products with FP32 accumulation, about 1e-5 relative L2."""
    ctx = gpu.Context(0)
    rng = np.random.default_rng(nx + 1)
    dx = 4.0 / nx
    f = rng.standard_normal(ngrid * (nx - 1) ** 2)
    for kind, alpha in ((0, 0.0), (1, -1.3e-5)):
        print("gemm solve nx=%d ngrid=%d kind=%d" % (nx, ngrid, kind), flush=True)
        u64 = ctx.elliptic_solve_precision(kind, nx, nx, ngrid, dx, alpha, f, 0)
        ug = ctx.elliptic_solve_precision(kind, nx, nx, ngrid, dx, alpha, f, 2)
        err = np.linalg.norm(ug - u64) / np.linalg.norm(u64)
        print("  rel L2 %.2e" % err, flush=True)
        assert 1e-8 < err <= 5e-5


def test_gemm_dst_only_for_square_grids_up_to_511(gpu):
    ctx = gpu.Context(0)
    f = np.zeros(1023 * 1023)
    with pytest.raises(gpu.IbpmError, match="GEMM-DST"):
        ctx.elliptic_solve_precision(0, 1024, 1024, 1, 4.0 / 1024, 0.0, f, 2)


def test_autotuner_and_gemm_step_run(gpu):
    """cylinder_200 (N = 199): the autotuner times both reduced-precision
    transforms and keeps the faster; the GEMM-DST forced on the step's solves
    runs 250 steps within 1e-2 of the oracle's Cd (2.5e-3 measured: its
    per-solve error is about 1e-5), and switching back to FP64 restores the
    FP64 run bitwise."""
    g = np.load(os.path.join(GOLD, "cylinder_200.npz"))
    flow = W.cylinder_200()[0]
    ctx = gpu.Context(0)
    b = gpu.Batch(ctx, [flow]).init().set_precision("reduced_tuned")
    rep = b.precision_report()
    print("autotuner:", rep, flush=True)
    assert list(rep) == [199]
    use_gemm, t_radix, t_gemm = rep[199]
    assert t_radix > 0 and t_gemm > 0 and use_gemm == (t_gemm < t_radix)
    b.close()
    b = gpu.Batch(ctx, [flow]).init().use_graph(True).set_precision("gemm_step")
    try:
        assert b.precision_report()[199][0]
        b.advance(int(g["nsteps"]))
        h = b.force_history()
    finally:
        b.close()
    cd = np.max(np.abs(h[:, 0] - g["forces"][:, 0]) / np.abs(g["forces"][:, 0]))
    print("gemm_step cylinder_200: Cd %.2e" % cd, flush=True)
    assert 1e-8 < cd <= 1e-2
    ref = gpu.Batch(ctx, [flow]).init().use_graph(True)
    ref.advance(10)
    sw = gpu.Batch(ctx, [flow]).init().use_graph(True)
    sw.advance(2)
    sw.set_precision("gemm_step").set_precision("fp64")
    sw.advance(8)
    assert np.array_equal(sw.omega(), ref.omega())
