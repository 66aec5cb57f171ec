"""The learned CG warm start: the network's loss and gradients computed by the
HIP training kernels against a CPU FP64 reference (tests/nn_reference.py),
the reference itself against finite differences, and an ablation of the force
solver on a pitching/plunging plate: the trained network's start and the
preconditioned solver against the reference's plain CG, all converging to the
same absolute residual tolerance."""
import numpy as np
import pytest

from tests import nn_reference as R
from tests.helpers import plate_geom


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(R.NP) * 0.3
    X = rng.standard_normal((40, R.NI))
    Y = rng.standard_normal((40, 2))
    loss, g = R.loss_grad(p, X, Y)
    h = 1e-6
    for q in list(range(0, R.NP, 37)) + [R.NP - 1, R.NP - 3]:
        e = np.zeros(R.NP)
        e[q] = h
        fd = (R.loss_grad(p + e, X, Y)[0] - R.loss_grad(p - e, X, Y)[0]) / (2 * h)
        assert abs(fd - g[q]) < 1e-7 * max(1, abs(fd))


def test_planted_gradient_error_is_detected():
    # A dropped tanh' factor in the reference would fail the check above.
    rng = np.random.default_rng(1)
    p = rng.standard_normal(R.NP) * 0.3
    X = rng.standard_normal((10, R.NI))
    Y = rng.standard_normal((10, 2))
    W1, b1, W2, b2 = R.unpack(p)
    H, Yp = R.forward(p, X)
    D2 = 2.0 * (Yp - Y) / 10
    wrong = (D2 @ W2).T @ X   # missing (1 - h^2)
    _, g = R.loss_grad(p, X, Y)
    assert np.abs(wrong.ravel() - g[:R.NH * R.NI]).max() > 1e-3


@pytest.mark.gpu
def test_gpu_loss_and_gradient_match_cpu_fp64(gpu):
    ctx = gpu.Context(0)
    rng = np.random.default_rng(2)
    for n in (1, 37, 5000):
        p = rng.standard_normal(R.NP) * 0.4
        X = rng.standard_normal((n, R.NI))
        Y = rng.standard_normal((n, 2))
        loss_g, grad_g = ctx.nn_loss_grad(p, X, Y)
        loss_c, grad_c = R.loss_grad(p, X, Y)
        assert abs(loss_g - loss_c) <= 1e-12 * abs(loss_c)
        assert np.abs(grad_g - grad_c).max() <= 1e-12 * np.abs(grad_c).max()


@pytest.mark.gpu
def test_force_solver_ablation(gpu):
    """Moving plate, same steps, same tolerance (CG residual 1e-10): CG
    iterations per step of the reference solver (plain CG started from the
    previous force), a zero start, the trained network's start, and the
    preconditioned solver (stale FP64 inverse of M; plus the low-rank update
    from the previous solve's search directions; plus the network's start).

    The point forces on this plate change by about 95% (relative L2) from one
    step to the next, so no per-point start, learned or extrapolated, can save
    much: the network must learn (loss halves) and must not cost more than 10%
    over the reference start.  The preconditioner must cut iterations to at
    most 35% of plain CG, and total M applications, preconditioner builds
    included, to at most half.  Every run converges to the same residual
    tolerance, so all forces agree to 1e-7 of their scale."""
    geom = plate_geom(33, 1.0, 0.3, 0.25, 0.4)
    flow = gpu.Flow(128, 128, 3, 4.0, -2.0, -2.0, 300.0, 0.005, geom, tol=1e-10)
    ctx = gpu.Context(0)

    def batch():
        return gpu.Batch(ctx, [flow]).init()
    runs = {
        "reference": batch(),
        "zero start": batch().set_warm_start("zero"),
        "learned start": batch().collect(True, 200000),
        "stale inverse": batch().set_precond("stale"),
        "low-rank": batch().set_precond("lowrank", 8),
        "low-rank + learned start": batch().set_precond("lowrank", 8),
    }
    ntrain, ntest = 40, 20
    for _ in range(ntrain):
        for b in runs.values():
            b.advance(1)
    lrn = runs["learned start"]
    assert lrn.num_samples() > 0
    loss0 = lrn.train(iters=1, lr=0.0)
    loss = lrn.train(iters=800, lr=3e-3)
    print("samples", lrn.num_samples(), "loss before %.3e after %.3e" % (loss0, loss), flush=True)
    assert loss < 0.5 * loss0
    lrn.collect(False).set_warm_start("learned")
    runs["low-rank + learned start"].nn_params(lrn.nn_params())
    runs["low-rank + learned start"].set_warm_start("learned")
    before = {k: b.solver_stats() for k, b in runs.items()}
    its = {k: [] for k in runs}
    for _ in range(ntest):
        for k, b in runs.items():
            b.advance(1)
            its[k].append(int(np.sum(b.cg_iters())))
    mean = {k: float(np.mean(v)) for k, v in its.items()}
    cost = {}
    for k, b in runs.items():
        st = b.solver_stats()
        cost[k] = (st["m_applications"] + st["build_m_applications"]) / (ntrain + ntest)
        print("%-26s CG iterations/step %6.2f  M applications/step (whole run, builds included) %7.2f  %s"
              % (k, mean[k], cost[k], {q: st[q] - before[k][q] for q in ("builds", "fallbacks")}), flush=True)
    ref = mean["reference"]
    assert mean["learned start"] <= 1.1 * ref
    assert mean["stale inverse"] <= 0.35 * ref
    assert mean["low-rank"] <= 0.35 * ref
    assert mean["low-rank + learned start"] <= 0.35 * ref
    assert cost["low-rank"] <= 0.5 * cost["reference"]
    cd_r, cl_r = runs["reference"].forces()
    scale = max(abs(cd_r), abs(cl_r))
    w_r = runs["reference"].omega()
    for k, b in runs.items():
        cd, cl = b.forces()
        assert abs(cd - cd_r) <= 1e-7 * scale and abs(cl - cl_r) <= 1e-7 * scale, k
        assert np.linalg.norm(b.omega() - w_r) <= 1e-8 * np.linalg.norm(w_r), k
