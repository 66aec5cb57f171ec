"""The learned CG warm start: the network's loss and gradients computed by the
HIP training kernels against a CPU FP64 reference (tests/nn_reference.py),
the reference itself against finite differences, and an ablation on a
pitching/plunging plate showing what the trained network gains over the
reference's own warm start (previous force) while CG still converges to the
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
def test_learned_warm_start_ablation(gpu):
    """Same plate, same steps: reference warm start vs the trained network.
    The network must reduce CG iterations; forces agree to the CG tolerance."""
    geom = plate_geom(33, 1.0, 0.3, 0.25, 0.4)
    flow = gpu.Flow(128, 128, 3, 4.0, -2.0, -2.0, 300.0, 0.005, geom, tol=1e-10)
    ctx = gpu.Context(0)
    ref = gpu.Batch(ctx, [flow]).init()
    lrn = gpu.Batch(ctx, [flow]).init().collect(True, 200000)
    zero = gpu.Batch(ctx, [flow]).init().set_warm_start("zero")
    ntrain, ntest = 40, 20
    for _ in range(ntrain):
        ref.advance(1)
        lrn.advance(1)
    assert lrn.num_samples() > 0
    loss0 = lrn.train(iters=1, lr=0.0)
    loss = lrn.train(iters=800, lr=3e-3)
    print("samples", lrn.num_samples(), "loss before %.3e after %.3e" % (loss0, loss), flush=True)
    assert loss < 0.5 * loss0
    lrn.collect(False).set_warm_start("learned")
    it_ref, it_lrn, it_zero = [], [], []
    for _ in range(ntest):
        ref.advance(1)
        lrn.advance(1)
        it_ref.append(ref.cg_iters())
        it_lrn.append(lrn.cg_iters())
    for _ in range(3):
        zero.advance(1)
        it_zero.append(zero.cg_iters())
    mr, ml, mz = np.mean(it_ref), np.mean(it_lrn), np.mean(it_zero)
    print("CG iterations per substep: zero %.1f, previous force (reference) %.1f, learned %.1f" % (mz, mr, ml),
          flush=True)
    assert ml < mr
    cd_r, cl_r = ref.forces()
    cd_l, cl_l = lrn.forces()
    scale = max(abs(cd_r), abs(cl_r))
    assert abs(cd_l - cd_r) <= 1e-7 * scale and abs(cl_l - cl_r) <= 1e-7 * scale
