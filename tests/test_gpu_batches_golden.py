"""Parity of the two batched workloads on the GPU, run whole and with the
bench's settings, against fixtures the FP64 oracle wrote for some of their
flows (tests/golden/make_golden.py):

  * mixed_64x3: the 64 stationary flows (graph replay) for 40 steps; one
    flow of each grid size (256, 384, 512) and body (cylinder, ellipse,
    plate) is checked;
  * tandem_16x1024x6: the 16 moving tandem-plate cases (preconditioned CG,
    local Helmholtz solves, restricted L^-1 passes) for 30 steps; case 0 is
    checked with P from the kernel table refreshed before every solve (the
    bench's solver; reference and learned warm starts) and with P from M's
    finest-level form plus 8 pairs.

Gates: Cd relative error <= 1e-6 at every step; Cl the same, relative to
max|Cd| for the mirror-symmetric cylinders at zero incidence (their Cl is
roundoff-seeded) and to max|Cl| otherwise; level-0 vorticity (every 4th
node) relative L2 <= 1e-5 and per-level norms to 1e-5 after the run.
"""
import os

import numpy as np
import pytest

from ibpm_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MIXED_FLOWS = (2, 5, 4, 20, 3, 11, 0, 1, 51)


def _fixture(name):
    path = os.path.join(GOLD, name + ".npz")
    if not os.path.exists(path):
        pytest.skip("fixture %s not generated (tests/golden/make_golden.py %s)" % (path, name))
    return np.load(path)


def _check(b, k, flow, g, label):
    hist = b.force_history(k)
    ref = g["forces"]
    assert hist.shape == ref.shape, (hist.shape, ref.shape)
    cd_err = np.abs(hist[:, 0] - ref[:, 0]) / np.abs(ref[:, 0])
    sym = "circle" in flow.geom
    cl_scale = np.abs(ref[:, 0]).max() if sym else np.abs(ref[:, 1]).max()
    cl_err = np.abs(hist[:, 1] - ref[:, 1]) / cl_scale
    w = b.omega(k)
    ns = (flow.nx - 1) * (flow.ny - 1)
    s = int(g["stride"])
    w0 = w[:ns].reshape(flow.nx - 1, flow.ny - 1)[::s, ::s]
    w_err = np.linalg.norm(w0 - g["omega0"]) / np.linalg.norm(g["omega0"])
    norms = np.array([np.linalg.norm(w[l * ns:(l + 1) * ns]) for l in range(flow.ngrid)])
    print("%s flow %d: Cd %.2e Cl %.2e omega %.2e" % (label, k, cd_err.max(), cl_err.max(), w_err), flush=True)
    assert cd_err.max() <= 1e-6
    assert cl_err.max() <= 1e-6
    assert w_err <= 1e-5
    assert np.all(np.abs(norms - g["omega_norms"]) <= 1e-5 * g["omega_norms"])


def test_mixed_batch_matches_oracle_fixtures(gpu):
    fixtures = {k: _fixture("mixed_64x3_f%d" % k) for k in MIXED_FLOWS}
    flows = W.mixed_64x3()
    nsteps = int(fixtures[MIXED_FLOWS[0]]["nsteps"])
    print("mixed_64x3: 64 flows, %d steps, graph replay" % nsteps, flush=True)
    ctx = gpu.Context(0)
    b = gpu.Batch(ctx, flows).init().use_graph(True)
    b.advance(nsteps)
    for k in MIXED_FLOWS:
        _check(b, k, flows[k], fixtures[k], "mixed_64x3")
    b.close()


SOLVERS = {"tracked": (("dense", 0), ("tracked", 0.1)), "finest_lowrank8": (("lowrank", 8), ("finest", 0.45))}


def _tandem(gpu, ctx, flows, nsteps, nn=None, solver="tracked"):
    (mode, pairs), (build, cost) = SOLVERS[solver]
    b = gpu.Batch(ctx, flows).init().set_precond(mode, pairs).set_precond_build(build, cost)
    if nn is not None:
        b.nn_params(nn)
        b.set_warm_start("learned")
    b.advance(nsteps)
    return b


@pytest.mark.parametrize("warm,solver", [("previous", "tracked"), ("learned", "tracked"),
                                         ("previous", "finest_lowrank8")])
def test_tandem_batch_matches_oracle_fixture(gpu, warm, solver):
    g = _fixture("tandem_16x1024x6_c0")
    flows = W.tandem_16x1024x6()
    nsteps = int(g["nsteps"])
    print("tandem_16x1024x6: 16 cases, %d steps, %s warm start, %s" % (nsteps, warm, solver), flush=True)
    ctx = gpu.Context(0)
    nn = None
    if warm == "learned":   # network trained on a separate run's first steps
        t = gpu.Batch(ctx, flows).init().set_precond("dense", 0).set_precond_build("tracked", 0.1)
        t.collect(True, 1 << 20)
        t.advance(10)
        t.train(iters=300, lr=3e-3)
        nn = t.nn_params()
        t.close()
    b = _tandem(gpu, ctx, flows, nsteps, nn, solver)
    st = b.solver_stats()
    print(st, flush=True)
    assert st["local_helmholtz_m_applications"] > 0   # the bench's local A^-1 path ran
    assert st["fallbacks"] == 0 and st["abandoned"] == 0
    if solver == "tracked":
        assert st["refreshes"] > 0 and st["build_m_applications"] == 0
    _check(b, 0, flows[0], g, "tandem_16x1024x6 (%s start, %s)" % (warm, solver))
    b.close()
