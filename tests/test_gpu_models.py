"""The linearized, adjoint, linear periodic and SFD models (reference
IBSolver.cc:203-279, ibpm.cc:185-235) on the GPU against the oracle's, after
a few steps from the same state: forces every step and the final vorticity
to 1e-10 relative with the Cholesky force solve, 1e-7 with CG (tolerance
1e-10).  One batch holds every
model at once, so each flow's kernels take its own model's branch."""
import numpy as np
import pytest

from ibpm_amd import Flow

pytestmark = pytest.mark.gpu
FIXED = "body P\n line_n -0.5 0 0.5 0 21\n center 0 0\n motion fixed 0 0 -0.3\nend\n"
MOVING = ("body P\n line_n -0.5 0 0.5 0 13\n center 0 0\n"   # points one cell apart
          "  motion pitchplunge 0.2 0.4 0 0.1 0.4 0\nend\n")
GRID = (48, 48, 2, 4.0, -2.0, -2.0)
RE, DT, NSTEPS = 100.0, 0.02, 4


@pytest.fixture(scope="module")
def base_states(oracle):
    s = oracle.Sim(*GRID, RE, DT, FIXED)
    s.init()
    s.advance(3)
    a = (s.omega(), s.q())
    s.advance(2)
    return a, (s.omega(), s.q())


CASES = [
    ("linear", FIXED, {}),
    ("adjoint", FIXED, {}),
    ("linearperiodic", FIXED, {"period": 2}),
    ("sfd", FIXED, {"chi": 0.5, "Delta": 2.0}),
    ("sfd", MOVING, {"chi": 0.5, "Delta": 2.0, "tol": 1e-10}),
    ("nonlinear", FIXED, {}),
]


def _oracle_run(oracle, model, geom, kw, bases):
    (wa, qa), (wb, qb) = bases
    s = oracle.Sim(*GRID, RE, DT, geom, model=model, period=kw.get("period", 1), chi=kw.get("chi", 0.02),
                   delta=kw.get("Delta", 15.0), tol=kw.get("tol", 1e-7))
    if model in ("linear", "adjoint", "linearperiodic"):
        s.set_base_state(wa, qa, 0)
        if model == "linearperiodic":
            s.set_base_state(wb, qb, 1)
    s.init()
    if model in ("linear", "adjoint", "linearperiodic"):
        s.set_omega(0.01 * wa)
    forces = []
    for _ in range(NSTEPS):
        s.advance(1)
        forces.append(s.forces())
    out = {"forces": np.array(forces), "omega": s.omega()}
    return out


@pytest.mark.parametrize("moving", [False, True])
def test_models_match_oracle_in_one_batch(gpu, oracle, base_states, moving):
    """Stationary bodies (Cholesky) in one batch of every model; the moving
    plate (CG) in a batch of its own, as a batch has one force solver."""
    (wa, qa), (wb, qb) = base_states
    cases = [c for c in CASES if (c[1] is MOVING) == moving]
    flows = []
    for model, geom, kw in cases:
        flows.append(Flow(*GRID, RE, DT, geom, model=model, period=kw.get("period", 1), chi=kw.get("chi", 0.02),
                          Delta=kw.get("Delta", 15.0), tol=kw.get("tol", 1e-7)))
    ctx = gpu.Context(0)
    b = gpu.Batch(ctx, flows)
    for k, (model, _, _) in enumerate(cases):
        if model in ("linear", "adjoint", "linearperiodic"):
            b.set_base_state(wa, qa, 0, k)
            if model == "linearperiodic":
                b.set_base_state(wb, qb, 1, k)
    b.init()
    for k, (model, _, _) in enumerate(cases):
        if model in ("linear", "adjoint", "linearperiodic"):
            b.set_omega(0.01 * wa, k)
    b.advance(NSTEPS)
    for k, (model, geom, kw) in enumerate(cases):
        print("model %s (%s)" % (model, "moving" if geom is MOVING else "fixed"), flush=True)
        ref = _oracle_run(oracle, model, geom, kw, base_states)
        h = b.force_history(k)
        scale = np.max(np.abs(ref["forces"]))
        ferr = np.max(np.abs(h - ref["forces"])) / scale
        w = b.omega(k)
        werr = np.linalg.norm(w - ref["omega"]) / np.linalg.norm(ref["omega"])
        print("  forces %.2e omega %.2e" % (ferr, werr), flush=True)
        # Cholesky: rounding; CG: its absolute residual tolerance 1e-10 (as the
        # nonlinear solver's CG parity tests)
        tol = 1e-7 if moving else 1e-10
        assert ferr <= tol and werr <= tol
