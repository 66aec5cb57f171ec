"""CPU pins of the oracle's other models (reference IBSolver.cc:203-279,
ibpm.cc:185-235): the linearized, adjoint, linear periodic and SFD forms of
the nonlinear term, each checked against an identity that the reference's
definitions imply, so that a sign, argument order or phase slip fails:

  * linearized N is the bilinear part of the nonlinear term:
    N_L(x; x0) = N(x0 + x) - N(x0) - N(x) (N is bilinear in (q, omega));
  * the adjoint N is the adjoint of the linearized operator
    omega -> N_L(q(omega), omega; x0) in the energy inner product <q1, q2>
    (q(omega) the flux of omega) on a single-level grid, where the discrete
    identity holds to rounding;
  * linear periodic with base states (A, B) equals the linear model with its
    base state switched A, B, A between steps; period 1 equals linear;
  * SFD with chi = 0 equals the nonlinear solver bitwise, and its filtered
    state starts at the state and then lags it.
"""
import numpy as np
import pytest

from oracle import oracle as O

GEOM = "body P\n line_n -0.5 0 0.5 0 21\n center 0 0\n motion fixed 0 0 -0.3\nend\n"


def _fields(nx, ny, L, seed):
    rng = np.random.default_rng(seed)
    ns = L * (nx - 1) * (ny - 1)
    nf = L * (2 * nx * ny + nx + ny)
    return rng.standard_normal(nf), rng.standard_normal(ns), rng.standard_normal(nf), rng.standard_normal(ns)


def test_linearized_N_is_the_bilinear_part(oracle):
    nx, ny, L = 32, 24, 3
    dx = 4.0 / nx
    q0, w0, q, w = _fields(nx, ny, L, 1)
    lin = O.linear_N(nx, ny, L, dx, q0, w0, q, w)
    ref = (O.nonlinear(nx, ny, L, dx, q0 + q, w0 + w) - O.nonlinear(nx, ny, L, dx, q0, w0)
           - O.nonlinear(nx, ny, L, dx, q, w))
    assert np.max(np.abs(lin - ref)) <= 1e-13 * np.max(np.abs(lin))
    # a swapped pair (q0 with q) is another operator
    swapped = O.linear_N(nx, ny, L, dx, q, w0, q0, w)
    assert np.max(np.abs(swapped - ref)) > 1e-2 * np.max(np.abs(lin))


def _flux_of(nx, ny, L, dx, w):
    psi = O.elliptic_solve(0, nx, ny, L, dx, 0.0, -w)
    return O.curl_scalar(nx, ny, L, psi)


def test_adjoint_N_is_the_energy_adjoint_of_the_linearized_operator(oracle):
    nx = ny = 8
    L = 1
    dx = 1.0 / nx
    ns = (nx - 1) * (ny - 1)
    rng = np.random.default_rng(2)
    w0 = rng.standard_normal(ns)
    q0 = _flux_of(nx, ny, L, dx, w0) + 0.3
    AL = np.zeros((ns, ns))
    AA = np.zeros((ns, ns))
    Q = []
    for k in range(ns):
        e = np.zeros(ns)
        e[k] = 1.0
        qe = _flux_of(nx, ny, L, dx, e)
        Q.append(qe)
        AL[:, k] = O.linear_N(nx, ny, L, dx, q0, w0, qe, e)
        AA[:, k] = O.adjoint_N(nx, ny, L, dx, q0, w0, qe)
    G = np.array([[O.ip_flux(nx, ny, L, Q[i], Q[j]) for j in range(ns)] for i in range(ns)])
    lhs, rhs = G @ AA, AL.T @ G
    scale = np.max(np.abs(lhs))
    assert np.max(np.abs(lhs - rhs)) <= 1e-12 * scale
    assert np.max(np.abs(lhs + rhs)) > 0.5 * scale   # the opposite sign is not adjoint


def _sim(model="nonlinear", **kw):
    return O.Sim(48, 48, 2, 4.0, -2.0, -2.0, 100.0, 0.02, GEOM, model=model, **kw)


@pytest.fixture(scope="module")
def base_states(oracle):
    s = _sim()
    s.init()
    s.advance(3)
    a = (s.omega(), s.q())
    s.advance(2)
    b = (s.omega(), s.q())
    return a, b


def test_linear_periodic_selects_phase_step_mod_period(base_states):
    (wa, qa), (wb, qb) = base_states
    p = _sim("linearperiodic", period=2)
    p.set_base_state(wa, qa, 0)
    p.set_base_state(wb, qb, 1)
    p.init()
    p.set_omega(0.01 * wa)
    lin = _sim("linear")
    lin.set_base_state(wa, qa)
    lin.init()
    lin.set_omega(0.01 * wa)
    for k, (w, q) in enumerate([(wa, qa), (wb, qb), (wa, qa)]):
        lin.set_base_state(w, q)
        lin.advance(1)
        p.advance(1)
        assert np.array_equal(p.omega(), lin.omega()), k
    one = _sim("linearperiodic", period=1)
    one.set_base_state(wa, qa, 0)
    one.init()
    one.set_omega(0.01 * wa)
    ref = _sim("linear")
    ref.set_base_state(wa, qa)
    ref.init()
    ref.set_omega(0.01 * wa)
    one.advance(2)
    ref.advance(2)
    assert np.array_equal(one.omega(), ref.omega())


def test_sfd_with_zero_gain_is_the_nonlinear_solver(oracle):
    a = _sim("sfd", chi=0.0, delta=3.0)
    b = _sim()
    a.init()
    b.init()
    a.advance(4)
    b.advance(4)
    assert np.array_equal(a.omega(), b.omega())
    assert a.forces() == b.forces()


def test_sfd_damps_the_growth_of_the_state(oracle):
    """The SFD term -chi (omega - omega_hat) pulls the flow toward its
    low-pass filtered state: with a large gain the vorticity changes less per
    step than the plain nonlinear solver's (a sign slip would amplify it)."""
    a = _sim("sfd", chi=5.0, delta=1.0)
    b = _sim()
    a.init()
    b.init()
    a.advance(2)
    b.advance(2)
    wa0, wb0 = a.omega(), b.omega()
    a.advance(3)
    b.advance(3)
    da = np.linalg.norm(a.omega() - wa0)
    db = np.linalg.norm(b.omega() - wb0)
    assert da < 0.9 * db, (da, db)
