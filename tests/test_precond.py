"""CPU checks of the preconditioned force solve's algebra, restated in numpy in
the order the PCG kernels apply it, on the oracle's M of a pitching plate:

  * the limited-memory update H = (I - S G Y^T) P (I - Y G S^T) + S G S^T
    (G = (S^T Y)^-1) maps every kept M-image back to its direction, H Y = S
    (exactly for symmetric M, to the asymmetry of a multi-level M), and is
    symmetric positive definite -- S and Y swapped in the right factor (a
    planted mistake) breaks H Y = S, swapped in the left one breaks symmetry;
  * with P the exact inverse, PCG converges in a handful of iterations;
  * PCG and plain CG stop on the same unpreconditioned test r.r <= tol^2 and
    give the same force to that tolerance;
  * an inverse of M built at an earlier body position, scaled by beta to
    another substep, still cuts the iteration count several-fold.
"""
import numpy as np
import pytest

from tests.helpers import plate_geom

TOL = 1e-10


def pcg(M, b, x0, apply_h, keep=0, maxit=5000):
    """Preconditioned CG as the kernels run it: begin (r = b - M x0), then per
    iteration alpha, x, r, the residual test on r.r, the pair (d, M d) kept for
    the first `keep` iterations, z = H r, beta = (r.z)_new / (r.z)_old."""
    x = x0.copy()
    r = b - M @ x
    if r @ r <= TOL * TOL:
        return x, 0, []
    z = apply_h(r)
    d = z.copy()
    rho = r @ z
    it, pairs = 0, []
    while it < maxit:
        q = M @ d
        alpha = rho / (d @ q)
        if it < keep:
            pairs.append((d.copy(), q.copy()))
        x += alpha * d
        r -= alpha * q
        it += 1
        if r @ r <= TOL * TOL:
            break
        z = apply_h(r)
        rnew = r @ z
        d = z + (rnew / rho) * d
        rho = rnew
    return x, it, pairs


def lmp(P, pairs, mistake=None):
    """The limited-memory update with the symmetrised Gram matrix.  mistake
    "left" or "right" swaps S and Y in that factor (planted errors)."""
    if not pairs:
        return lambda r: P @ r
    S = np.stack([p[0] for p in pairs], 1)
    Y = np.stack([p[1] for p in pairs], 1)
    W = S.T @ Y
    G = np.linalg.inv(0.5 * (W + W.T))
    A, B = (Y, S) if mistake == "left" else (S, Y)     # left factor  I - A G B^T
    C, D = (S, Y) if mistake == "right" else (Y, S)    # right factor I - C G D^T

    def h(r):
        u = P @ (r - C @ (G @ (D.T @ r)))
        return u - A @ (G @ (B.T @ u)) + S @ (G @ (S.T @ r))
    return h


@pytest.fixture(scope="module")
def plate_mats(oracle):
    """Dense M of substeps 0 and 2 of a moving plate at step 3, and M of
    substep 0 at step 15 (the body has moved)."""
    s = oracle.Sim(64, 64, 2, 4.0, -2.0, -2.0, 300.0, 0.01, plate_geom(17, 1.0, 0.3, 0.25, 0.4), tol=TOL)
    s.init()
    s.advance(3)
    n = 2 * s.nb
    E = np.eye(n)
    M0 = np.stack([s.M(0, e) for e in E], 1)
    M2 = np.stack([s.M(2, e) for e in E], 1)
    betas = [s.beta(i) for i in range(3)]
    s.advance(12)
    M0late = np.stack([s.M(0, e) for e in E], 1)
    return M0, M2, M0late, betas


def test_update_maps_images_back_and_is_spd(plate_mats):
    M0, _, _, _ = plate_mats
    n = len(M0)
    rng = np.random.default_rng(0)
    b = M0 @ rng.standard_normal(n)
    _, _, pairs = pcg(M0, b, np.zeros(n), lambda r: r, keep=6)
    P = np.eye(n) / np.mean(np.diag(M0))   # on the scale of M^-1
    H = np.stack([lmp(P, pairs)(e) for e in np.eye(n)], 1)
    S = np.stack([p[0] for p in pairs], 1)
    Y = np.stack([p[1] for p in pairs], 1)
    # exact for symmetric M; multi-level M is non-symmetric at ~1e-6 here
    asym = np.abs(M0 - M0.T).max() / np.abs(M0).max()
    assert np.abs(H @ Y - S).max() <= max(1e-12, 10 * asym) * np.abs(S).max()
    assert np.abs(H - H.T).max() <= 1e-10 * np.abs(H).max()
    assert np.linalg.eigvalsh(0.5 * (H + H.T)).min() > 0
    Hr = np.stack([lmp(P, pairs, mistake="right")(e) for e in np.eye(n)], 1)
    assert np.abs(Hr @ Y - S).max() > 1e-3 * np.abs(S).max()
    Hl = np.stack([lmp(P, pairs, mistake="left")(e) for e in np.eye(n)], 1)
    assert np.abs(Hl - Hl.T).max() > 1e-3 * np.abs(Hl).max()


def test_exact_inverse_converges_at_once_and_matches_plain_cg(plate_mats):
    M0, _, _, _ = plate_mats
    n = len(M0)
    b = M0 @ np.random.default_rng(1).standard_normal(n)
    x_cg, it_cg, _ = pcg(M0, b, np.zeros(n), lambda r: r)
    Pinv = np.linalg.inv(0.5 * (M0 + M0.T))
    x_p, it_p, _ = pcg(M0, b, np.zeros(n), lambda r: Pinv @ r)
    assert it_p <= 4 < it_cg
    assert np.linalg.norm(b - M0 @ x_p) <= 2 * TOL
    assert np.abs(x_p - x_cg).max() <= 1e-6 * np.abs(x_cg).max()


def test_stale_beta_scaled_inverse_cuts_iterations(plate_mats):
    M0, M2, M0late, betas = plate_mats
    n = len(M0)
    rng = np.random.default_rng(2)
    P0 = np.linalg.inv(0.5 * (M0 + M0.T))
    for M, scale in ((M2, betas[0] / betas[2]), (M0late, 1.0)):
        b = M @ rng.standard_normal(n)
        _, it_cg, _ = pcg(M, b, np.zeros(n), lambda r: r)
        _, it_p, _ = pcg(M, b, np.zeros(n), lambda r: scale * (P0 @ r))
        assert it_p * 3 <= it_cg, (it_p, it_cg)


def test_lowrank_pairs_from_previous_solve_help_a_stale_inverse(plate_mats):
    M0, _, M0late, _ = plate_mats
    n = len(M0)
    rng = np.random.default_rng(3)
    P0 = np.linalg.inv(0.5 * (M0 + M0.T))
    b1 = M0late @ rng.standard_normal(n)
    _, it_stale, pairs = pcg(M0late, b1, np.zeros(n), lambda r: P0 @ r, keep=8)
    b2 = M0late @ rng.standard_normal(n)
    _, it2_stale, _ = pcg(M0late, b2, np.zeros(n), lambda r: P0 @ r)
    _, it2_lmp, _ = pcg(M0late, b2, np.zeros(n), lmp(P0, pairs))
    assert it2_lmp <= it2_stale


# ---------------------------------------------------------------------------
# The kernel-table matrix: M from G's response to one unit flux per component.

@pytest.fixture(scope="module")
def table_case(oracle):
    """A pitching plate of 25 points on 128^2 x 3: the oracle's dense M of
    substep 2 after 4 steps, the association lists of its points and the
    kernel tables of that substep."""
    from tests import kernel_table as KT
    nx, ngrid, length = 128, 3, 4.0
    dx = length / nx
    s = oracle.Sim(nx, nx, ngrid, length, -2.0, -2.0, 300.0, 0.005, plate_geom(25, 1.0, 0.3, 0.2, 0.4), tol=TOL)
    s.init()
    s.advance(4)
    n = 2 * s.nb
    M = np.stack([s.M(2, e) for e in np.eye(n)], 1)
    pts = s.points()
    assoc = oracle.reg_build(nx, nx, dx, -2.0, -2.0, pts[:s.nb], pts[s.nb:])
    tab = KT.tables(oracle, nx, nx, ngrid, dx, s.beta(2), 300.0)
    return dict(s=s, M=M, assoc=assoc, tab=tab, nx=nx, n=n, beta=s.beta(2))


def test_kernel_table_matrix_approximates_m(table_case):
    """E G E^T from the translated unit response agrees with the exact
    multi-level M to a few parts in 10^6 (3e-6 here), and its inverse
    clusters the spectrum of P M within 1e-3 of 1; a table read about an
    origin one node off, or with the X and Y responses swapped, misses by
    order one (planted mistakes)."""
    from tests import kernel_table as KT
    c = table_case
    nx, n = c["nx"], c["n"]
    MK = KT.upper_mirrored(KT.assemble(c["tab"], nx, nx, c["assoc"], n, c["beta"]))
    rel = np.linalg.norm(MK - c["M"]) / np.linalg.norm(c["M"])
    assert rel <= 1e-4, rel
    ev = np.linalg.eigvals(np.linalg.solve(MK, c["M"])).real
    assert 0.999 <= ev.min() and ev.max() <= 1.001, (ev.min(), ev.max())
    off = KT.assemble(c["tab"], nx, nx, c["assoc"], n, c["beta"], centre=(nx // 2 + 1, nx // 2))
    assert np.linalg.norm(off - c["M"]) / np.linalg.norm(c["M"]) > 0.1
    swapped = KT.assemble(c["tab"][::-1], nx, nx, c["assoc"], n, c["beta"])
    assert np.linalg.norm(swapped - c["M"]) / np.linalg.norm(c["M"]) > 0.1


def test_newton_schulz_step_tracks_the_moving_table_matrix(table_case, oracle):
    """One step P <- P (2 I - M~ P) from the inverse at the previous step's
    points squares the error I - M~ P at the new points (I - M~ P1 = (I - M~
    P)^2: 0.072 to 0.0022 here), and PCG with the refreshed P needs 3
    iterations where plain CG needs 26."""
    from tests import kernel_table as KT
    c = table_case
    s, nx, n = c["s"], c["nx"], c["n"]
    dx = 4.0 / nx
    M_old = KT.upper_mirrored(KT.assemble(c["tab"], nx, nx, c["assoc"], n, c["beta"]))
    P = np.linalg.inv(M_old)
    s.advance(1)
    pts = s.points()
    assoc = oracle.reg_build(nx, nx, dx, -2.0, -2.0, pts[:s.nb], pts[s.nb:])
    M_new = KT.upper_mirrored(KT.assemble(c["tab"], nx, nx, assoc, n, c["beta"]))
    e0 = np.linalg.norm(np.eye(n) - M_new @ P, 2)
    P1 = P @ (2 * np.eye(n) - M_new @ P)
    e1 = np.linalg.norm(np.eye(n) - M_new @ P1, 2)
    assert 1e-2 < e0 < 0.5 and e1 <= e0 * e0 * (1 + 1e-6), (e0, e1)
    M = np.stack([s.M(2, e) for e in np.eye(n)], 1)
    b = M @ np.random.default_rng(4).standard_normal(n)
    _, it_cg, _ = pcg(M, b, np.zeros(n), lambda r: r)
    _, it_p, _ = pcg(M, b, np.zeros(n), lambda r: P1 @ r)
    assert it_p <= 4 and it_p * 5 <= it_cg, (it_p, it_cg)
