"""The premise of the force solve's local Helmholtz path, on the CPU: for a
right-hand side that lives near the body (beta B f: level 0 only, carried to
the coarser levels by coarsify), the multi-level solve A^-1 (the oracle's DST
solve, EllipticSolver.cc:40-70 with HelmholtzSolver) equals, level by level,
Jacobi sweeps over the support grown by S cells, S chosen so that
rho^S <= 1e-18 for the Jacobi contraction rho = 4|c| / |1 - 4c|, c = alpha/dx^2,
with the right-hand side zero off the support and no boundary terms.  This is
the rule the library's planner applies (it falls back to the DST solve when a
box would reach a boundary or a ring); the GPU test compares the library's
two paths.  Half the sweeps on level 0 leave errors far above rounding, so
the comparison has teeth."""
import math

import numpy as np
import pytest

NGRID = 3
LENGTH = 4.0


def _support_chain(s0, s1, t0, t1, ngrid, nx, ny):
    """Support box per level: level 0 given; a coarse point sees fine rows
    2 (ic - ex) - 1 .. + 1, clamped to the central region."""
    ex, ey = nx // 4, ny // 4
    out = [(s0, s1, t0, t1)]
    for _ in range(1, ngrid):
        s0 = max(s0 // 2 + ex, ex + 1)
        s1 = min((s1 + 1) // 2 + ex, ex + nx // 2 - 1)
        t0 = max(t0 // 2 + ey, ey + 1)
        t1 = min((t1 + 1) // 2 + ey, ey + ny // 2 - 1)
        out.append((s0, s1, t0, t1))
    return out


def _sweeps(c):
    rho = abs(4 * c) / abs(1 - 4 * c)
    return math.ceil(math.log(1e-18) / math.log(rho))


def _local_solve(rhs_lev, c, sup, S, nx, ny):
    """S Jacobi sweeps of (1 + alpha Lap) u = r from u = 0 on the support grown
    by S cells; u = 0 off that box.  Arrays are (nx - 1) x (ny - 1), index
    [i - 1, j - 1]."""
    s0, s1, t0, t1 = sup
    b0, b1, c0, c1 = s0 - S, s1 + S, t0 - S, t1 + S
    assert b0 >= 2 and b1 <= nx - 2 and c0 >= 2 and c1 <= ny - 2
    r = np.zeros_like(rhs_lev)
    r[s0 - 1:s1, t0 - 1:t1] = rhs_lev[s0 - 1:s1, t0 - 1:t1]
    box = np.zeros_like(rhs_lev, dtype=bool)
    box[b0 - 1:b1, c0 - 1:c1] = True
    u = np.zeros_like(rhs_lev)
    for _ in range(S):
        s4 = np.zeros_like(u)
        s4[1:, :] += u[:-1, :]
        s4[:-1, :] += u[1:, :]
        s4[:, 1:] += u[:, :-1]
        s4[:, :-1] += u[:, 1:]
        u = np.where(box, (r - c * s4) / (1 - 4 * c), 0.0)
    return u


@pytest.mark.parametrize("cfac,n", [(-0.0164, 64), (-0.068, 128)])
def test_local_sweeps_equal_multilevel_dst_solve(oracle, cfac, n):
    NX = NY = n
    dx = LENGTH / NX
    alpha = cfac * dx * dx   # c on level 0 as in the moving-plate workloads
    ns = (NX - 1) * (NY - 1)
    rng = np.random.default_rng(7)
    f = np.zeros(NGRID * ns)
    s0, s1, t0, t1 = n // 2 - 3, n // 2 + 3, n // 2 - 5, n // 2 + 5
    f0 = np.zeros((NX - 1, NY - 1))
    f0[s0 - 1:s1, t0 - 1:t1] = rng.standard_normal((s1 - s0 + 1, t1 - t0 + 1))
    f[:ns] = f0.ravel()
    u_dst = oracle.elliptic_solve(1, NX, NY, NGRID, dx, alpha, f).reshape(NGRID, NX - 1, NY - 1)
    rhs = oracle.coarsify(NX, NY, NGRID, f).reshape(NGRID, NX - 1, NY - 1)
    sups = _support_chain(s0, s1, t0, t1, NGRID, NX, NY)
    ex = NX // 4
    for lev in range(NGRID):
        c = alpha / (dx * (1 << lev)) ** 2
        S = _sweeps(c)
        sup = sups[lev]
        if lev >= 1:   # the box stays inside the ring of the finer level
            assert sup[0] - S > ex and sup[1] + S < ex + NX // 2
        # the right-hand side is zero off the support
        off = rhs[lev].copy()
        off[sup[0] - 1:sup[1], sup[2] - 1:sup[3]] = 0.0
        assert np.abs(off).max() == 0.0
        u = _local_solve(rhs[lev], c, sup, S, NX, NY)
        ref = u_dst[lev]
        scale = np.abs(ref).max()
        assert np.abs(u - ref).max() <= 1e-14 * scale, (lev, np.abs(u - ref).max() / scale)
        if lev == 0:   # half the sweeps: far from rounding
            short = _local_solve(rhs[lev], c, sup, S // 2, NX, NY)
            assert np.abs(short - ref).max() > 1e-11 * scale
