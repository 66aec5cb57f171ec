"""Test-side restatement of the kernel-table preconditioner matrix, built on
the oracle's operators (test infrastructure; the library forms the same
matrix on the GPU in k_pc_assemble).

M = beta E G E^T (ProjectionSolver.cc:96-101, NavierStokesModel.cc:101-140),
with G the level-0 flux-to-flux operator Curl L^-1 (-A^-1 Curl).  On an
unbounded lattice G is translation invariant, so G(a, b) is read from the
response to one unit flux of b's component placed at node (nx/2, ny/2),
evaluated at a's node shifted by (nx/2, ny/2) minus b's node, and taken as
zero past the level.
"""
import numpy as np


def tables(oracle, nx, ny, ngrid, dx, beta, reynolds):
    """Level-0 responses of G (A^-1 with beta's alpha, no beta factor) to a
    unit X flux and a unit Y flux at the centre, each as (X part (nx+1, ny),
    Y part (nx, ny+1))."""
    nf1 = 2 * nx * ny + nx + ny
    nfx = (nx + 1) * ny
    units = ((nx // 2) * ny + ny // 2, nfx + (nx // 2) * (ny + 1) + ny // 2)
    out = []
    for idx in units:
        q = np.zeros(nf1 * ngrid)
        q[idx] = 1.0
        w = oracle.curl_flux(nx, ny, ngrid, dx, q)
        w = oracle.elliptic_solve(1, nx, ny, ngrid, dx, -beta / 2.0 / reynolds, w)
        psi = oracle.elliptic_solve(0, nx, ny, ngrid, dx, 0.0, -w)
        r = oracle.curl_scalar(nx, ny, ngrid, psi)[:nf1]
        out.append((r[:nfx].reshape(nx + 1, ny), r[nfx:].reshape(nx, ny + 1)))
    return out


def assemble(tab, nx, ny, assoc, n, scale, centre=None):
    """scale * E G E^T from the association lists (boundary index, level-0
    flux index, weight) of the points; `centre` moves the table's origin (a
    planted mistake when it differs from the unit flux's node)."""
    bi, fi, w = assoc
    cx, cy = centre if centre is not None else (nx // 2, ny // 2)
    nfx = (nx + 1) * ny
    comp = (fi >= nfx).astype(int)
    ii = np.where(comp == 0, fi // ny, (fi - nfx) // (ny + 1))
    jj = np.where(comp == 0, fi % ny, (fi - nfx) % (ny + 1))
    M = np.zeros((n, n))
    for cb in (0, 1):
        for ca in (0, 1):
            A = np.where(comp == ca)[0]
            B = np.where(comp == cb)[0]
            T = tab[cb][ca]
            di = ii[A][:, None] - ii[B][None, :] + cx
            dj = jj[A][:, None] - jj[B][None, :] + cy
            ok = (di >= 0) & (di < T.shape[0]) & (dj >= 0) & (dj < T.shape[1])
            g = np.where(ok, T[np.clip(di, 0, T.shape[0] - 1), np.clip(dj, 0, T.shape[1] - 1)], 0.0)
            g *= w[A][:, None] * w[B][None, :]
            np.add.at(M, (np.repeat(bi[A][:, None], len(B), 1), np.repeat(bi[B][None, :], len(A), 0)), g)
    return scale * M


def upper_mirrored(M):
    """The matrix the library forms: each entry above the diagonal computed
    once and mirrored."""
    U = np.triu(M)
    return U + np.triu(M, 1).T
