"""Test-side restatement of the kernel-table preconditioner matrix, built on
the oracle's operators (test infrastructure; the library forms the same
matrix on the GPU in k_pc_assemble).

M = beta E G E^T (ProjectionSolver.cc:96-101, NavierStokesModel.cc:101-140),
with G the level-0 flux-to-flux operator Curl L^-1 (-A^-1 Curl).  On an
unbounded lattice G is translation invariant, so G(a, b) is read from the
response to one unit flux of b's component placed at a source node s,
evaluated at a's node shifted by s minus b's node, and taken as zero past
the level; the library places sources on a lattice over the points
(`sources`) and each column reads the nearest.
"""
import numpy as np


def sources(points, nx, ny, dx, xoff, yoff):
    """The library's source nodes of a table (table_sources in engine.cpp):
    a lattice nx/8 by ny/8 nodes apart (at least 8; doubled until at most 64
    nodes) centred on the level-0 extent of the points at set-up."""
    x, y = points
    i0, i1 = int(np.floor((x.min() - xoff) / dx)), int(np.ceil((x.max() - xoff) / dx))
    j0, j1 = int(np.floor((y.min() - yoff) / dx)), int(np.ceil((y.max() - yoff) / dx))
    di, dj = max(8, nx // 8), max(8, ny // 8)
    while True:
        ni, nj = max(0, i1 - i0) // di + 1, max(0, j1 - j0) // dj + 1
        if ni * nj <= 64:
            break
        di, dj = 2 * di, 2 * dj
    out = []
    for a in range(ni):
        for b in range(nj):
            i = (i0 + i1) // 2 - ((ni - 1) * di) // 2 + a * di
            j = (j0 + j1) // 2 - ((nj - 1) * dj) // 2 + b * dj
            out.append((min(max(i, 1), nx - 2), min(max(j, 1), ny - 2)))
    return out


def tables(oracle, nx, ny, ngrid, dx, beta, reynolds, node=None):
    """Level-0 responses of G (A^-1 with beta's alpha, no beta factor) to a
    unit X flux and a unit Y flux at `node` (default the centre), each as
    (X part (nx+1, ny), Y part (nx, ny+1))."""
    nf1 = 2 * nx * ny + nx + ny
    nfx = (nx + 1) * ny
    ci, cj = node if node is not None else (nx // 2, ny // 2)
    units = (ci * ny + cj, nfx + ci * (ny + 1) + cj)
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


def assemble(tab, nx, ny, assoc, n, scale, centre=None, srcs=None):
    """scale * E G E^T from the association lists (boundary index, level-0
    flux index, weight) of the points.  One table: `tab` read about `centre`
    (default the centre node; another origin is a planted mistake).  Several:
    `tab` a list of tables and `srcs` their nodes, each column reading the
    table whose node is nearest to the middle node of its 3 x 3 entries
    (first of equals)."""
    bi, fi, w = assoc
    nfx = (nx + 1) * ny
    comp = (fi >= nfx).astype(int)
    ii = np.where(comp == 0, fi // ny, (fi - nfx) // (ny + 1))
    jj = np.where(comp == 0, fi % ny, (fi - nfx) % (ny + 1))
    if srcs is None:
        tabs, srcs = [tab], [centre if centre is not None else (nx // 2, ny // 2)]
    else:
        tabs = tab
    S = np.array(srcs)
    near = np.zeros(n, dtype=int)   # per column: the source nearest to its first node + (1, 1)
    for col in range(n):
        m = bi == col
        if m.any():
            d = (ii[m].min() + 1 - S[:, 0]) ** 2 + (jj[m].min() + 1 - S[:, 1]) ** 2
            near[col] = int(np.argmin(d))
    M = np.zeros((n, n))
    for cb in (0, 1):
        for ca in (0, 1):
            A = np.where(comp == ca)[0]
            B = np.where(comp == cb)[0]
            for s_ in set(near[bi[B]].tolist()):
                Bs = B[near[bi[B]] == s_]
                T = tabs[s_][cb][ca]
                di = ii[A][:, None] - ii[Bs][None, :] + S[s_, 0]
                dj = jj[A][:, None] - jj[Bs][None, :] + S[s_, 1]
                ok = (di >= 0) & (di < T.shape[0]) & (dj >= 0) & (dj < T.shape[1])
                g = np.where(ok, T[np.clip(di, 0, T.shape[0] - 1), np.clip(dj, 0, T.shape[1] - 1)], 0.0)
                g *= w[A][:, None] * w[Bs][None, :]
                np.add.at(M, (np.repeat(bi[A][:, None], len(Bs), 1), np.repeat(bi[Bs][None, :], len(A), 0)), g)
    return scale * M


def upper_mirrored(M):
    """The matrix the library forms: each entry above the diagonal computed
    once and mirrored."""
    U = np.triu(M)
    return U + np.triu(M, 1).T
