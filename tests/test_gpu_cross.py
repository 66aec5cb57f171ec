"""The cross products on the GPU, through the C ABI (ibpm_cross_product,
ibpm_cross_flux_flux): the reference's own cross-product identities
(VectorOperationsTest.cc:593-693) evaluated with the HIP kernels alone, so the
sign of the nonlinear term is pinned on the GPU without the oracle, then
parity with the oracle on random fields (1e-14 relative)."""
import numpy as np
import pytest

from tests.helpers import G

pytestmark = pytest.mark.gpu

VG = G(8, 12, 3, 2.0, -1.0, -3.0)
NS = (VG.nx - 1) * (VG.ny - 1) + 1


@pytest.fixture(scope="module")
def ctx(gpu):
    return gpu.Context(0)


@pytest.fixture(scope="module")
def fields(ctx):
    g = VG
    sc = [np.zeros(g.ngrid * g.ns)]
    for m in range(1, NS):
        sc.append(g.single_wavenumber((m - 1) % (g.nx - 1) + 1, (m - 1) // (g.nx - 1) + 1))
    fl = [ctx.curl_scalar(g.nx, g.ny, g.ngrid, g.dx, s) for s in sc]
    return sc, fl


def test_gpu_cross_identities(ctx, oracle, fields):
    print("gpu cross identities 8x12x3", flush=True)
    sc, fl = fields
    g = VG
    for q in fl[::7]:   # q x q = 0
        assert np.all(ctx.cross_flux_flux(g.nx, g.ny, g.ngrid, g.dx, q, q) == 0.0)
    for m in range(0, NS, 11):   # antisymmetry
        for n in range(0, NS, 5):
            a = ctx.cross_flux_flux(g.nx, g.ny, g.ngrid, g.dx, fl[m], fl[n])
            b = ctx.cross_flux_flux(g.nx, g.ny, g.ngrid, g.dx, fl[n], fl[m])
            assert np.array_equal(a, -b)
    # (u, 0) x (0, v) = u v for constants, boundary of the coarsest level aside
    u = np.full(g.ngrid * g.ns, 3.0)
    v = np.full(g.ngrid * g.ns, 5.0)
    q1 = oracle.xvel_to_flux(g.nx, g.ny, g.ngrid, g.dx, u)
    q2 = oracle.yvel_to_flux(g.nx, g.ny, g.ngrid, g.dx, v)
    c = ctx.cross_flux_flux(g.nx, g.ny, g.ngrid, g.dx, q1, q2)
    top = g.ngrid - 1
    for lev in range(g.ngrid):
        for i in range(1, g.nx):
            for j in range(1, g.ny):
                if lev == top and (i in (1, g.nx - 1) or j in (1, g.ny - 1)):
                    continue
                assert abs(c[g.si(lev, i, j)] - 15.0) < 1e-13
    # triple product <a, q1 x q2> = <q1, q2 x a>, both products from the GPU
    worst = big = 0.0
    for m in range(0, NS, 4):
        for n in range(0, NS, 3):
            p1, p2 = fl[n], fl[NS - n - 1]
            ips = oracle.ip_scalar(g.nx, g.ny, g.ngrid, g.dx, sc[m],
                                   ctx.cross_flux_flux(g.nx, g.ny, g.ngrid, g.dx, p1, p2))
            ipf = oracle.ip_flux(g.nx, g.ny, g.ngrid, p1, ctx.cross_product(g.nx, g.ny, g.ngrid, g.dx, p2, sc[m]))
            worst = max(worst, abs(ips - ipf))
            big = max(big, abs(ips))
    assert worst < 5e-14 and big > 1e-2


@pytest.mark.parametrize("nx,ny,ngrid", [(8, 12, 3), (32, 32, 4), (200, 200, 1), (64, 96, 2)])
def test_gpu_cross_products_match_oracle(ctx, oracle, nx, ny, ngrid):
    print("cross parity", nx, ny, ngrid, flush=True)
    g = G(nx, ny, ngrid, 2.0, -1.0, -3.0)
    rng = np.random.default_rng(7 * nx + ngrid)
    w = rng.standard_normal(ngrid * g.ns)
    q1 = rng.standard_normal(ngrid * g.nf)
    q2 = rng.standard_normal(ngrid * g.nf)
    a = ctx.cross_product(nx, ny, ngrid, g.dx, q1, w)
    b = oracle.cross_product(nx, ny, ngrid, g.dx, q1, w)
    assert np.linalg.norm(a - b) <= 1e-14 * np.linalg.norm(b)
    a = ctx.cross_flux_flux(nx, ny, ngrid, g.dx, q1, q2)
    b = oracle.cross_flux_flux(nx, ny, ngrid, g.dx, q1, q2)
    assert np.linalg.norm(a - b) <= 1e-14 * np.linalg.norm(b)
