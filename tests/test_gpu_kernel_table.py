"""The kernel-table preconditioner on the GPU: the matrix k_pc_assemble forms
from the library's own unit-flux responses, against the test-side
restatement on the oracle's operators (tests/kernel_table.py) and against
the oracle's exact M; and the P that the Newton-Schulz refresh (k_dgemm)
keeps, against the inverse of that matrix."""
import numpy as np
import pytest

from tests import kernel_table as KT
from tests.helpers import plate_geom

pytestmark = pytest.mark.gpu

NX, NGRID, LENGTH, RE, DT = 128, 3, 4.0, 300.0, 0.005
GEOM = plate_geom(25, 1.0, 0.3, 0.2, 0.4)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("steps", [1, 4])
def test_gpu_table_matrix_and_refreshed_p(gpu, oracle, steps):
    dx = LENGTH / NX
    print("kernel table: pitching plate 128^2 x 3, %d steps" % steps, flush=True)
    ctx = gpu.Context(0)
    f = gpu.Flow(NX, NX, NGRID, LENGTH, -2.0, -2.0, RE, DT, GEOM, tol=1e-10)
    b = gpu.Batch(ctx, [f]).init().set_precond("dense", 0).set_precond_build("tracked", 0.1)
    b.advance(steps)
    mt, p = b.precond_matrices(0, 2)
    s = oracle.Sim(NX, NX, NGRID, LENGTH, -2.0, -2.0, RE, DT, GEOM, tol=1e-10)
    s.init()
    p0 = s.points()   # the points at set-up place the table's sources
    srcs = KT.sources((p0[:s.nb], p0[s.nb:]), NX, NX, dx, -2.0, -2.0)
    s.advance(steps)
    pts = b.points(0)
    assert np.array_equal(pts, s.points())   # the same substep's points
    n = 2 * s.nb
    assoc = oracle.reg_build(NX, NX, dx, -2.0, -2.0, pts[:s.nb], pts[s.nb:])
    tabs = [KT.tables(oracle, NX, NX, NGRID, dx, s.beta(2), RE, node) for node in srcs]
    ref = KT.upper_mirrored(KT.assemble(tabs, NX, NX, assoc, n, s.beta(2), srcs=srcs))
    print("sources", srcs, flush=True)
    M = np.stack([s.M(2, e) for e in np.eye(n)], 1)
    e = np.linalg.norm(np.eye(n) - mt @ p, 2)
    st = b.solver_stats()
    print("table vs restatement %.2e, vs exact M %.2e, |I - M~ P| %.2e" % (rel(mt, ref), rel(mt, M), e), st,
          flush=True)
    assert np.array_equal(mt, mt.T)
    assert rel(mt, ref) <= 1e-12
    assert len(srcs) > 1
    assert rel(mt, M) <= 2e-5
    assert e <= 1e-2
    assert st["build_m_applications"] == 0 and st["builds"] == 1
    assert st["refreshes"] == 3 * steps - 1   # every solve but the one after the build
    assert np.all(b.cg_iters() <= 4)
    b.close()
