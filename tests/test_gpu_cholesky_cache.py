"""The Cholesky factor cache in the reference's .cholesky format
(CholeskySolver::load/save, src/CholeskySolver.cc:113-197; file per substep
<basename>_0i.cholesky, src/IBSolver.cc:77-108)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
CYL40 = "body C\n circle_n 0 0 0.5 40\nend\n"


def parse(path):
    tok = open(path).read().split()
    n = int(tok[0])
    ab = float(tok[1])
    d = np.array([float(t) for t in tok[2:2 + n]])
    assert tok[2 + n] == "#"
    low = np.array([float(t) for t in tok[3 + n:]])
    assert low.size == n * (n - 1) // 2
    L = np.zeros((n, n))
    L[np.tril_indices(n, -1)] = low
    return n, ab, d, L


def test_cholesky_cache_matches_oracle_factor_and_round_trips(gpu, oracle, tmp_path):
    base = str(tmp_path / "cyl")
    ctx = gpu.Context(0)
    f = gpu.Flow(64, 64, 2, 4.0, -2.0, -2.0, 100.0, 0.02, CYL40)
    a = gpu.Batch(ctx, [f]).cholesky_cache(base).init()
    assert a.cholesky_loaded() == 0
    s = oracle.Sim(64, 64, 2, 4.0, -2.0, -2.0, 100.0, 0.02, CYL40)
    s.init()
    for i in range(3):
        n, ab, d, L = parse("%s_%02d.cholesky" % (base, i + 1))
        assert n == 80
        assert ab == (1.0 / 100.0) * s.beta(i)          # CholeskySolver.cc:38, written at 17 digits
        do, Lo = s.cholesky(i)
        assert np.abs(d - do).max() <= 1e-10 * np.abs(do).max()
        Lo = np.tril(Lo, -1)
        assert np.abs(L - Lo).max() <= 1e-10 * np.abs(Lo).max()
    b = gpu.Batch(ctx, [f]).cholesky_cache(base).init()
    assert b.cholesky_loaded() == 3
    for _ in range(4):
        a.advance(1)
        b.advance(1)
    assert np.array_equal(a.omega(), b.omega()) and np.array_equal(a.f(), b.f())
    # another dt changes alpha*beta: the files do not match and are rebuilt
    g = gpu.Flow(64, 64, 2, 4.0, -2.0, -2.0, 100.0, 0.01, CYL40)
    c = gpu.Batch(ctx, [g]).cholesky_cache(base).init()
    assert c.cholesky_loaded() == 0
    assert parse(base + "_01.cholesky")[1] == (1.0 / 100.0) * (8.0 / 15.0 * 0.01)
