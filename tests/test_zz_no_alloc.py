"""The allocation-free multi-domain solve (ibpm_elliptic_solve_noalloc) against
ibpm_elliptic_solve, through the C ABI: the same kernels on the same input,
so the two results are bitwise equal.  Two calls in a row, the second on a
smaller grid, check that the fixed scratch array keeps no state that changes
an answer; a grid beyond the array is refused with a message."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(seed, kind, nx, ny, ngrid, dx, alpha):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(ngrid * (nx - 1) * (ny - 1))
    return (kind, nx, ny, ngrid, dx, alpha, f)


def test_noalloc_solve_equals_allocating_solve_twice(gpu):
    ctx = gpu.Context(0)
    for case in (_case(1, 1, 512, 512, 5, 4.0 / 512, -0.0125),   # Helmholtz, register-held passes
                 _case(2, 0, 128, 96, 3, 4.0 / 128, 0.0)):       # Poisson, smaller, LDS passes
        print("noalloc case kind=%d %dx%d x%d" % case[:4], flush=True)
        want = ctx.elliptic_solve(*case)
        got = ctx.elliptic_solve_noalloc(*case)
        assert np.array_equal(got, want), np.max(np.abs(got - want))
        assert np.all(np.isfinite(got)) and np.linalg.norm(got) > 0


def test_noalloc_solve_refuses_grids_beyond_its_scratch(gpu):
    ctx = gpu.Context(0)
    kind, nx, ny, ngrid, dx, alpha, f = _case(3, 0, 1024, 1024, 6, 8.0 / 1024, 0.0)
    with pytest.raises(gpu.IbpmError, match="scratch"):
        ctx.elliptic_solve_noalloc(kind, nx, ny, ngrid, dx, alpha, f)
