"""The reference's solver-class tests through the C++ classes of
ibpm_amd/host/ibpm_gpu.hpp (EllipticSolver2d, PoissonSolver, Regularizer,
NavierStokesModel, ProjectionSolver / CholeskySolver /
ConjugateGradientSolver), compiled by the build into tests/cpp/boundary_test:
EllipticSolver2dTest.cc:43-172, EllipticSolverTest.cc:86-112 and
ProjectionSolverTest.cc:70-165, at the reference's 1e-10 tolerance."""
import os
import subprocess
import tempfile

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "boundary_test")


def test_reference_solver_class_tests(gpu):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/boundary_test is not built (python -m ibpm_amd.build)")
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all ok" in r.stdout
    assert r.stdout.count(" ok ") >= 20
