"""Full-length parity gates on the GPU against fixtures written by the FP64
oracle (tests/golden/make_golden.py):

  * Cd relative error <= 1e-6 at every step,
  * Cl error <= 1e-6 relative to max|Cd| (the configured cylinder is
    mirror-symmetric, so its Cl is roundoff-seeded and has no relative scale),
  * vorticity relative L2 <= 1e-5 after the full run (level 0, and per-level
    norms),
  * for CG cases, iterations per substep within +-1 of the oracle's.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from ibpm_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
PLATE = ("body Plate\n  line_n -0.5 0 0.5 0 33\n  center 0 0\n"
         "  motion pitchplunge 0.2617993877991494 0.4 0 0.2 0.4 0\nend\n")


def _fixture(name):
    path = os.path.join(GOLD, name + ".npz")
    if not os.path.exists(path):
        pytest.skip("fixture %s not generated (tests/golden/make_golden.py %s)" % (path, name))
    return np.load(path)


def _flow(name):
    if name == "plate":
        return W.Flow(128, 128, 3, 4.0, -2.0, -2.0, 300.0, 0.005, PLATE, tol=1e-10)
    if name == "plate_512x5_short":
        return W.plate_512x5()[0]
    return W.WORKLOADS[name]()[0]


@pytest.mark.parametrize("name", ["cylinder_200", "cylinder_400x4", "plate", "plate_512x5_short"])
def test_full_run_matches_golden(gpu, name):
    g = _fixture(name)
    f = _flow(name)
    nsteps = int(g["nsteps"])
    print("golden", name, nsteps, "steps", flush=True)
    ctx = gpu.Context(0)
    b = gpu.Batch(ctx, [f]).init()
    moving = "pitchplunge" in f.geom
    if not moving:
        b.use_graph(True)
    iters = []
    for k in range(nsteps):
        b.advance(1)
        if moving:
            iters.append(b.cg_iters())
    hist = b.force_history()
    ref = g["forces"]
    assert hist.shape == ref.shape
    cdmax = np.abs(ref[:, 0]).max()
    cd_err = np.abs(hist[:, 0] - ref[:, 0]) / np.abs(ref[:, 0])
    cl_err = np.abs(hist[:, 1] - ref[:, 1]) / (cdmax if not moving else np.abs(ref[:, 1]).max())
    print(name, "max Cd rel err %.3e, max Cl err %.3e" % (cd_err.max(), cl_err.max()), flush=True)
    assert cd_err.max() <= 1e-6
    assert cl_err.max() <= 1e-6
    w = b.omega()
    ns = (f.nx - 1) * (f.ny - 1)
    s = int(g["stride"])
    w0 = w[:ns].reshape(f.nx - 1, f.ny - 1)[::s, ::s]
    assert np.linalg.norm(w0 - g["omega0"]) / np.linalg.norm(g["omega0"]) <= 1e-5
    norms = np.array([np.linalg.norm(w[l * ns:(l + 1) * ns]) for l in range(f.ngrid)])
    assert np.all(np.abs(norms - g["omega_norms"]) <= 1e-5 * g["omega_norms"])
    if moving:
        it = np.array(iters)
        assert np.abs(it - g["cg_iters"]).max() <= 1, (it, g["cg_iters"])


def test_ibpm_driver_force_file(gpu, oracle):
    exe = os.path.join(os.path.dirname(HERE), "ibpm_amd", "ibpm")
    geom = "body C\n circle_n 0 0 0.5 40\nend\n"
    with tempfile.TemporaryDirectory() as d:
        gpath = os.path.join(d, "cyl.geom")
        open(gpath, "w").write(geom)
        r = subprocess.run([exe, "-geom", gpath, "-nx", "64", "-ny", "64", "-ngrid", "2", "-Re", "100",
                            "-dt", "0.02", "-nsteps", "6", "-outdir", d, "-name", "cyl"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        rows = [l.split() for l in open(os.path.join(d, "cyl.force")).read().splitlines()]
        assert len(rows) == 6 and rows[0][0] == "1"
        full = np.loadtxt(os.path.join(d, "cyl.force17"))
    s = oracle.Sim(64, 64, 2, 4.0, -2.0, -2.0, 100.0, 0.02, geom)
    s.init()
    for k in range(6):
        s.advance(1)
        cd, cl = s.forces()
        assert abs(full[k, 2] - cd) <= 1e-9 * abs(cd)
        assert abs(full[k, 3] - cl) <= 1e-9 * abs(cd)
        assert abs(float(rows[k][2]) - cd) <= 1e-5 * abs(cd)


@pytest.mark.parametrize("pcbuild,precond", [("all", "lowrank"), ("finest", "lowrank"), ("tracked", "dense")])
def test_ibpm_driver_moving_plate_precond(gpu, oracle, pcbuild, precond):
    """The driver on a pitching plate with -precond, -tol and each
    preconditioner build: forces agree with the oracle's plain CG to the CG
    tolerance's effect."""
    exe = os.path.join(os.path.dirname(HERE), "ibpm_amd", "ibpm")
    with tempfile.TemporaryDirectory() as d:
        gpath = os.path.join(d, "plate.geom")
        open(gpath, "w").write(PLATE)
        r = subprocess.run([exe, "-geom", gpath, "-nx", "128", "-ny", "128", "-ngrid", "3", "-Re", "300",
                            "-dt", "0.005", "-nsteps", "5", "-tol", "1e-10", "-precond", precond,
                            "-pcbuild", pcbuild, "-outdir", d, "-name", "plate"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        full = np.loadtxt(os.path.join(d, "plate.force17"))
    s = oracle.Sim(128, 128, 3, 4.0, -2.0, -2.0, 300.0, 0.005, PLATE, tol=1e-10)
    s.init()
    for k in range(5):
        s.advance(1)
        cd, cl = s.forces()
        scale = max(abs(cd), abs(cl))
        assert abs(full[k, 2] - cd) <= 1e-7 * scale and abs(full[k, 3] - cl) <= 1e-7 * scale


def test_ibpm_driver_unsteady_base_flow(gpu, oracle):
    """The driver with -ubf 1 on a pitching plate and -alpha 5: the force file
    holds drag and lift rotated by the base-flow angle of each step
    (ibpm.cc:347-351), equal to the oracle's body-frame run."""
    exe = os.path.join(os.path.dirname(HERE), "ibpm_amd", "ibpm")
    with tempfile.TemporaryDirectory() as d:
        gpath = os.path.join(d, "plate.geom")
        open(gpath, "w").write(PLATE)
        r = subprocess.run([exe, "-geom", gpath, "-nx", "128", "-ny", "128", "-ngrid", "3", "-Re", "300",
                            "-dt", "0.005", "-nsteps", "5", "-alpha", "5", "-ubf", "1",
                            "-outdir", d, "-name", "plate"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        full = np.loadtxt(os.path.join(d, "plate.force17"))
    s = oracle.Sim(128, 128, 3, 4.0, -2.0, -2.0, 300.0, 0.005, PLATE, alpha_deg=5.0, ubf=True)
    assert s.ubf == 1
    s.init()
    for k in range(5):
        s.advance(1)
        cd, cl = s.forces()
        scale = max(abs(cd), abs(cl))
        assert abs(full[k, 2] - cd) <= 1e-9 * scale and abs(full[k, 3] - cl) <= 1e-9 * scale


REBUILD_COST = {"all": 1.0, "finest": 0.45, "table": 0.1, "tracked": 0.1}


@pytest.mark.parametrize("name", ["plate", "plate_512x5_short"])
@pytest.mark.parametrize("pairs", [0, 8])
@pytest.mark.parametrize("build", ["all", "finest", "table", "tracked"])
def test_preconditioned_cg_matches_golden(gpu, name, pairs, build):
    """The preconditioned force solve (stale FP64 inverse of M, with and
    without the low-rank update; P from M, from its finest-level form, from
    the kernel table, or from the table refreshed every substep) against the
    oracle's plain-CG fixtures, at the same gates: only iteration counts may
    change."""
    g = _fixture(name)
    f = _flow(name)
    nsteps = int(g["nsteps"])
    print("preconditioned", name, "pairs", pairs, "build", build, nsteps, "steps", flush=True)
    ctx = gpu.Context(0)
    b = gpu.Batch(ctx, [f]).init().set_precond("lowrank" if pairs else "stale", pairs)
    b.set_precond_build(build, REBUILD_COST[build])
    iters = []
    for k in range(nsteps):
        b.advance(1)
        iters.append(b.cg_iters())
    hist = b.force_history()
    ref = g["forces"]
    cd_err = np.abs(hist[:, 0] - ref[:, 0]) / np.abs(ref[:, 0])
    cl_err = np.abs(hist[:, 1] - ref[:, 1]) / np.abs(ref[:, 1]).max()
    st = b.solver_stats()
    print(name, "max Cd rel err %.3e, max Cl err %.3e" % (cd_err.max(), cl_err.max()),
          "iterations/step %.2f (plain CG %.2f)" % (np.sum(iters) / nsteps, g["cg_iters"].sum() / nsteps),
          st, flush=True)
    assert cd_err.max() <= 1e-6
    assert cl_err.max() <= 1e-6
    w = b.omega()
    ns = (f.nx - 1) * (f.ny - 1)
    s = int(g["stride"])
    w0 = w[:ns].reshape(f.nx - 1, f.ny - 1)[::s, ::s]
    assert np.linalg.norm(w0 - g["omega0"]) / np.linalg.norm(g["omega0"]) <= 1e-5
    norms = np.array([np.linalg.norm(w[l * ns:(l + 1) * ns]) for l in range(f.ngrid)])
    assert np.all(np.abs(norms - g["omega_norms"]) <= 1e-5 * g["omega_norms"])
    assert st["fallbacks"] == 0 and st["abandoned"] == 0
    assert st["iterations"] == int(np.sum(iters))
    assert np.sum(iters) <= 0.4 * g["cg_iters"].sum()


def _read_state(path, nx, ny, ngrid, npts):
    import ctypes
    from ibpm_amd import load
    lib = load()
    lib.ibpm_state_load.argtypes = [ctypes.c_char_p] + [ctypes.c_void_p] * 12
    hdr = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_double(),
           ctypes.c_double(), ctypes.c_int()]
    q = np.zeros(ngrid * (2 * nx * ny + nx + ny))
    w = np.zeros(ngrid * (nx - 1) * (ny - 1))
    f = np.zeros(2 * npts)
    ts, tm = ctypes.c_int(), ctypes.c_double()
    assert lib.ibpm_state_load(path.encode(), *[ctypes.byref(h) for h in hdr], q.ctypes.data, w.ctypes.data,
                               f.ctypes.data, ctypes.byref(ts), ctypes.byref(tm)) == 0
    return w, q


@pytest.mark.parametrize("model", ["linear", "adjoint", "sfd"])
def test_ibpm_driver_other_models(gpu, oracle, model):
    """The driver's -model linear|adjoint (base flow and initial condition from
    a nonlinear run's restart file) and -model sfd (that file as initial
    condition, no filtered-state file: filtered vorticity zero, as
    loadFilteredState leaves it): forces equal the oracle's model run."""
    exe = os.path.join(os.path.dirname(HERE), "ibpm_amd", "ibpm")
    geom = "body C\n circle_n 0.1 0.05 0.5 30\nend\n"
    args = ["-nx", "48", "-ny", "48", "-ngrid", "2", "-Re", "100", "-dt", "0.02"]
    with tempfile.TemporaryDirectory() as d:
        gpath = os.path.join(d, "cyl.geom")
        open(gpath, "w").write(geom)
        r = subprocess.run([exe, "-geom", gpath] + args + ["-nsteps", "5", "-restart", "5", "-outdir", d,
                                                          "-name", "base"], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        bin5 = os.path.join(d, "base00005.bin")
        extra = ["-chi", "0.5", "-Delta", "2"] if model == "sfd" else ["-baseflow", bin5]
        r = subprocess.run([exe, "-geom", gpath] + args + ["-model", model, "-ic", bin5, "-nsteps", "3",
                                                          "-restart", "0", "-outdir", d, "-name", "m"] + extra,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        full = np.loadtxt(os.path.join(d, "m.force17"))
        w0, q0 = _read_state(bin5, 48, 48, 2, 30)
    s = oracle.Sim(48, 48, 2, 4.0, -2.0, -2.0, 100.0, 0.02, geom, model=model, chi=0.5, delta=2.0)
    if model != "sfd":
        s.set_base_state(w0, q0)
    s.init()
    s.set_omega(w0)
    if model == "sfd":
        s.set_filtered(np.zeros_like(w0))
    for k in range(3):
        s.advance(1)
        cd, cl = s.forces()
        scale = max(abs(cd), abs(cl))
        print(model, k, full[k, 2], cd, full[k, 3], cl, flush=True)
        assert abs(full[k, 2] - cd) <= 1e-9 * scale and abs(full[k, 3] - cl) <= 1e-9 * scale


def test_precond_build_arguments(gpu):
    """ibpm_batch_set_precond_build refuses an unknown source of P, a
    non-positive rebuild cost, and the finest-level build without the local
    Helmholtz path; every refusal leaves an error message."""
    from ibpm_amd import IbpmError, load
    ctx = gpu.Context(0)
    b = gpu.Batch(ctx, [_flow("plate")]).init().set_precond("lowrank", 8)
    lib = load()
    assert lib.ibpm_batch_set_precond_build(b.h, 4, 1.0) != 0
    assert b"preconditioner build" in lib.ibpm_last_error()
    assert lib.ibpm_batch_set_precond_build(b.h, 1, 0.0) != 0
    assert b"rebuild cost" in lib.ibpm_last_error()
    with pytest.raises(IbpmError):
        b.set_precond_build("finest", -1.0)
    b.set_precond_build("finest", 0.45).set_precond_build("all", 1.0)
    b.set_local_helmholtz(False)
    with pytest.raises(IbpmError, match="local Helmholtz"):
        b.set_precond_build("finest", 0.45)
    b.close()
