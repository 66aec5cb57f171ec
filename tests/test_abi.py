"""CPU checks of the C ABI: the library loads and exports every function that
include/ibpm_gpu.h declares; host-only entry points (geometry parsing) agree
with the oracle.  No GPU call is made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ibpm_gpu.h")
LIB = os.path.join(ROOT, "ibpm_amd", "libibpm_gpu.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ibpm_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libibpm_gpu.so is not built (python -m ibpm_amd.build)")
    return ctypes.CDLL(LIB)


def test_every_declared_symbol_is_exported(lib):
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("geom", [
    "body C\n circle_n 0 0 0.5 157\nend\n",
    "body C\n circle 0 0 0.5 0.02\nend\n",
    "body P\n line 0 0 1 0 0.05\n motion pitchplunge 0.1 0.2 0 0.1 0.2 0\nend\n"
    "body Q\n line_aoa 1 0.25 0 10 21\n point 3 3\nend\n",
])
def test_geometry_point_counts_match_oracle(lib, oracle, geom):
    lib.ibpm_geom_num_points.argtypes = [ctypes.c_char_p]
    n = lib.ibpm_geom_num_points(geom.encode())
    xy, _ = oracle.geom_points(geom, 0.0)
    assert n == len(xy) // 2


def test_bad_geometry_is_an_error_not_an_abort(lib):
    lib.ibpm_geom_num_points.argtypes = [ctypes.c_char_p]
    lib.ibpm_last_error.restype = ctypes.c_char_p
    assert lib.ibpm_geom_num_points(b"body P\n motion pitchplunge 0 0 0.5 0.2\nend\n") == -1
    assert b"pitchplunge" in lib.ibpm_last_error()


def test_driver_is_built():
    assert os.access(os.path.join(ROOT, "ibpm_amd", "ibpm"), os.X_OK)


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def test_restart_file_round_trips_with_reference_state(lib, tmp_path):
    """ibpm_state_save -> reference State::load and reference State::save ->
    ibpm_state_load, bit-exact (src/State.cc:65-223)."""
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libibpm_refpart.so")
    if not os.path.exists(ref_path):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    R = ctypes.CDLL(ref_path)
    import numpy as np
    nx, ny, ng, L, x0, y0, npts = 16, 12, 2, 2.0, -1.0, -0.5, 7
    rng = np.random.default_rng(4)
    q = rng.standard_normal(ng * (2 * nx * ny + nx + ny))
    w = rng.standard_normal(ng * (nx - 1) * (ny - 1))
    f = rng.standard_normal(2 * npts)
    D = ctypes.c_double
    a = str(tmp_path / "ours.bin").encode()
    lib.ibpm_state_save.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, D, D, D,
                                    ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, D]
    assert lib.ibpm_state_save(a, nx, ny, ng, L / nx, x0, y0, npts, q.ctypes.data, w.ctypes.data,
                               f.ctypes.data, 17, D(0.34)) == 0
    q2, w2, f2 = np.zeros_like(q), np.zeros_like(w), np.zeros_like(f)
    ts, tt = ctypes.c_int(), ctypes.c_double()
    assert R.ref_state_load(a, nx, ny, ng, D(L), D(x0), D(y0), npts, _p(q2), _p(w2), _p(f2),
                            ctypes.byref(ts), ctypes.byref(tt)) == 0
    assert np.array_equal(q, q2) and np.array_equal(w, w2) and np.array_equal(f, f2)
    assert ts.value == 17 and tt.value == 0.34
    b = str(tmp_path / "ref.bin").encode()
    assert R.ref_state_save(b, nx, ny, ng, D(L), D(x0), D(y0), npts, _p(q), _p(w), _p(f), 23, D(1.5)) == 0
    assert open(a, "rb").read()[:-12] == open(b, "rb").read()[:-12]   # same bytes up to step/time
    lib.ibpm_state_load.argtypes = [ctypes.c_char_p] + [ctypes.c_void_p] * 12
    hdr = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), D(), D(), D(), ctypes.c_int()]
    q3, w3, f3 = np.zeros_like(q), np.zeros_like(w), np.zeros_like(f)
    assert lib.ibpm_state_load(b, *[ctypes.byref(h) for h in hdr], q3.ctypes.data, w3.ctypes.data,
                               f3.ctypes.data, ctypes.byref(ts), ctypes.byref(tt)) == 0
    assert [h.value for h in hdr] == [nx, ny, ng, L / nx, x0, y0, npts]
    assert np.array_equal(q, q3) and np.array_equal(w, w3) and np.array_equal(f, f3)
    assert ts.value == 23 and tt.value == 1.5
