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
