"""Builds ibpm_amd/libibpm_gpu.so (gfx950 HIP kernels + host engine + C ABI)
in-tree with hipcc.  Run: python -m ibpm_amd.build"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libibpm_gpu.so")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("IBPM_OFFLOAD_ARCH", "gfx950")

DEVICE_SOURCES = ["elliptic.hip", "ops.hip", "learn.hip", "scratch.hip"]
HOST_SOURCES = ["engine.cpp", "geometry.cpp", "restart.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]
# Packed FP32 instructions are kept out of every code object.
DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src, defines=(), obj_dir=OBJ):
    os.makedirs(obj_dir, exist_ok=True)
    obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + DEVICE_FLAGS + list(defines) + ["-c", path, "-o", obj]
    else:   # host code (no kernels): its device pass targets the same architecture
        cmd = [HIPCC] + COMMON + ["--offload-arch=" + ARCH] + list(defines) + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stdout + r.stderr).splitlines()
            if "packed-fp32-ops" not in l and "not a recognized feature" not in l and l.strip()]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s" % (src, "\n".join(msgs)))
    return obj, msgs


def build(verbose=False, defines=(), out=OUT, obj_dir=OBJ):
    """Compiles every source (always) and links libibpm_gpu.so (or `out`,
    with extra -D `defines`, for a tuning variant)."""
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(lambda s: _compile(s, defines, obj_dir), DEVICE_SOURCES + HOST_SOURCES))
    objs = [o for o, _ in results]
    for _, msgs in results:
        for m in msgs:
            if verbose or "warning" in m:
                print(m, file=sys.stderr)
    tmp = out + ".tmp"
    subprocess.check_call([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, out)
    if out == OUT:
        build_driver()
        build_boundary_test()
    return out


DRIVER = os.path.join(HERE, "ibpm")


def build_driver():
    """The `ibpm` command-line driver (C++ host code over the C ABI)."""
    src = os.path.join(HERE, "host", "ibpm_main.cpp")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall", "-o", DRIVER + ".tmp", src,
                           "-L" + HERE, "-libpm_gpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


BOUNDARY_TEST = os.path.join(os.path.dirname(HERE), "tests", "cpp", "boundary_test")


def build_boundary_test():
    """The reference's solver-class tests restated over the C++ classes of
    host/ibpm_gpu.hpp (run by tests/test_gpu_boundary.py)."""
    src = BOUNDARY_TEST + ".cpp"
    if not os.path.exists(src):
        return None
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall", "-o", BOUNDARY_TEST + ".tmp", src,
                           "-L" + HERE, "-libpm_gpu", "-Wl,-rpath,$ORIGIN/../../ibpm_amd"])
    os.replace(BOUNDARY_TEST + ".tmp", BOUNDARY_TEST)
    return BOUNDARY_TEST


if __name__ == "__main__":
    print(build(verbose=True))
