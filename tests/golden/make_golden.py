"""Writes the golden fixtures of the GPU parity tests from the FP64 CPU oracle
(oracle/ibpm_oracle.c).  Run in this repository (takes minutes to ~1.5 h):

    python tests/golden/make_golden.py cylinder_200        # 250 steps, ~1 min
    python tests/golden/make_golden.py cylinder_400x4        # 2000 steps, ~1.5 h
    python tests/golden/make_golden.py plate       # moving plate, 30 steps
    python tests/golden/make_golden.py plate_512x5_short   # configured moving plate, 12 steps
    python tests/golden/make_golden.py mixed_64x3_f0 ...   # one flow of the 64-flow batch, 40 steps
    python tests/golden/make_golden.py tandem_16x1024x6_c0 # first tandem case, 30 steps, ~1 h

Each fixture holds the per-step Cd/Cl series (full precision), CG iterations
per substep where CG is used, per-level vorticity norms and the level-0
vorticity (every 2nd node of grids larger than 200^2, every 4th node for the
fixtures of single flows drawn from the batched workloads).

The mixed_64x3 fixtures cover one flow of each grid size and body type of the
64-flow batch (flow indices MIXED_FLOWS); the GPU test runs them inside the
whole batch.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from ibpm_amd import workloads as W  # noqa: E402

PLATE = ("body Plate\n  line_n -0.5 0 0.5 0 33\n  center 0 0\n"
         "  motion pitchplunge 0.2617993877991494 0.4 0 0.2 0.4 0\nend\n")

# first flow of each (grid size, body) pair of mixed_64x3:
# 256: cylinder 2, ellipse 5, plate 4; 384: 20, 3, 11; 512: 0, 1, 51
MIXED_FLOWS = (2, 5, 4, 20, 3, 11, 0, 1, 51)

CASES = {
    "cylinder_200": (W.cylinder_200()[0], 250),
    "cylinder_400x4": (W.cylinder_400x4()[0], 2000),
    "plate": (W.Flow(128, 128, 3, 4.0, -2.0, -2.0, 300.0, 0.005, PLATE, tol=1e-10), 30),
    "plate_512x5_short": (W.plate_512x5()[0], 12),
    "tandem_16x1024x6_c0": (W.tandem_16x1024x6()[0], 30),
}
for _k in MIXED_FLOWS:
    CASES["mixed_64x3_f%d" % _k] = (W.mixed_64x3()[_k], 40)
STRIDE4 = ("mixed_64x3_", "tandem_")


def run(name):
    f, nsteps = CASES[name]
    t0 = time.time()
    s = O.Sim(f.nx, f.ny, f.ngrid, f.length, f.xoffset, f.yoffset, f.Re, f.dt, f.geom,
              scheme=f.scheme, tol=f.tol, force_cg=f.force_cg)
    s.init()
    print(name, "init %.1fs" % (time.time() - t0), flush=True)
    forces, iters = [], []
    for k in range(nsteps):
        s.advance(1)
        forces.append(s.forces())
        if s.uses_cg:
            iters.append(s.cg_iters())
        if (k + 1) % 50 == 0 or k + 1 == nsteps:
            print(name, "step", k + 1, "%.1fs" % (time.time() - t0), forces[-1], flush=True)
    w = s.omega()
    ns = (f.nx - 1) * (f.ny - 1)
    norms = [float(np.linalg.norm(w[l * ns:(l + 1) * ns])) for l in range(f.ngrid)]
    w0 = w[:ns].reshape(f.nx - 1, f.ny - 1)
    stride = 1 if f.nx <= 200 else 2
    if name.startswith(STRIDE4):
        stride = 4
    out = os.path.join(HERE, name + ".npz")
    np.savez_compressed(out, forces=np.array(forces), cg_iters=np.array(iters, dtype=np.int32),
                        omega_norms=np.array(norms), omega0=w0[::stride, ::stride], stride=stride,
                        nsteps=nsteps, f_final=s.f())
    print("wrote", out, "%.1fs" % (time.time() - t0), flush=True)


if __name__ == "__main__":
    for n in sys.argv[1:]:
        run(n)
