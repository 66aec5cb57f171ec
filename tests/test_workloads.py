"""CPU checks of the configured workloads: sizes, point spacing (points much
closer than a cell make M = C A^-1 B singular) and determinism."""
import math

import numpy as np
import pytest

from ibpm_amd import workloads as W


def _points(geom, t=0.0):
    from oracle import oracle as O
    xy, _ = O.geom_points(geom, t)
    n = len(xy) // 2
    return xy[:n], xy[n:]


@pytest.mark.parametrize("name", ["cylinder_200", "cylinder_400x4", "mixed_64x3", "plate_512x5", "tandem_16x1024x6"])
def test_workload_point_spacing(oracle, name):
    flows = W.WORKLOADS[name]()
    for f in flows[:16]:
        dx = f.length / f.nx
        x, y = _points(f.geom)
        pts = np.stack([x, y], 1)
        d = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
        np.fill_diagonal(d, np.inf)
        assert d.min() > 0.75 * dx, (name, d.min() / dx)


def test_workload_sizes_and_determinism():
    assert [f.nx for f in W.cylinder_200()] == [200] and W.cylinder_200()[0].ngrid == 1
    assert W.cylinder_400x4()[0].nx == 400 and W.cylinder_400x4()[0].ngrid == 4
    c3 = W.mixed_64x3()
    assert len(c3) == 64 and {f.nx for f in c3} <= {256, 384, 512}
    assert all(f.ngrid == 3 and f.dt == 0.005 and 50 <= f.Re <= 300 for f in c3)
    assert [f.geom for f in W.mixed_64x3()] == [f.geom for f in c3]
    c5 = W.tandem_16x1024x6()
    assert len(c5) == 16 and all(f.geom.count("body") == 3 and f.nx == 1024 and f.ngrid == 6 for f in c5)
    assert W.grid_points(W.cylinder_400x4()[0]) == 4 * 399 * 399
