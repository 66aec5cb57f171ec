"""Batches on separate contexts (one HIP stream each) advanced at the same time
from separate host threads (ibpm_amd.advance_concurrently): each batch's
vorticity, flux, forces and iteration counts equal, bit for bit, those of the
same batch advanced alone, so sharing the GPU changes nothing but the time."""
import numpy as np
import pytest

from tests.helpers import plate_geom

pytestmark = pytest.mark.gpu

NX, NGRID, LENGTH, RE, DT, STEPS = 128, 3, 4.0, 300.0, 0.005, 6
GEOMS = [plate_geom(25, 1.0, 0.3, 0.2, 0.4), plate_geom(31, 0.8, 0.2, 0.3, 0.5, center=(0.1, -0.05))]


def _batch(gpu, ctx, geom):
    f = gpu.Flow(NX, NX, NGRID, LENGTH, -2.0, -2.0, RE, DT, geom, tol=1e-10)
    b = gpu.Batch(ctx, [f]).init()
    b.set_precond("dense", 0).set_local_helmholtz(True).set_precond_build("tracked", 0.1).set_warm_start("zero")
    return b


def _state(b):
    st = b.solver_stats()
    return b.omega(0), b.q(0), np.array(b.forces(0)), st["iterations"], st["m_applications"]


def test_concurrent_batches_equal_lone_runs(gpu):
    print("concurrent: two moving plates 128^2 x 3, %d steps" % STEPS, flush=True)
    alone = []
    for geom in GEOMS:
        ctx = gpu.Context(0)
        b = _batch(gpu, ctx, geom)
        b.advance(STEPS)
        alone.append(_state(b))
        b.close()
        ctx.close()
    ctxs = [gpu.Context(0) for _ in GEOMS]
    bs = [_batch(gpu, c, g) for c, g in zip(ctxs, GEOMS)]
    gpu.advance_concurrently(bs, STEPS)
    for k, b in enumerate(bs):
        w, q, fc, its, mapps = _state(b)
        w0, q0, fc0, its0, mapps0 = alone[k]
        print("flow", k, "iterations", its, its0, "forces", fc, fc0, flush=True)
        assert its == its0 and mapps == mapps0
        assert np.array_equal(w, w0) and np.array_equal(q, q0) and np.array_equal(fc, fc0)
    for b in bs:
        b.close()


CYL = "body C\n circle_n 0 0 0.5 40\nend\n"


def test_concurrent_graph_replays_equal_lone_runs(gpu):
    # stationary flows (Cholesky force solve): each batch replays its own captured step graph
    print("concurrent: two stationary cylinders 96^2 x 2 and 64^2 x 3, graph replay, %d steps" % STEPS, flush=True)
    shapes = [(96, 2), (64, 3)]

    def mk(ctx, nx, ngrid):
        f = gpu.Flow(nx, nx, ngrid, LENGTH, -2.0, -2.0, 100.0, 0.01, CYL)
        return gpu.Batch(ctx, [f]).init().use_graph(True)

    alone = []
    for nx, ngrid in shapes:
        ctx = gpu.Context(0)
        b = mk(ctx, nx, ngrid)
        b.advance(STEPS)
        alone.append((b.omega(0), b.q(0), np.array(b.forces(0))))
        b.close()
        ctx.close()
    ctxs = [gpu.Context(0) for _ in shapes]
    bs = [mk(c, nx, ngrid) for c, (nx, ngrid) in zip(ctxs, shapes)]
    gpu.advance_concurrently(bs, STEPS)
    for k, b in enumerate(bs):
        w0, q0, fc0 = alone[k]
        assert np.array_equal(b.omega(0), w0) and np.array_equal(b.q(0), q0)
        assert np.array_equal(np.array(b.forces(0)), fc0)
    for b in bs:
        b.close()


def test_concurrent_batches_need_their_own_contexts(gpu):
    ctx = gpu.Context(0)
    bs = [_batch(gpu, ctx, g) for g in GEOMS]
    with pytest.raises(ValueError):
        gpu.advance_concurrently(bs, 1)
    for b in bs:
        b.close()
