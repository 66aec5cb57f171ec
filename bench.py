"""Benchmark of the MI355X projection-step path: grid-point.steps/s of the
configured workload on one GPU, with the dominant kernel's roofline and the
CPU oracle timed beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg2]

A step is one full RK3 timestep (3 substeps) of every flow in the workload;
grid-pt counts the stored interior nodes of all levels.  Inputs are resident
in HBM when the timed region starts; the timed region is K hipGraph replays
(stationary bodies) or K eager steps with CG (moving bodies).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
DEFAULT_STEPS = {"cfg1": 500, "cfg2": 200, "cfg3": 20, "cfg4": 20, "cfg5": 5}
CONFIG_TEXT = {
    "cfg1": "stationary cylinder D=1 (157 pts), 200x200, 1 level, Re 100, dt 0.02, Cholesky",
    "cfg2": "stationary cylinder D=1 (314 pts), 400x400 finest, 4 levels, Re 200, dt 0.01, Cholesky",
    "cfg3": "64 stationary flows (cylinder/ellipse/plate), n in {256,384,512}, 3 levels, dt 0.005",
    "cfg4": "pitching/plunging plate (100 pts), 512x512, 5 levels, Re 300, dt 0.0025, CG tol 1e-10",
    "cfg5": "16 x 3 tandem pitching/plunging plates, 1024x1024, 6 levels, Re 500, dt 0.001, CG tol 1e-10",
}


def cpu_baseline(flows, budget_s):
    """Times the FP64 C oracle on the host (1 core) on the same workload's
    first flow: a bounded sample of the step loop after a set-up that skips
    the one-time Cholesky build."""
    from oracle import oracle as O
    f = flows[0]
    s = O.Sim(f.nx, f.ny, f.ngrid, f.length, f.xoffset, f.yoffset, f.Re, f.dt, f.geom,
              scheme=f.scheme, tol=f.tol, force_cg=f.force_cg)
    s.init_timing()
    t0 = time.perf_counter()
    s.advance(1)
    t1 = time.perf_counter()
    per = t1 - t0
    n = max(1, min(200, int(budget_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    s.advance(n)
    t1 = time.perf_counter()
    from ibpm_amd.workloads import grid_points
    g = grid_points(f)
    return {"value": g * n / (t1 - t0), "unit": "grid-pt*steps/s", "cores": 1, "kind": "port",
            "sample": "%d steps of the first flow (%dx%d, %d levels) after 1 warm step; "
                      "Cholesky build excluded (identity factors, same triangular-solve loops)"
                      % (n, f.nx, f.ny, f.ngrid)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cfg2", choices=sorted(CONFIG_TEXT))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--warm", default="previous", choices=["previous", "zero", "learned"],
                    help="CG warm start of moving-body workloads (previous = the reference's)")
    ap.add_argument("--train-steps", type=int, default=30,
                    help="steps of sample collection before training the learned warm start")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import ibpm_amd
    from ibpm_amd.workloads import WORKLOADS, grid_points
    steps = args.steps or DEFAULT_STEPS[args.workload]
    flows = WORKLOADS[args.workload]()
    ctx = ibpm_amd.Context(0)
    b = ibpm_amd.Batch(ctx, flows)
    t_init0 = time.perf_counter()
    b.init()
    t_init = time.perf_counter() - t_init0
    moving = any("pitchplunge" in f.geom for f in flows)
    if not args.no_graph and not moving:
        b.use_graph(True)
    learn = None
    if moving and args.warm == "learned":
        # sample collection with the reference warm start, then training on the GPU
        b.collect(True, 4 << 20)
        its = []
        for _ in range(args.train_steps):
            b.advance(1)
            its.append(sum(int(b.cg_iters(k).sum()) for k in range(len(flows))) / len(flows))
        t_tr = time.perf_counter()
        loss0 = b.train(iters=1, lr=0.0)
        loss = b.train(iters=800, lr=3e-3)
        learn = {"train_steps": args.train_steps, "samples": b.num_samples(), "loss_before": loss0,
                 "loss_after": loss, "train_s": round(time.perf_counter() - t_tr, 3),
                 "cg_iters_per_step_previous_warm_start": sum(its[-10:]) / len(its[-10:])}
        b.collect(False).set_warm_start("learned")
    elif moving and args.warm == "zero":
        b.set_warm_start("zero")
    b.advance(args.warmup)
    ctx_sync = ibpm_amd.load().ibpm_ctx_synchronize
    ctx_sync(ctx.h)
    iters = []
    t0 = time.perf_counter()
    if moving:
        for _ in range(steps):
            b.advance(1)
            iters.append(sum(int(b.cg_iters(k).sum()) for k in range(len(flows))) / len(flows))
    else:
        b.advance(steps)
    ctx_sync(ctx.h)
    t1 = time.perf_counter()
    sec = t1 - t0
    gp = sum(grid_points(f) for f in flows)
    value = gp * steps / sec
    cg = (sum(iters) / len(iters)) if moving else None
    # live per-kernel timing (HIP events on the library stream, eager steps)
    stats = b.kernel_stats(2 if moving else 3)
    dom = max(stats.items(), key=lambda kv: kv[1][1])
    name, (cnt, ms, by) = dom
    avg_s = ms / cnt / 1e3
    achieved = by / cnt / avg_s / 1e9
    cd, cl = b.forces(0)
    line = {
        "metric": "grid-pt*steps/s (projection step, %s)" % ("moving-body CG" if moving else "Cholesky"),
        "value": value,
        "unit": "grid-pt*steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": sec / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (configured geometry and parameters, zero initial vorticity)",
        "config": {"workload": args.workload, "description": CONFIG_TEXT[args.workload],
                   "flows": len(flows), "grid_points": gp, "parallelism": "single GPU",
                   "warm_start": args.warm if moving else None,
                   "cg_iters_per_step": cg, "learned": learn, "init_s": round(t_init, 3),
                   "Cd_flow0": cd, "Cl_flow0": cl},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "avg_launch_us": avg_s * 1e6, "bytes_per_launch": by / cnt},
        "kernels": {k: {"launches": c, "ms": round(m, 4), "GB/s": (bb / (m / 1e3) / 1e9) if (bb and m) else None}
                    for k, (c, m, bb) in sorted(stats.items(), key=lambda kv: -kv[1][1])},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(flows, args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
