"""Benchmark of the MI355X projection-step path: grid-point.steps/s of the
configured workload on one GPU, with the dominant kernel's roofline and the
CPU oracle timed beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload tandem_16x1024x6]

A step is one full RK3 timestep (3 substeps) of every flow in the workload;
grid-pt counts the stored interior nodes of all levels.  Inputs are resident
in HBM when the timed region starts; the timed region is K hipGraph replays
(stationary bodies) or K eager steps with the CG force solve (moving bodies).
For moving bodies the batched, preconditioned form is timed, then the plain
GPU form (one flow, plain CG as the reference) on the first flow for the same
steps, with the speed-up and the Cd/Cl difference between the two.  The
default workload is the batch of 16 tandem pitching/plunging plate cases
(1024x1024, 6 levels).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
DEFAULT_STEPS = {"cylinder_200": 500, "cylinder_400x4": 200, "mixed_64x3": 20, "plate_512x5": 40, "tandem_16x1024x6": 40}
METRIC = "grid-pt\u00b7steps/s learned+batched vs plain GPU at Cd/Cl err\u22641e-6; CG iters/step"
CONFIG_TEXT = {
    "cylinder_200": "stationary cylinder D=1 (157 pts), 200x200, 1 level, Re 100, dt 0.02, Cholesky",
    "cylinder_400x4": "stationary cylinder D=1 (314 pts), 400x400 finest, 4 levels, Re 200, dt 0.01, Cholesky",
    "mixed_64x3": "64 stationary flows (cylinder/ellipse/plate), n in {256,384,512}, 3 levels, dt 0.005",
    "plate_512x5": "pitching/plunging plate (100 pts), 512x512, 5 levels, Re 300, dt 0.0025, CG tol 1e-10",
    "tandem_16x1024x6": "16 x 3 tandem pitching/plunging plates, 1024x1024, 6 levels, Re 500, dt 0.001, CG tol 1e-10",
}


def _static_geom(geom):
    return "".join(l + "\n" for l in geom.splitlines() if not l.strip().startswith("motion"))


def cpu_baseline(flows, budget_s, cg_iters_per_step=None):
    """Times the FP64 C oracle on the host (1 core) on the workload's first
    flow.  Stationary bodies: a bounded sample of the step loop after a set-up
    that skips the one-time Cholesky build.  Moving bodies (CG): one oracle
    step of the same grid with the body held still (everything but the force
    solve) plus the oracle's M application timed on the moving body, times the
    plain-CG iterations per step that the GPU plain form measured (the same
    algorithm and counts as the oracle)."""
    from oracle import oracle as O
    from ibpm_amd.workloads import grid_points
    import numpy as np
    f = flows[0]
    g = grid_points(f)
    moving = "pitchplunge" in f.geom
    s = O.Sim(f.nx, f.ny, f.ngrid, f.length, f.xoffset, f.yoffset, f.Re, f.dt,
              _static_geom(f.geom) if moving else f.geom, scheme=f.scheme, tol=f.tol,
              force_cg=f.force_cg and not moving)
    s.init_timing()
    n = 1
    if not moving:   # one warm step, then as many as fit the budget
        t0 = time.perf_counter()
        s.advance(1)
        per = time.perf_counter() - t0
        n = max(1, min(200, int(budget_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    s.advance(n)
    t_step = (time.perf_counter() - t0) / n
    if not moving:
        return {"value": g / t_step, "unit": "grid-pt*steps/s", "cores": 1, "kind": "port",
                "sample": "%d steps of the first flow (%dx%d, %d levels) after 1 warm step; "
                          "Cholesky build excluded (identity factors, same triangular-solve loops)"
                          % (n, f.nx, f.ny, f.ngrid)}
    m = O.Sim(f.nx, f.ny, f.ngrid, f.length, f.xoffset, f.yoffset, f.Re, f.dt, f.geom,
              scheme=f.scheme, tol=f.tol)
    m.init()
    x = np.random.default_rng(0).standard_normal(2 * m.nb)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        m.M(0, x)
    t_m = (time.perf_counter() - t0) / reps
    its = cg_iters_per_step or 0.0
    t = t_step + its * t_m
    return {"value": g / t, "unit": "grid-pt*steps/s", "cores": 1, "kind": "port",
            "sample": "first flow (%dx%d, %d levels): 1 oracle step with the body held still "
                      "(%.2f s) + %.1f plain-CG iterations/step x oracle M application (%.3f s, mean of %d)"
                      % (f.nx, f.ny, f.ngrid, t_step, its, t_m, reps)}


def traffic_per_launch(workload, kernel, alg_bytes):
    """HBM bytes per launch of the dominant kernel: the ratio of measured
    traffic (rocprofv3 FETCH_SIZE and WRITE_SIZE, calibrated on a plain copy)
    to algorithmic bytes, from the committed profile of this workload and
    kernel, times this run's algorithmic bytes per launch; None without one."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    t = json.load(open(path))
    if t.get("workload") != workload or t.get("kernel") != kernel:
        return None
    return t["traffic_over_algorithmic"] * alg_bytes


def timed_steps(b, nflows, steps, moving):
    """Advances `steps` steps; returns (seconds, CG iterations per step per flow)."""
    import ibpm_amd
    sync = ibpm_amd.load().ibpm_ctx_synchronize
    sync(b.ctx.h)
    iters = []
    t0 = time.perf_counter()
    if moving:
        for _ in range(steps):
            b.advance(1)
            iters.append(sum(int(b.cg_iters(k).sum()) for k in range(nflows)) / nflows)
    else:
        b.advance(steps)
    sync(b.ctx.h)
    return time.perf_counter() - t0, (sum(iters) / len(iters) if iters else None)


def plain_form(ctx, flow, total_steps, steps):
    """The plain GPU form on one flow: plain CG from the previous force with
    every A^-1 of M by the DST solve (the reference's solver), no batching;
    advances total_steps and times the last `steps`.  Returns throughput,
    iterations per step and final Cd, Cl."""
    import ibpm_amd
    from ibpm_amd.workloads import grid_points
    b = ibpm_amd.Batch(ctx, [flow]).init().set_local_helmholtz(False)
    b.advance(total_steps - steps)
    sec, its = timed_steps(b, 1, steps, True)
    cd, cl = b.forces(0)
    b.close()
    return {"value": grid_points(flow) * steps / sec, "ms_per_step": sec / steps * 1e3,
            "cg_iters_per_step": its, "Cd": cd, "Cl": cl}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="tandem_16x1024x6", choices=sorted(CONFIG_TEXT),
                    help="tandem_16x1024x6 (default): the moving-plate batch the metric is quoted on")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--precond", default="lowrank", choices=["none", "stale", "lowrank"],
                    help="force solve of moving bodies: plain CG (the reference's), CG preconditioned "
                         "by a stale FP64 inverse of M, or that plus the low-rank update")
    ap.add_argument("--pairs", type=int, default=8)
    ap.add_argument("--warm", default="previous", choices=["previous", "zero", "learned"],
                    help="CG start of moving-body workloads (previous = the reference's)")
    ap.add_argument("--train-steps", type=int, default=30,
                    help="steps of sample collection before training the learned warm start")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain GPU form (moving bodies) run beside the batched one")
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
    done_steps = 0
    if moving:
        b.set_precond(args.precond, args.pairs)
    if moving and args.warm == "learned":
        # sample collection, then training on the GPU
        b.collect(True, 4 << 20)
        its = []
        for _ in range(args.train_steps):
            b.advance(1)
            its.append(sum(int(b.cg_iters(k).sum()) for k in range(len(flows))) / len(flows))
        done_steps += args.train_steps
        t_tr = time.perf_counter()
        loss0 = b.train(iters=1, lr=0.0)
        loss = b.train(iters=800, lr=3e-3)
        learn = {"train_steps": args.train_steps, "samples": b.num_samples(), "loss_before": loss0,
                 "loss_after": loss, "train_s": round(time.perf_counter() - t_tr, 3),
                 "cg_iters_per_step_previous_warm_start": sum(its[-10:]) / len(its[-10:])}
        b.collect(False).set_warm_start("learned")
    elif moving and args.warm == "zero":
        b.set_warm_start("zero")
    t_w0 = time.perf_counter()
    b.advance(args.warmup)
    t_warm = time.perf_counter() - t_w0
    done_steps += args.warmup
    st0 = b.solver_stats() if moving else None
    sec, cg = timed_steps(b, len(flows), steps, moving)
    done_steps += steps
    gp = sum(grid_points(f) for f in flows)
    value = gp * steps / sec
    cd, cl = b.forces(0)
    solver = None
    if moving:
        st1 = b.solver_stats()
        d = {k: st1[k] - st0[k] for k in st1}
        solver = {"precond": args.precond, "pairs": args.pairs if args.precond == "lowrank" else 0,
                  "warm_start": args.warm, "cg_iters_per_step": cg,
                  "batched_m_applications_per_step": (d["m_applications"] + d["build_m_applications"]) / steps,
                  "builds_in_timed_steps": d["builds"], "fallbacks": d["fallbacks"],
                  "host_ms_per_step_kinematics": round(d["host_us_kinematics"] / steps / 1e3, 3),
                  "host_ms_per_step_polls": round(d["host_us_polls"] / steps / 1e3, 3),
                  "m_applications_with_local_helmholtz": d["local_helmholtz_m_applications"] / max(
                      1, d["m_applications"] + d["build_m_applications"]),
                  "warmup_s_including_first_build": round(t_warm, 3)}
    # live per-kernel timing (HIP events on the library stream, eager steps)
    stats = b.kernel_stats(2 if moving else 3)
    b.close()
    # dominant kernel: most time among the classes with algorithmic bytes
    dom = max(stats.items(), key=lambda kv: kv[1][1] if kv[1][2] else 0.0)
    name, (cnt, ms, by) = dom
    avg_s = ms / cnt / 1e3
    achieved = by / cnt / avg_s / 1e9
    plain = None
    if moving and not args.no_plain:
        plain = plain_form(ctx, flows[0], done_steps, steps)
        plain["speedup_of_this_form"] = (value / plain["value"])
        plain["Cd_rel_err_vs_this_form_flow0"] = abs(cd - plain["Cd"]) / abs(plain["Cd"])
        plain["Cl_rel_err_vs_this_form_flow0"] = abs(cl - plain["Cl"]) / max(abs(plain["Cl"]), 1e-300)
    line = {
        "metric": METRIC if moving else "grid-pt\u00b7steps/s (stationary bodies, Cholesky force solve)",
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
                   "solver": solver, "learned": learn, "plain_gpu_form": plain,
                   "init_s": round(t_init, 3), "Cd_flow0": cd, "Cl_flow0": cl},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_per_launch(args.workload, name, by / cnt),
                     "avg_launch_us": avg_s * 1e6, "bytes_per_launch": by / cnt},
        "kernels": {k: {"launches": c, "ms": round(m, 4), "GB/s": (bb / (m / 1e3) / 1e9) if (bb and m) else None}
                    for k, (c, m, bb) in sorted(stats.items(), key=lambda kv: -kv[1][1])},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(flows, args.cpu_seconds,
                                            plain["cg_iters_per_step"] if plain else None)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
