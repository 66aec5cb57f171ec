"""Benchmark of the MI355X projection-step path: grid-point.steps/s of the
configured workload on one GPU, with the dominant kernel's roofline and the
CPU oracle timed beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload tandem_16x1024x6]

A step is one full RK3 timestep (3 substeps) of every flow in the workload;
grid-pt counts the stored interior nodes of all levels.  Inputs are resident
in HBM when the timed region starts; the timed region is K hipGraph replays
(stationary bodies) or K eager steps with the CG force solve (moving bodies).

Moving bodies (the default workload, 16 tandem pitching/plunging plate cases
at 1024x1024 with 6 levels): the headline form is the batched solver with CG
preconditioned by the dense P from the kernel table, refreshed before every
solve, from a zero CG start (any other start costs one M application for its
residual), from a zero initial vorticity: W warm-up steps (the first builds
P), then K timed steps with every rebuild that falls due inside them.
Beside it the same steps are timed with the reference's warm start (the
previous force) and with the learned one (the network is first trained on a
separate run of the workload's first steps), for batched plain CG, and for
the plain GPU form (one flow, plain CG, every A^-1 by the DST solve: the
reference's solver).
Flow 0's Cd/Cl at every step is compared with the FP64 oracle's fixture
(tests/golden), and the run exits with status 3 if the error exceeds 1e-6.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
DEFAULT_STEPS = {"cylinder_200": 500, "cylinder_400x4": 200, "mixed_64x3": 20, "plate_512x5": 150,
                 "tandem_16x1024x6": 150}
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
    ratio = t.get("traffic_over_algorithmic", {}).get(kernel)
    if t.get("workload") != workload or ratio is None:
        return None
    return ratio * alg_bytes


def log(msg):
    """Progress line on stderr (the JSON result line goes to stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def timed_steps(b, nflows, steps, moving):
    """Advances `steps` steps in one call; returns (seconds, CG iterations per
    step per flow, from the solver's counters, for moving bodies)."""
    import ibpm_amd
    sync = ibpm_amd.load().ibpm_ctx_synchronize
    sync(b.ctx.h)
    it0 = b.solver_stats()["iterations"] if moving else 0
    t0 = time.perf_counter()
    b.advance(steps)
    sync(b.ctx.h)
    sec = time.perf_counter() - t0
    its = (b.solver_stats()["iterations"] - it0) / nflows / steps if moving else None
    return sec, its


GOLDEN = os.path.join(ROOT, "tests", "golden")
# Oracle fixtures (FP64 CPU restatement, per-step Cd/Cl) of flows of the workloads:
# workload -> {flow index: fixture file}
FIXTURES = {
    "cylinder_200": {0: "cylinder_200.npz"},
    "cylinder_400x4": {0: "cylinder_400x4.npz"},
    "plate_512x5": {0: "plate_512x5_short.npz"},
    "tandem_16x1024x6": {0: "tandem_16x1024x6_c0.npz"},
    "mixed_64x3": {k: "mixed_64x3_f%d.npz" % k for k in (2, 5, 4, 20, 3, 11, 0, 1, 51)},
}
GATE = 1e-6


def accuracy_vs_oracle(b, workload):
    """Largest relative Cd and Cl error, over the steps both cover, of the
    flows of `workload` that have an oracle fixture (Cl relative to max|Cd|
    for the mirror-symmetric cylinders, where Cl is roundoff-seeded, and to
    max|Cl| otherwise)."""
    import numpy as np
    out = {"steps_compared": None, "Cd_rel_err": 0.0, "Cl_rel_err": 0.0, "flows": []}
    for k, name in sorted(FIXTURES.get(workload, {}).items()):
        path = os.path.join(GOLDEN, name)
        if not os.path.exists(path) or k >= len(b.flows):
            continue
        ref = np.load(path)["forces"]
        got = b.force_history(k)
        n = min(len(ref), len(got))
        if n == 0:
            continue
        r, g = ref[:n], got[:n]
        cd = float(np.max(np.abs(g[:, 0] - r[:, 0]) / np.abs(r[:, 0])))
        sym = "circle" in b.flows[k].geom   # mirror-symmetric cylinder: Cl is roundoff-seeded
        cl_scale = np.max(np.abs(r[:, 0])) if sym else np.max(np.abs(r[:, 1]))
        cl = float(np.max(np.abs(g[:, 1] - r[:, 1]) / cl_scale))
        out["flows"].append(k)
        out["steps_compared"] = n if out["steps_compared"] is None else min(n, out["steps_compared"])
        out["Cd_rel_err"] = max(out["Cd_rel_err"], cd)
        out["Cl_rel_err"] = max(out["Cl_rel_err"], cl)
    if not out["flows"]:
        return None
    out["gate"] = GATE
    out["ok"] = out["Cd_rel_err"] <= GATE and out["Cl_rel_err"] <= GATE
    return out


REBUILD_COST = {"all": 1.0, "finest": 0.45, "table": 0.1, "tracked": 0.1}
P_SOURCE = {"all": "M", "finest": "M's finest-level form",
            "table": "the kernel-table M at the build's points",
            "tracked": "the kernel-table M, refreshed every substep by a Newton-Schulz step"}


def moving_form(ctx, flows, warmup, steps, precond="dense", pairs=0, local=True, warm="previous",
                nn=None, build="tracked", rebuild_cost=0.1):
    """One form of the moving-body solver on `flows` from a zero initial
    vorticity: warmup steps (the first builds the preconditioner), then
    `steps` timed steps, with the preconditioner rebuilds that fall due inside
    them.  build / rebuild_cost: how P is built (Batch.set_precond_build).
    Returns the form's figures and the batch."""
    import ibpm_amd
    from ibpm_amd.workloads import grid_points
    b = ibpm_amd.Batch(ctx, flows).init()
    b.set_precond(precond, pairs).set_local_helmholtz(local)
    if precond != "none":
        b.set_precond_build(build, rebuild_cost)
    if nn is not None:
        b.nn_params(nn)
    if warm != "previous":
        b.set_warm_start(warm)
    sync = ibpm_amd.load().ibpm_ctx_synchronize
    sync(ctx.h)
    b.advance(max(1, warmup))
    st0 = b.solver_stats()
    sec, its = timed_steps(b, len(flows), steps, True)
    st1 = b.solver_stats()
    gp = sum(grid_points(f) for f in flows)
    d = {k: st1[k] - st0[k] for k in st1}
    ms = sec / steps * 1e3
    out = {"flows": len(flows), "precond": precond, "pairs": pairs if precond == "lowrank" else 0,
           "local_helmholtz": local, "warm_start": warm,
           "value": gp * steps / sec, "ms_per_step": ms, "cg_iters_per_step": its,
           "batched_m_applications_per_step": d["m_applications"] / steps, "fallbacks": d["fallbacks"],
           "host_ms_per_step_kinematics": round(d["host_us_kinematics"] / steps / 1e3, 3),
           "host_ms_per_step_polls": round(d["host_us_polls"] / steps / 1e3, 3)}
    if precond != "none":
        nb = d["builds"]
        out["preconditioner"] = {
            "build": build, "rebuild_cost": rebuild_cost,
            "first_build_s": round(st0["build_us"] / 1e6 / max(1, st0["builds"]), 3),
            "builds_in_timed_steps": nb,
            "build_s_each": round(d["build_us"] / 1e6 / nb, 3) if nb else None,
            "build_host_s_each": round(d["build_host_us"] / 1e6 / nb, 3) if nb else None,
            "build_m_applications_per_step": d["build_m_applications"] / steps,
            "share_of_timed_s": round(d["build_us"] / 1e6 / sec, 4)}
    log("form %s/%s/%s warm start, %d flows: %.1f ms/step over %d steps (%d rebuilds), %.1f CG iterations/step"
        % (precond, "local" if local else "dst", warm, len(flows), ms, steps, d["builds"], its or 0.0))
    return out, b


def streams_form(flows, warmup, steps, nstreams=2, build="tracked", rebuild_cost=0.1, moving=True):
    """The headline form split into `nstreams` batches of consecutive flows,
    each on a context (HIP stream) of its own, advanced at the same time from
    one host thread per batch (ibpm_amd.advance_concurrently): moving bodies
    with the preconditioned CG from a zero start, stationary ones by graph
    replay.  Returns the form's figures for the whole job."""
    import ibpm_amd
    from ibpm_amd.workloads import grid_points
    sync = ibpm_amd.load().ibpm_ctx_synchronize
    k = (len(flows) + nstreams - 1) // nstreams
    groups = [flows[i:i + k] for i in range(0, len(flows), k)]
    ctxs = [ibpm_amd.Context(0) for _ in groups]
    bs = []
    for c, g in zip(ctxs, groups):
        b = ibpm_amd.Batch(c, g).init()
        if moving:
            b.set_precond("dense", 0).set_local_helmholtz(True).set_precond_build(build, rebuild_cost)
            b.set_warm_start("zero")
        else:
            b.use_graph(True)
        bs.append(b)
    ibpm_amd.advance_concurrently(bs, max(1, warmup))
    for c in ctxs:
        sync(c.h)
    it0 = [b.solver_stats()["iterations"] if moving else 0 for b in bs]
    t0 = time.perf_counter()
    ibpm_amd.advance_concurrently(bs, steps)
    for c in ctxs:
        sync(c.h)
    sec = time.perf_counter() - t0
    its = (sum(b.solver_stats()["iterations"] - i for b, i in zip(bs, it0)) / len(flows) / steps) if moving else None
    for b in bs:
        b.close()
    for c in ctxs:
        c.close()
    gp = sum(grid_points(f) for f in flows)
    log("form %s, %d flows on %d streams: %.2f ms/step"
        % ("dense/local/zero warm start" if moving else "graph replay", len(flows), len(groups), sec / steps * 1e3))
    out = {"flows": len(flows), "streams": len(groups), "flows_per_stream": [len(g) for g in groups],
           "value": gp * steps / sec, "ms_per_step": sec / steps * 1e3}
    if moving:
        out.update({"warm_start": "zero", "cg_iters_per_step": its})
    return out


def plain_form(ctx, flow, warmup, steps):
    """The plain GPU form on one flow: plain CG from the previous force with
    every A^-1 of M by the DST solve (the reference's solver), no batching."""
    out, b = moving_form(ctx, [flow], warmup, steps, precond="none", local=False, build=None)
    b.close()
    return out


def train_network(ctx, flows, train_steps):
    """Trains the warm-start network on (kinematics, right-hand side, force)
    samples of a separate run of the workload's flows (its first train_steps
    steps), by the library's own training kernels; returns its parameters."""
    import ibpm_amd
    b = ibpm_amd.Batch(ctx, flows).init().set_precond("dense", 0).set_precond_build("tracked", 0.1)
    b.collect(True, 4 << 20)
    b.advance(train_steps)
    t0 = time.perf_counter()
    loss0 = b.train(iters=1, lr=0.0)
    loss = b.train(iters=800, lr=3e-3)
    info = {"train_steps": train_steps, "samples": b.num_samples(), "loss_before": loss0,
            "loss_after": loss, "train_s": round(time.perf_counter() - t0, 3)}
    params = b.nn_params()
    b.close()
    log("network trained on %d samples in %.1f s" % (info["samples"], info["train_s"]))
    return params, info


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="tandem_16x1024x6", choices=sorted(CONFIG_TEXT),
                    help="tandem_16x1024x6 (default): the moving-plate batch the metric is quoted on")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--precond", default="dense", choices=["none", "dense", "stale", "lowrank"],
                    help="force solve of moving bodies: plain CG (the reference's), CG preconditioned "
                         "by the dense FP64 P alone (dense; stale is its former name), or P plus the "
                         "low-rank update of --pairs Krylov pairs")
    ap.add_argument("--pairs", type=int, default=8)
    ap.add_argument("--warm", default="zero", choices=["previous", "zero", "learned"],
                    help="CG start of the headline moving-body form (previous = the reference's; the "
                         "other two are timed beside it unless --no-forms)")
    ap.add_argument("--train-steps", type=int, default=30,
                    help="steps of the training run of the learned warm start")
    ap.add_argument("--precond-build", default="tracked", choices=sorted(P_SOURCE),
                    help="P of the preconditioned forms: " + "; ".join("%s: %s" % kv for kv in sorted(P_SOURCE.items())))
    ap.add_argument("--rebuild-cost", type=float, default=None,
                    help="a build's cost in M applications per unknown, the rebuild rule's break-even "
                         "(default per build: " + ", ".join("%s %g" % kv for kv in sorted(REBUILD_COST.items())) + ")")
    ap.add_argument("--no-forms", action="store_true",
                    help="time only the headline form (moving bodies: no plain, batched-plain, other-start "
                         "or two-stream forms beside it; stationary batches: no two-stream form)")
    args = ap.parse_args()
    if args.rebuild_cost is None:
        args.rebuild_cost = REBUILD_COST[args.precond_build]
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import ibpm_amd
    from ibpm_amd.workloads import WORKLOADS, grid_points
    steps = args.steps or DEFAULT_STEPS[args.workload]
    flows = WORKLOADS[args.workload]()
    ctx = ibpm_amd.Context(0)
    moving = any("pitchplunge" in f.geom for f in flows)
    gp = sum(grid_points(f) for f in flows)
    forms = None
    learn = None
    if moving:
        nn = None
        if args.warm == "learned":
            nn, learn = train_network(ctx, flows, args.train_steps)
        t_i0 = time.perf_counter()
        head, b = moving_form(ctx, flows, args.warmup, steps, args.precond, args.pairs, True, args.warm,
                              nn, args.precond_build, args.rebuild_cost)
        t_init = time.perf_counter() - t_i0
        value, sec, cg = head["value"], head["ms_per_step"] * steps / 1e3, head["cg_iters_per_step"]
        accuracy = accuracy_vs_oracle(b, args.workload)
        cd, cl = b.forces(0)
        stats = b.kernel_stats(2)
        b.close()
        if not args.no_forms:
            forms = {}
            iters = {args.warm: cg}
            for w in ("zero", "previous", "learned"):
                if w == args.warm:
                    continue
                if w == "learned" and nn is None:
                    nn, learn = train_network(ctx, flows, args.train_steps)
                fw, bw = moving_form(ctx, flows, args.warmup, steps, args.precond, args.pairs, True, w,
                                     nn if w == "learned" else None, args.precond_build, args.rebuild_cost)
                bw.close()
                forms["batched_preconditioned_%s_start" % w] = fw
                iters[w] = fw["cg_iters_per_step"]
            if learn is not None:
                for w, it in iters.items():
                    learn["cg_iters_per_step_%s_start" % w] = it
            forms["batched_preconditioned_zero_start_two_streams"] = streams_form(
                flows, args.warmup, steps, 2, args.precond_build, args.rebuild_cost)
            bpl, bb = moving_form(ctx, flows, 1, 3, precond="none", local=False, build=None)
            bb.close()
            forms["batched_plain_cg"] = bpl
            plain = plain_form(ctx, flows[0], args.warmup, steps)
            forms["plain_gpu_form"] = plain
            forms["speedup_vs_plain_gpu_form"] = value / plain["value"]
            if "batched_preconditioned_learned_start" in forms:
                forms["speedup_learned_form_vs_plain_gpu_form"] = (
                    forms["batched_preconditioned_learned_start"]["value"] / plain["value"])
        solver = head
    else:
        b = ibpm_amd.Batch(ctx, flows)
        t_i0 = time.perf_counter()
        b.init()
        t_init = time.perf_counter() - t_i0
        if not args.no_graph:
            b.use_graph(True)
        b.advance(args.warmup)
        sec, cg = timed_steps(b, len(flows), steps, False)
        value = gp * steps / sec
        log("%d steps timed: %.3f ms/step" % (steps, sec / steps * 1e3))
        accuracy = accuracy_vs_oracle(b, args.workload)
        cd, cl = b.forces(0)
        solver = None
        stats = b.kernel_stats(3)
        b.close()
        if len(flows) > 1 and not args.no_forms and not args.no_graph:
            forms = {"graph_replay_two_streams": streams_form(flows, args.warmup, steps, 2, moving=False)}
    # dominant kernel: most time among the classes with algorithmic bytes
    dom = max(stats.items(), key=lambda kv: kv[1][1] if kv[1][2] else 0.0)
    name, (cnt, ms, by) = dom
    avg_s = ms / cnt / 1e3
    achieved = by / cnt / avg_s / 1e9
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
                   "form": ("batched, preconditioned (P from %s), %s warm start, rebuilds inside the timed steps"
                            % (P_SOURCE[args.precond_build], args.warm))
                           if moving else "graph replay",
                   "solver": solver, "learned": learn, "forms": forms,
                   "accuracy_vs_oracle_fixture": accuracy,
                   "init_s": round(t_init, 3), "Cd_flow0": cd, "Cl_flow0": cl},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_per_launch(args.workload, name, by / cnt),
                     "avg_launch_us": avg_s * 1e6, "bytes_per_launch": by / cnt},
        "kernels": {k: {"launches": c, "ms": round(m, 4), "GB/s": (bb / (m / 1e3) / 1e9) if (bb and m) else None}
                    for k, (c, m, bb) in sorted(stats.items(), key=lambda kv: -kv[1][1])},
    }
    if not args.no_cpu_baseline:
        log("CPU baseline")
        plain_its = forms["plain_gpu_form"]["cg_iters_per_step"] if forms and "plain_gpu_form" in forms else None
        line["cpu_baseline"] = cpu_baseline(flows, args.cpu_seconds, plain_its)
    print(json.dumps(line))
    if accuracy is not None and not accuracy["ok"]:
        print("bench.py: flow Cd/Cl error against the oracle fixture exceeds %g" % GATE, file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
