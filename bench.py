"""Benchmark: simulated milliseconds per wall-clock millisecond of the ABD Newton step on one
MI355X, on the 32x32 torus chain net (1024 links of 288 verts / 576 tris, four corners pinned,
FP64, dt 8 ms, dhat 1e-3, kappa 1e7, 16 Newton iterations, tolerance 1e-5).

Each method (plain: the reference's Newton step; warm: every PCG after a step's first Newton
iteration starts from the part of the previous direction its line search did not take; stall: a
scene's Newton loop ends after a step whose line search accepted alpha < 1e-8, a step pinned by
CCD, reported as stalled; learned: warm plus the first iterate of every step moved toward a learned
per-body prediction) runs the workload's steps from its initial state after a few untimed warmup
steps; `value` is the first method's rate, `methods` holds every method's rate, Newton iterations,
converged and stalled fractions, worst final residual and, for every method after the first, how
far its final state lies from the first's.  The CPU baseline is the project's FP64
oracle (oracle/, "port") timed on this host on a bounded sample of the same workload (whole steps
from the GPU's state at the middle of the timed run), single-threaded and with OpenMP.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_FLOPS = 78.6e12  # MI355X dense FP64
# the reference's own published rate for a workload, sim-ms per wall-ms: chains.json, converged
# cubic barrier, ~80 ms wall per 8 ms step on an RTX 3080 (README.md:38,47)
REF_PUBLISHED = {"ref_chain_144": 0.10}
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (each calibrated by a
# 512 MiB device copy in the same run), per workload and precision
TRAFFIC_FILES = {("chain_net_32x32", "fp64"): "profiles/chain_net_32x32_plain_fp64_hbm_bytes.json"}
TIMER_KERNEL = {"pcg_solve": "abd::k_pcg_persist<{t}>", "pcg_iter": "abd::k_pcg_cg<{t}>",
                "broad_phase_query": "abd::k_pair_queries", "contact_hess": "abd::k_contact_hess",
                "body_assemble": "abd::k_body_assemble", "offblock_sum": "abd::k_offblock_sum",
                "narrow_phase": "abd::k_cand_eval", "ccd": "abd::k_ccd", "energy": "abd::k_body_energy"}


def measured_traffic(workload, precision, timer):
    """HBM bytes per launch of the kernel behind `timer` from the committed counter summary."""
    path = TRAFFIC_FILES.get((workload, precision))
    if not path or timer not in TIMER_KERNEL or not os.path.exists(os.path.join(ROOT, path)):
        return None, None
    with open(os.path.join(ROOT, path)) as f:
        summary = json.load(f)
    name = TIMER_KERNEL[timer].format(t="double" if precision == "fp64" else "float")
    k = summary["kernels"].get(name)
    if not k:
        return None, None
    fb = k.get("fetch_bytes_profiled", k["fetch_bytes_mean"])
    wb = k.get("write_bytes_profiled", k["write_bytes_mean"])
    return float(fb + wb), path
METRIC = "sim-ms per wall-ms + Newton iters to 1e-5 residual, learned vs plain ABD"


def workload(name, precision="fp64"):
    from colb_amd import scene as S
    from colb_amd.params import north_star
    prec = {"fp64": 0, "mixed": 1}[precision]
    if name == "chain_net_32x32":
        sc = S.chain_net(n=32)
        p = north_star(dt=8e-3, dhat=1e-3, kappa=1e7, k_contact=1e20, max_newton=16, newton_tol=1e-5, precision=prec)
        return sc, p, {"workload": "chain_net_32x32", "links": 1024, "verts_per_link": 288,
                       "tris_per_link": 576, "pinned": 4, "dt_ms": 8.0, "dhat": 1e-3, "kappa": 1e7,
                       "k_contact": 1e20, "max_newton": 16, "newton_tol": 1e-5, "precision": precision}
    if name == "box_pile_4096":
        sc = S.box_pile(n=4096)
        q = sc.q0  # lowest vertex 5e-4 above the ground plane y = 0
        y = q[sc.vert_body, 1] + (sc.rest_x * q[sc.vert_body][:, [4, 7, 10]]).sum(1)
        sc.q0[:, 1] -= y.min() - 5e-4
        p = north_star(dt=8e-3, ground_enabled=True, dhat=1e-3, kappa=1e7, k_contact=1e20, precision=prec)
        return sc, p, {"workload": "box_pile_4096", "bodies": 4096, "edge": [0.2, 1.0], "seed": 4, "dt_ms": 8.0,
                       "dhat": 1e-3, "kappa": 1e7, "k_contact": 1e20, "ground": True, "precision": precision}
    if name == "hull_piles_256":
        sc = S.batched_hull_piles(n_scenes=256, seed=3, lo=4, hi=512)
        q = sc.q0  # lowest vertex of every pile 5e-4 above the ground plane y = 0
        y = q[sc.vert_body, 1] + (sc.rest_x * q[sc.vert_body][:, [4, 7, 10]]).sum(1)
        sc.q0[:, 1] -= y.min() - 5e-4
        p = north_star(dt=4e-3, ground_enabled=True, dhat=1e-3, kappa=1e7, k_contact=1e20, precision=prec)
        return sc, p, {"workload": "hull_piles_256", "scenes": 256, "bodies": int(sc.n_bodies), "seed": 3,
                       "bodies_per_scene": "log-uniform in [4, 512]", "dt_ms": 4.0, "dhat": 1e-3, "kappa": 1e7,
                       "k_contact": 1e20, "ground": True, "precision": precision}
    if name == "ref_chain_144":
        from colb_amd.params import reference_compat
        here = os.path.dirname(os.path.abspath(__file__))
        sc = S.load_json_scene(os.path.join(here, "tests", "golden", "ref_assets", "chains.json"))
        p = reference_compat(precision=prec)
        return sc, p, {"workload": "ref_chain_144", "links": int(sc.n_bodies), "scene": "the reference's chains.json",
                       "dt_ms": 8.0, "dhat": p.dhat, "kappa": p.kappa, "k_contact": p.k_contact,
                       "params": "reference_compat", "precision": precision}
    if name == "cubes_8":
        sc = S.cubes_on_ground()
        p = north_star(dt=8e-3, ground_enabled=True, dhat=1e-3, kappa=1e7, k_contact=1e20, precision=prec)
        return sc, p, {"workload": "cubes_8_on_ground", "bodies": 8, "dt_ms": 8.0, "dhat": 1e-3, "precision": precision}
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(sc, p, state, budget_s, threads):
    """The FP64 oracle (oracle/) on this host: whole steps from `state` until budget_s has passed."""
    from oracle import oracle as O
    O.build(force=False)
    orc = O.Oracle(sc, p, threads=threads)
    orc.set_state(*state)
    t0 = time.perf_counter()
    steps, iters = 0, 0
    while True:
        st = orc.step()[0]
        steps += 1
        iters += st["newton_iters"]
        if time.perf_counter() - t0 > budget_s:
            break
    wall = time.perf_counter() - t0
    return {"value": steps * p.dt * 1e3 / (wall * 1e3), "unit": "sim-ms/wall-ms", "cores": threads,
            "kind": "port", "wall_s": round(wall, 3), "steps": steps, "newton_iters_per_step": iters / steps}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


METHODS = {
    # name -> (parameter changes, description)
    "plain": ({}, "the reference's Newton step: first iterate q_now (abd.py:674-699)"),
    "extrapolate": ({"predictor": 1}, "first iterate moved toward q_now + dt qdot (CCD-safe, energy-decreasing)"),
    "warm": ({"pcg_x0": 1}, "PCG from the untaken part (1 - alpha) du of the previous Newton direction"),
    "stall": ({"stall_alpha": 1e-8},
              "the reference's step, a scene stopped once its accepted line-search step falls below 1e-8 (CCD-pinned)"),
    "warm_stall": ({"pcg_x0": 1, "stall_alpha": 1e-8}, "warm and stall together"),
    "learned": ({"predictor": 3, "pcg_x0": 1},
                "warm, and the first iterate moved toward the learned per-body prediction of the step"),
}


def train_learned(sc, p, args):
    """The learned predictor, trained on the GPU on the (features, step) pairs of the converged steps
    that plain runs of perturbed copies of the workload log (other initial velocities: trajectories
    the timed run never sees; a step that ends at the Newton cap is no target)."""
    from colb_amd.abd import NewtonAbd
    from colb_amd.learned import StepPredictor
    free = int((sc.body_pinned == 0).sum())  # rows per step (one scene)
    Fs, Ds, conv_steps = [], [], 0
    t0 = time.perf_counter()
    for run in range(args.train_runs):
        rng = np.random.default_rng(12345 + run)
        qd = sc.qdot0.copy()
        qd[:, :3] += rng.normal(scale=0.05, size=qd[:, :3].shape) * (1 - sc.body_pinned[:, None])
        g = NewtonAbd(p.dt, sc, params=p)
        g.set_state(sc.q0, sc.q0, qd)
        g.log_steps(True)
        conv = []
        for _ in range(args.train_steps):
            sts = g.step()
            conv.append(all(x["converged"] for x in sts))
        F, D = g.step_samples()
        g.close()
        keep = np.repeat(np.array(conv, dtype=bool), free) if sc.n_scenes == 1 else np.ones(len(F), dtype=bool)
        Fs.append(F[keep])
        Ds.append(D[keep])
        conv_steps += int(sum(conv))
    F, D = np.concatenate(Fs), np.concatenate(Ds)
    t_log = time.perf_counter() - t0
    sp = StepPredictor(seed=0)
    t1 = time.perf_counter()
    losses = sp.fit(F, D, p.dt, steps=args.train_iters, batch=1024, lr=1e-3)
    info = {"model": "49-128-128-128-12 MLP, BF16 weights, FP32 accumulation", "training_rows": int(len(F)),
            "training_runs": args.train_runs, "training_steps_logged": args.train_runs * args.train_steps,
            "training_steps_converged": conv_steps, "adam_steps": args.train_iters,
            "loss_first": round(float(losses[:20].mean()), 4), "loss_last": round(float(losses[-100:].mean()), 4),
            "log_s": round(t_log, 2), "train_s": round(time.perf_counter() - t1, 2)}
    return sp, info


def run_method(name, sc, p, args, learned_model=None, timing=False):
    from colb_amd.abd import NewtonAbd
    changes, desc = METHODS[name]
    pm = p.replace(**changes)
    solver = NewtonAbd(pm.dt, sc, params=pm)
    if learned_model is not None:
        solver.set_predictor(learned_model)
    for k in range(args.warmup):
        solver.step()
    tune = None
    if args.pcg_layout:
        parts = [int(x) for x in args.pcg_layout.split(",")]
        solver.set_pcg_layout(parts[0], bool(parts[1]) if len(parts) > 1 else False)
        tune = {"bodies_per_tile": parts[0], "per_iteration_launch": bool(parts[1]) if len(parts) > 1 else False,
                "fixed": True}
    elif not args.no_autotune and args.solver != "direct":
        bpt, multi, ms = solver.autotune(repeats=3)
        tune = {"bodies_per_tile": bpt, "per_iteration_launch": bool(multi),
                "ms_per_solve": [[round(float(v), 4) for v in row] for row in ms]}
    if not args.from_warmup:
        # the timed run starts from the workload's initial state (buffers already sized by the warmup)
        solver.set_state(sc.q0, sc.q0, sc.qdot0)
        solver.set_stiffness(np.full(sc.n_bodies, pm.k_contact))
    if timing:
        solver.enable_timing(2)  # device wall-clock stamps of the timed steps' PCG solves only
    recs = []
    mid = None
    paused = 0.0
    t0 = time.perf_counter()
    for k in range(args.steps):
        sts = solver.step()
        recs.append((max(x["newton_iters"] for x in sts), all(x["converged"] for x in sts),
                     max(x["residual_last"] for x in sts), max(x["cg_iters"] for x in sts),
                     sts[0]["traversals"], sts[0]["detections"], sum(x["pcg_fallbacks"] for x in sts),
                     np.mean([x["stalled"] for x in sts])))
        if k == args.steps // 2 - 1 and timing:
            tp = time.perf_counter()
            mid = solver.get_state()  # the CPU baseline's start state, copied outside the timed work
            paused += time.perf_counter() - tp
    t1 = time.perf_counter()
    wall_ms = (t1 - t0 - paused) * 1e3
    final_q = solver.get_state()[0]  # after the clock stops: the trajectory comparison across methods
    newton = np.array([r[0] for r in recs])
    conv = np.array([r[1] for r in recs])
    out = {"description": desc, "sim_ms_per_wall_ms": round(args.steps * pm.dt * 1e3 / wall_ms, 6),
           "ms_per_step": round(wall_ms / args.steps, 3), "newton_iters_per_step": round(float(newton.mean()), 3),
           "ms_per_newton_iteration": round(wall_ms / max(int(newton.sum()), 1), 4),
           "converged_fraction": round(float(conv.mean()), 4),
           "stalled_fraction": round(float(np.mean([r[7] for r in recs])), 4),
           "worst_final_residual": float(max(r[2] for r in recs)),
           "cg_iters_per_step": round(float(np.mean([r[3] for r in recs])), 1), "steps": args.steps,
           "traversals_per_step": round(float(np.mean([r[4] for r in recs])), 3),
           "detections_per_step": round(float(np.mean([r[5] for r in recs])), 3),
           "pcg_fallbacks": int(sum(r[6] for r in recs)), "autotune": tune}
    if timing:
        out["pcg_times"] = solver.kernel_times()  # the timed steps' PCG solves (device wall clock)
        solver.enable_timing(0)
    print(f"# {name}: {out}", file=sys.stderr, flush=True)
    return solver, out, mid, final_q


def kernel_roofline(solver, out, cfg, precision, steps):
    """Roofline of the dominant kernel, the PCG solve: its times over the timed steps (out["pcg_times"]:
    device wall-clock stamps of the persistent solves, HIP events of per-iteration launches) give the
    average launch duration and the algorithmic bytes per launch.  Then
    a per-stage breakdown over `steps` extra steps after the timed region (printed only)."""
    kt = out.pop("pcg_times", {}) or {}
    dom = next((k for k in ("pcg_solve", "pcg_iter") if k in kt and kt[k]["calls"] > 0), None)
    roof = None
    if dom == "pcg_iter" and cfg.get("scenes", 1) > 1:
        # a per-iteration launch skips the scenes that have converged, so the whole batch's operator
        # bytes overstate what the launch moves
        print("# roofline: per-iteration PCG launches over a batch of scenes; not reported", file=sys.stderr)
        dom = None
    if dom:
        d = kt[dom]
        avg_ms = d["ms"] / d["calls"]
        bytes_per = d["bytes"] / d["calls"]
        ach = bytes_per / (avg_ms * 1e-3) / 1e9
        traffic, src = measured_traffic(cfg["workload"], precision, dom)
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": traffic, "avg_us": round(avg_ms * 1e3, 3),
                "launches": d["calls"], "bytes_per_launch": bytes_per, "traffic_source": src,
                "timed_over": ("the timed steps: device wall-clock stamps at each persistent solve's first "
                               "workgroup start and last workgroup end (graph replays included)"
                               if dom == "pcg_solve" else "the timed steps (HIP events on the solver's stream)")}
    pcg_ms = sum(kt[k]["ms"] for k in ("pcg_solve", "pcg_iter") if k in kt)
    timed_cg = out["cg_iters_per_step"] * out["steps"]
    pcg_us = round(pcg_ms * 1e3 / timed_cg, 3) if timed_cg else None
    if steps > 0:
        solver.enable_timing(True)
        for _ in range(steps):
            solver.step()
        stages = solver.kernel_times()
        solver.enable_timing(False)
        for k, v in stages.items():
            print(f"# stage {k:20s} {v['ms']:9.3f} ms {v['calls']:6d} calls "
                  f"{v['bytes'] / max(v['ms'], 1e-9) / 1e6:8.1f} GB/s", file=sys.stderr, flush=True)
    return roof, pcg_us


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed steps from the initial state (the C2 run: 200)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps that size the buffers before the run")
    ap.add_argument("--workload", default="chain_net_32x32", choices=["chain_net_32x32", "box_pile_4096",
                                                                   "hull_piles_256", "ref_chain_144", "cubes_8"])
    ap.add_argument("--methods", default="plain,warm,stall,learned",
                    help="comma-separated: plain (the reference's step), extrapolate, warm, stall, warm_stall, "
                         "learned (trained first)")
    ap.add_argument("--solver", default="pcg", choices=["pcg", "direct", "fallback"],
                    help="linear solve: block-Jacobi PCG, dense Cholesky per scene, or PCG with the direct "
                         "solve for scenes whose PCG ends at its iteration cap")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="mixed: FP32 PCG, FP64 near-contact distances, CCD and energies")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of oracle stepping per CPU leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--train-steps", type=int, default=100, help="steps of each perturbed run the learned model uses")
    ap.add_argument("--train-runs", type=int, default=2, help="perturbed runs logged for the learned model")
    ap.add_argument("--train-iters", type=int, default=3000, help="Adam steps of the learned model")
    ap.add_argument("--pcg-layout", default=None, metavar="BPT[,multi]",
                    help="fix the PCG layout (bodies per tile, 1 = one launch per iteration) instead of autotuning")
    ap.add_argument("--from-warmup", action="store_true",
                    help="time the steps that follow the warmup instead of restarting from the initial state "
                         "(for a workload whose contact-rich phase comes later, e.g. the hull piles once landed)")
    ap.add_argument("--no-graph", action="store_true",
                    help="eager launches instead of replaying each Newton iteration from a captured HIP graph "
                         "(the same kernels; for kernel tracers that do not follow graph launches)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 512 MiB device-to-device copy (reference kernel for --pmc byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build targets one GPU; --gpus must be 1 (multi-GPU sharding is design-only)"}))
        sys.exit(2)
    methods = [m for m in args.methods.split(",") if m]
    for m in methods:
        if m not in METHODS:
            raise SystemExit(f"unknown method {m}")
    if args.calib_copy:
        import torch
        a = torch.empty(512 << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    sc, p, cfg = workload(args.workload, args.precision)
    p = p.replace(solver={"pcg": 0, "direct": 1, "fallback": 2}[args.solver])
    cfg["solver"] = args.solver
    if args.no_graph:
        p = p.replace(debug=p.debug | 64)
    cfg["graph"] = not args.no_graph
    cfg["timed_from"] = f"the state after {args.warmup} warmup steps" if args.from_warmup else "the initial state"
    print(f"# workload {cfg['workload']}: {sc.n_bodies} bodies, {sc.n_verts} verts, {sc.n_tris} tris, "
          f"{sc.n_edges} edges", file=sys.stderr, flush=True)
    results, learned_info, mid_state, roof, pcg_us, first_solver = {}, None, None, None, None, None
    for m in methods:
        model = None
        if m == "learned":
            model, learned_info = train_learned(sc, p, args)
            print(f"# learned: {learned_info}", file=sys.stderr, flush=True)
        solver, out, mid, final_q = run_method(m, sc, p, args, model, timing=(m == methods[0]))
        results[m] = out
        if m == methods[0]:
            first_q = final_q
        else:
            # the same steps from the same initial state: how far this method's trajectory ends from the first's
            out[f"final_q_rel_diff_vs_{methods[0]}"] = float(np.abs(final_q - first_q).max() / np.abs(first_q).max())
        if m == methods[0]:
            mid_state = mid
            roof, pcg_us = kernel_roofline(solver, out, cfg, args.precision, args.profile_steps)
            out["pcg_us_per_iteration"] = pcg_us
        solver.close()
        if model is not None:
            model.mlp.close()  # device memory released before the interpreter's teardown
    head = results[methods[0]]
    cpu = None
    if not args.no_cpu_baseline and mid_state is not None:
        threads = min(16, len(os.sched_getaffinity(0)))
        single = cpu_baseline(sc, p, mid_state, args.cpu_budget, 1)
        multi = cpu_baseline(sc, p, mid_state, args.cpu_budget, threads) if threads > 1 else single
        sample = (f"{{n}} step(s) of the same workload from the GPU's state at timed step {args.steps // 2} "
                  f"(the middle of the run), FP64 oracle")
        cpu = dict(multi, sample=sample.format(n=multi["steps"]) + f", OpenMP over {threads} threads",
                   cpu_model=cpu_model(),
                   single_thread={"value": single["value"], "cores": 1, "steps": single["steps"],
                                  "wall_s": single["wall_s"], "newton_iters_per_step": single["newton_iters_per_step"],
                                  "sample": sample.format(n=single["steps"]) + ", one thread"})
        print(f"# cpu baseline: {cpu}", file=sys.stderr, flush=True)
    ratio = None
    if "learned" in results and "plain" in results:
        ratio = round(results["learned"]["sim_ms_per_wall_ms"] / results["plain"]["sim_ms_per_wall_ms"], 4)
    vs_plain = {m: round(results[m]["sim_ms_per_wall_ms"] / results["plain"]["sim_ms_per_wall_ms"], 4)
                for m in results if m != "plain"} if "plain" in results else None
    result = {
        "metric": METRIC, "value": head["sim_ms_per_wall_ms"], "unit": "sim-ms/wall-ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": (round(head["sim_ms_per_wall_ms"] / REF_PUBLISHED[args.workload], 3)
                        if args.workload in REF_PUBLISHED else None),
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": ("the reference's own scene (chains.json and its meshes)" if args.workload == "ref_chain_144"
                 else "synthetic (procedural scene)"), "config": cfg,
        "method": methods[0], "newton_iters_per_step": head["newton_iters_per_step"],
        "converged_fraction": head["converged_fraction"], "worst_final_residual": head["worst_final_residual"],
        "cg_iters_per_step": head["cg_iters_per_step"], "methods": results, "learned_vs_plain": ratio,
        "vs_plain": vs_plain,
        "learned": learned_info, "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
