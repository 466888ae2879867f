"""Benchmark: simulated milliseconds per wall-clock millisecond of the ABD Newton step on one
MI355X, on the 32x32 torus chain net (1024 links of 288 verts / 576 tris, four corners pinned,
FP64, dt 8 ms, dhat 1e-3, kappa 1e7, 16 Newton iterations, tolerance 1e-5).

Prints one JSON line (see the `result` dict).  The CPU baseline is the project's FP64 oracle
(oracle/, "port") timed on this host on a bounded sample of the same workload: it starts from the
GPU's state at the beginning of the timed region and runs whole steps until ~20 s have passed.
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
            "kind": "port", "wall_s": round(wall, 3), "newton_iters_per_step": iters / steps,
            "sample": f"{steps} step(s) of the same workload from the GPU's state at the start of the timed "
                      f"region, FP64 oracle, OpenMP over {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="chain_net_32x32", choices=["chain_net_32x32", "box_pile_4096", "hull_piles_256", "ref_chain_144",
                                                                   "cubes_8"])
    ap.add_argument("--solver", default="pcg", choices=["pcg", "direct", "fallback"],
                    help="linear solve: block-Jacobi PCG, dense Cholesky per scene, or PCG with the direct "
                         "solve for scenes whose PCG ends at its iteration cap")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="mixed: FP32 PCG, FP64 near-contact distances, CCD and energies")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--method", default="plain", choices=["plain", "learned"],
                    help="learned: train the warm start on the iterates logged during warmup (on the GPU), "
                         "then time the steps with it")
    ap.add_argument("--pcg-layout", default=None, metavar="BPT[,multi]",
                    help="fix the PCG layout (bodies per tile, 1 = one launch per iteration) instead of autotuning")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 512 MiB device-to-device copy (reference kernel for --pmc byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build targets one GPU; --gpus must be 1 (multi-GPU sharding is design-only)"}))
        sys.exit(2)

    from colb_amd.abd import NewtonAbd
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
    print(f"# workload {cfg['workload']}: {sc.n_bodies} bodies, {sc.n_verts} verts, {sc.n_tris} tris, "
          f"{sc.n_edges} edges", file=sys.stderr, flush=True)
    solver = NewtonAbd(p.dt, sc, params=p)
    learned = None
    if args.method == "learned":
        solver.log_samples(True)
    for k in range(args.warmup):
        st = solver.step()[0]
        print(f"# warmup {k}: {st}", file=sys.stderr, flush=True)
    if args.method == "learned":
        from colb_amd.learned import WarmStart
        F, D = solver.samples()
        solver.log_samples(False)
        t_fit = time.perf_counter()
        ws = WarmStart(seed=0)
        losses = ws.fit(F, D, steps=2000, batch=512, lr=1e-3)
        learned = {"model": "warm start 49-128-128-128-12, BF16 weights", "training_rows": int(len(F)),
                   "adam_steps": 2000, "loss_first": round(float(losses[:20].mean()), 4),
                   "loss_last": round(float(losses[-100:].mean()), 4),
                   "train_s": round(time.perf_counter() - t_fit, 2)}
        solver.set_warm_start(ws)
        print(f"# learned: {learned}", file=sys.stderr, flush=True)
    tune = None
    if args.pcg_layout:
        parts = [int(x) for x in args.pcg_layout.split(",")]
        solver.set_pcg_layout(parts[0], bool(parts[1]) if len(parts) > 1 else False)
        tune = {"bodies_per_tile": parts[0], "per_iteration_launch": bool(parts[1]) if len(parts) > 1 else False,
                "fixed": True}
    elif not args.no_autotune and args.solver != "direct":
        # choose PCG tile size and launch form at the warmed-up state (outside the timed region)
        st_keep = solver.get_state()
        bpt, multi, ms = solver.autotune(repeats=3)
        solver.set_state(*st_keep)
        tune = {"bodies_per_tile": bpt, "per_iteration_launch": bool(multi),
                "ms_per_solve": [[round(float(v), 4) for v in row] for row in ms]}
        print(f"# autotune: {tune}", file=sys.stderr, flush=True)
    state0 = solver.get_state()
    iters, cg = [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        sts = solver.step()
        st = sts[0]
        # several scenes step together: the batch runs as long as its slowest scene
        iters.append(max(x["newton_iters"] for x in sts))
        cg.append(max(x["cg_iters"] for x in sts))
    t1 = time.perf_counter()
    wall_ms = (t1 - t0) * 1e3
    ms_per_step = wall_ms / args.steps
    sim_per_wall = args.steps * p.dt * 1e3 / wall_ms
    print(f"# timed: {ms_per_step:.2f} ms/step, newton {np.mean(iters):.2f}, cg {np.mean(cg):.1f}, "
          f"last {st}", file=sys.stderr, flush=True)

    # per-kernel HIP-event timing over a few extra steps (kept out of the timed region)
    solver.enable_timing(True)
    prof_cg = 0
    for k in range(args.profile_steps):
        prof_cg += sum(st_["cg_iters"] for st_ in solver.step())
    kt = solver.kernel_times()
    solver.enable_timing(False)
    single = ["pcg_solve", "pcg_iter", "broad_phase_query", "narrow_phase", "contact_hess", "offblock_sum", "body_assemble",
              "ccd", "energy"]
    dom = max((k for k in kt if k in single), key=lambda k: kt[k]["ms"], default=None)
    roof = None
    if dom:
        d = kt[dom]
        avg_ms = d["ms"] / d["calls"]
        bytes_per = d["bytes"] / d["calls"]
        ach = bytes_per / (avg_ms * 1e-3) / 1e9
        traffic, src = measured_traffic(cfg["workload"], args.precision, dom)
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": traffic, "avg_us": round(avg_ms * 1e3, 3),
                "bytes_per_launch": bytes_per, "traffic_source": src}
    direct = None
    if "direct_factor" in kt:
        # dense Cholesky flops per factorization: sum over scenes of N^3 / 3 (N = 12 x bodies padded to 32)
        nbs = np.bincount(sc.body_scene, minlength=sc.n_scenes)
        N = (12 * nbs + 31) // 32 * 32
        flops = float(np.sum(N.astype(np.float64) ** 3) / 3.0)
        f, s_ = kt["direct_factor"], kt.get("direct_solve", {"ms": 0.0, "calls": 1})
        fus = f["ms"] * 1e3 / f["calls"]
        direct = {"factor_us": round(fus, 2), "solve_us": round(s_["ms"] * 1e3 / s_["calls"], 2),
                  "factorizations": f["calls"], "order": [int(N.min()), int(N.max())],
                  "gflops": round(flops / (fus * 1e-6) / 1e9, 2),
                  "frac_fp64_peak": round(flops / (fus * 1e-6) / FP64_PEAK_FLOPS, 5)}
    pcg_ms = sum(kt[k]["ms"] for k in ("pcg_solve", "pcg_iter") if k in kt)
    pcg_us_per_iter = round(pcg_ms * 1e3 / prof_cg, 3) if prof_cg else None
    for k, v in kt.items():
        print(f"# kernel {k:20s} {v['ms']:9.3f} ms {v['calls']:6d} calls {v['bytes'] / max(v['ms'], 1e-9) / 1e6:8.1f} GB/s",
              file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        threads = len(os.sched_getaffinity(0))
        cpu = cpu_baseline(sc, p, state0, args.cpu_budget, threads)
        print(f"# cpu baseline: {cpu}", file=sys.stderr, flush=True)

    result = {
        "metric": METRIC, "value": round(sim_per_wall, 6), "unit": "sim-ms/wall-ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": (round(sim_per_wall / REF_PUBLISHED[args.workload], 3) if args.workload in REF_PUBLISHED
                        else None),
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": ("the reference's own scene (chains.json and its meshes)" if args.workload == "ref_chain_144"
                 else "synthetic (procedural scene)"), "config": cfg,
        "newton_iters_per_step": round(float(np.mean(iters)), 3), "cg_iters_per_step": round(float(np.mean(cg)), 1),
        "method": args.method, "learned": learned, "autotune": tune, "pcg_us_per_iteration": pcg_us_per_iter,
        "roofline": roof, "direct": direct, "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
