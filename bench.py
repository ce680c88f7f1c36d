#!/usr/bin/env python3
"""Benchmark: particle-substeps/s of the adaptive water drop (1M particles:
900k coarse pool particles r = 1h and a 100k fine-particle drop r = 0.5h, split,
merge and kernel-width spreading on, 4 solver iterations, dt = 1/120) on one
MI355X, with the mean density error of the timed substeps.

One step = one substep (the reference's waterdrop::update) over the whole
scene, inputs resident in HBM. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")

# Algorithmic bytes of each timeable kernel, per particle and per neighbour
# pair of the solver's list, from the fixed byte model of the substep
# (unpadded elements: xyz 12 B, scalars 4 B, pair entry 8 B; inputs read and
# outputs written once; neighbour-side gathers assumed cache-served).
BYTE_MODEL = {
    # density/lambda pass (36 B + 8 B/pair) plus the pair reads of the
    # position-delta pass (8 B/pair): the row pass carries both
    "solve_rows": (36.0, 16.0),
    # position-delta pass, particle part (28 B) + the next box collision (28 B)
    "solve_apply": (56.0, 0.0),
    "permute": (100.0, 0.0),     # permute 48 B of state: index 4 + gather 48 + write 48
    "fsm12": (38.0, 16.0),       # split/merge find, flood fill and nearest neighbour
}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def committed_traffic(kernel: str, workload_name: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by its copy-kernel
    calibration), or None when no summary covers this kernel and workload."""
    path = os.path.join(PROFILES, f"{kernel}_{workload_name}_traffic.txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split()[1])
    except OSError:
        pass
    return None


def calibrate_copy(n_bytes: int = 1 << 30) -> None:
    """A plain device copy of known size (read n_bytes, write n_bytes), three
    times, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    src = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n_bytes, "copy_bytes_written": n_bytes, "launches": 3}), flush=True)


def timed_run(params, state, steps: int, mode: str, model=None, iterations=None):
    """`steps` substeps from `state` (after 2 untimed ones) with the plain solver
    or the learned warm start; wall time, error and per-iteration errors."""
    import torch
    from apbf_amd import learned as Lm
    from apbf_amd.scene import Scene
    p = dict(params)
    if iterations is not None:
        p["solver_iterations"] = iterations
    s = Scene(p)
    s.set_state(state)
    s.set_error_measurement(True)
    if mode == "warm":
        Lm.set_mode(s, model, "warm")
    for _ in range(2):
        s.update(p["dt"])
    s.iteration_errors(reset=True)
    s.error_stats(reset=True)
    s.particle_substeps(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.update(p["dt"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    before = s.iteration_errors(max_iters=p["solver_iterations"]).tolist()
    stats = s.error_stats()
    ps, _ = s.particle_substeps()
    s.close()
    return {"iterations": p["solver_iterations"], "warm_start": mode == "warm", "ms_per_step": el / steps * 1e3,
            "particle_substeps_per_s": ps / el, "mean_density_error": stats[0], "mean_abs_density_error": stats[3],
            "mean_error_before_iteration": before}


def learned_comparison(scene, params, args, log):
    """Harvest (feature, total-lambda) pairs from plain substeps, train the
    warm-start network on the GPU, then run the plain solver and the warm
    start from the same state."""
    from apbf_amd import learned as Lm
    model, rep = Lm.fit_warm_start(scene, params["dt"], harvest_substeps=args.learned_harvest,
                                   epochs=args.learned_epochs, damping=args.learned_damping)
    log(f"warm-start network trained on {rep['samples']} pairs in {rep['train_s']:.2f} s "
        f"(relative rms error {rep['relative_rms_error']:.3f})")
    st = scene.get_state(pairs=False)
    st["fluid_index"] = st["fluid_index"][: len(st["radius"])]
    steps = max(10, min(args.steps, 30))
    iters = params["solver_iterations"]
    plain = timed_run(params, st, steps, "off")
    warm = timed_run(params, st, steps, "warm", model, iterations=args.learned_iterations)
    log(f"plain {plain['ms_per_step']:.2f} ms err {plain['mean_density_error']:.2e}; "
        f"learned {warm['ms_per_step']:.2f} ms err {warm['mean_density_error']:.2e}")
    # iterations to the plain solver's error: every plain iteration count below
    # the workload's and every warm-started one up to it, from the same state
    table = {"plain": {iters: plain}, "learned": {args.learned_iterations: warm}}
    for k in range(1, iters + 1):
        if k not in table["plain"]:
            table["plain"][k] = timed_run(params, st, steps, "off", iterations=k)
        if k not in table["learned"]:
            table["learned"][k] = timed_run(params, st, steps, "warm", model, iterations=k)
    target = plain["mean_density_error"]
    plain_k = min(k for k, r in table["plain"].items() if r["mean_density_error"] <= target)
    warm_ks = [k for k, r in table["learned"].items() if r["mean_density_error"] <= target]
    warm_k = min(warm_ks) if warm_ks else None
    log("iterations to the plain error %.2e: plain %d, learned %s" % (target, plain_k, warm_k))
    model.close()
    brief = {side: {str(k): {"mean_density_error": r["mean_density_error"],
                             "mean_abs_density_error": r["mean_abs_density_error"], "ms_per_step": r["ms_per_step"]}
                    for k, r in sorted(rows.items())} for side, rows in table.items()}
    return {"steps": steps, "training": rep, "plain": plain, "learned": warm, "tolerance": 1e-3,
            "by_iterations": brief, "plain_error_target": target,
            "plain_iterations_to_target": plain_k, "learned_iterations_to_target": warm_k,
            "iteration_ratio": (plain_k / warm_k) if warm_k else None}


def run_batch(args, log):
    """64 independent scenes (10k-500k particles, log-uniform) stepped as one
    batch: eager (every scene's substep enqueued on its own stream) and as a
    captured two-substep graph of all scenes."""
    import torch
    from apbf_amd import config as cfg
    from apbf_amd.scene import Batch, Scene
    cases = cfg.batch_scenes(64)
    dt = cases[0][0]["dt"]
    scenes = []
    for p, st in cases:
        s = Scene(p)
        s.set_state(st)
        scenes.append(s)
    n_total = sum(len(st["radius"]) for _, st in cases)
    log(f"{len(scenes)} scenes, {n_total} particles")
    batch = Batch(scenes)
    for _ in range(args.warmup):
        batch.update(dt)
    batch.synchronize()
    steps = max(2, args.steps - args.steps % 2)
    for s in scenes:
        s.particle_substeps(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.update(dt)
    batch.synchronize()
    eager_s = time.perf_counter() - t0
    eager_ps = sum(s.particle_substeps(reset=True)[0] for s in scenes)
    batch.capture(dt)
    batch.replay(1)  # first launch of the graph, untimed
    batch.synchronize()
    for s in scenes:
        s.particle_substeps(reset=True)
    t0 = time.perf_counter()
    batch.replay(steps // 2)
    batch.synchronize()
    graph_s = time.perf_counter() - t0
    graph_ps = sum(s.particle_substeps(reset=True)[0] for s in scenes)
    log(f"eager {eager_s / steps * 1e3:.2f} ms/substep, graph {graph_s / steps * 1e3:.2f} ms/substep")
    eager = {"value": eager_ps / eager_s, "ms_per_step": eager_s / steps * 1e3}
    graph = {"value": graph_ps / graph_s, "ms_per_step": graph_s / steps * 1e3}
    best, mode = (eager, "eager streams") if eager["value"] >= graph["value"] else (graph, "graph replay")
    out = {
        "metric": "particle-substeps/s, 64 independent scenes of 10k-500k particles stepped as one batch",
        "value": best["value"], "unit": "particle-substeps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": best["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+i32", "data": "synthetic",
        "config": {"workload": "batch64", "scenes": len(scenes), "particles": n_total,
                   "kinds": "alternating plain dam break / adaptive water drop", "seed": 0, "mode": mode},
        "eager": eager, "graph": graph,
    }
    batch.close()
    for s in scenes:
        s.close()
    print(json.dumps(out), flush=True)


def workload(name: str):
    from apbf_amd import config as cfg
    if name == "water_drop":
        return cfg.water_drop()
    if name == "dam_break":
        return cfg.dam_break(n_side=32)
    if name == "small_water_drop":
        return cfg.small_water_drop()
    if name == "large_pool":
        return cfg.large_pool()
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(params, state, seconds: float):
    """The scalar oracle (single thread) on the same workload state, for a bounded
    number of substeps (at least one, then until `seconds` have passed)."""
    from oracle import oracle as O
    O.build()
    sc = O.OracleScene(params)
    sc.set_state(state)
    t0 = time.perf_counter()
    ps = 0
    steps = 0
    while True:
        n = sc.counts()[1]
        sc.update(params["dt"])
        ps += n
        steps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": ps / dt, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} substep(s) of the same scene from the post-warmup state, "
                      f"{ps} particle-substeps in {dt:.1f} s on 1 thread of '{cpu_model()}' "
                      f"(nproc {os.cpu_count()}), g++ -O2 -ffp-contract=off"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="water_drop")
    ap.add_argument("--timed-kernel", default="solve_rows", choices=sorted(BYTE_MODEL))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--iterations", type=int, default=None,
                    help="override the workload's solver iterations (studies of iterations to tolerance)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start comparison")
    ap.add_argument("--learned-iterations", type=int, default=1,
                    help="Jacobi iterations after the learned warm start")
    ap.add_argument("--learned-damping", type=float, default=1.0, help="scale of the predicted lambdas")
    ap.add_argument("--learned-epochs", type=int, default=12)
    ap.add_argument("--learned-harvest", type=int, default=6, help="plain substeps harvested for training")
    ap.add_argument("--hash-res", type=int, default=None,
                    help="override the workload's neighbour grid resolution (cells per axis = 2^res)")
    ap.add_argument("--launch-shape", default=None,
                    help="walk lanes,solver lanes (e.g. 8,16); every shape computes identical results")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the launch shape with the library's autotuner after the warm-up")
    ap.add_argument("--graph", action="store_true",
                    help="time HIP-graph replays of captured substeps instead of eager updates")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 1 GiB device copy (for scaling rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build is single-GPU (the multi-GPU design is "
              "domain decomposition, not implemented); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from apbf_amd.scene import Scene

    if args.workload == "batch64":
        run_batch(args, lambda msg: print(f"[bench] {msg}", file=sys.stderr, flush=True))
        return
    params, state = workload(args.workload)
    if args.iterations is not None:
        params["solver_iterations"] = args.iterations
    if args.hash_res is not None:
        params["hash_res_log2"] = args.hash_res
    dt = params["dt"]
    torch.cuda.init()
    scene = Scene(params)
    scene.set_state(state)
    scene.set_error_measurement(True)

    log = lambda msg: print(f"[bench] {msg}", file=sys.stderr, flush=True)
    log(f"scene {params.get('workload')} with {len(state['radius'])} particles ready")
    shape = "walk 8 lanes (4 when rows average under 62 entries), solver 16 lanes (library default)"
    if args.launch_shape:
        w, v = (int(x) for x in args.launch_shape.split(","))
        scene.set_launch_shape(w, v)
        shape = f"walk {w}, solver {v} lanes"
    for _ in range(args.warmup):
        scene.update(dt)
    if args.autotune:  # before the timed region; the tuner times every shape from this state and puts it back
        w, v, ms = scene.autotune(dt, substeps=2)
        log(f"autotuned launch shape: walk {w} lanes, solver {v} lanes ({ms:.3f} ms/substep)")
        shape = f"walk {w}, solver {v} lanes (autotuned)"
    torch.cuda.synchronize()
    log("warmup done")
    scene.iteration_errors(reset=True)
    scene.error_stats(reset=True)
    scene.particle_substeps(reset=True)
    scene.time_kernel(args.timed_kernel, max_launches=args.steps * (params["solver_iterations"] + 2))

    if args.graph:  # timing through events is not captured
        scene.time_kernel(None)
        scene.capture(dt)
        args.steps -= args.steps % 2
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if args.graph:
        scene.replay(args.steps // 2)
    else:
        for _ in range(args.steps):
            scene.update(dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()

    elapsed = t1 - t0
    log(f"timed {args.steps} substeps in {elapsed:.3f} s")
    ps, pairs = scene.particle_substeps()
    k_ms, k_launches = scene.kernel_time()
    scene.time_kernel(None)
    iters = params["solver_iterations"]
    before_it = scene.iteration_errors(max_iters=max(iters, 1)).tolist()[:iters]
    mean_err, max_err, last_err, mean_abs, err_steps = scene.error_stats()
    n_now, nt_now, np_now = scene.counts()
    diag = scene.diagnostics()
    if diag["overflow_flags"]:
        log(f"WARNING: capacity clamp hit (flags {diag['overflow_flags']}, peak pairs "
            f"{diag['peak_pairs_before_clamp']} of {params['cap_pairs']})")
    value = ps / elapsed

    # roofline of the timed kernel: algorithmic bytes per launch / mean launch time
    per_p, per_pair = BYTE_MODEL[args.timed_kernel]
    launches_per_step = k_launches / args.steps if args.steps else 0
    mean_n = ps / args.steps
    mean_pairs = pairs / args.steps
    bytes_per_launch = per_p * mean_n + per_pair * mean_pairs
    avg_launch_s = (k_ms / 1e3) / k_launches if k_launches else float("nan")
    achieved = bytes_per_launch / avg_launch_s / 1e9 if k_launches else None

    learned = None
    if not args.no_learned:
        learned = learned_comparison(scene, params, args, log)

    cpu = None
    if not args.no_cpu_baseline:
        st = scene.get_state(pairs=False)
        st["fluid_index"] = st["fluid_index"][: len(st["radius"])]
        cpu = cpu_baseline(params, st, args.cpu_seconds)

    out = {
        "metric": "particle-substeps/s, 1M adaptive particles at <=0.1% density error, plain vs learned",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+i32",
        "data": "synthetic",
        "config": {
            "workload": params.get("workload", args.workload),
            "particles_start": int(len(state["radius"])),
            "particles_end": int(n_now),
            "mean_particles": mean_n,
            "mean_pairs": mean_pairs,
            "transfers_end": int(nt_now),
            "solver_iterations": params["solver_iterations"],
            "dt": dt,
            "hash_res_log2": params["hash_res_log2"],
            "solver": "plain (cold-start lambda)",
            "launch_shape": shape,
            "peak_walk_pairs": diag["peak_pairs_before_clamp"],
            "pair_capacity": params["cap_pairs"],
            "capacity_clamped": bool(diag["overflow_flags"]),
        },
        "density_error": {"mean_of_substep_means": mean_err, "max": max_err,
                          "mean_abs_of_substep_means": mean_abs, "substeps": err_steps,
                          "mean_before_iteration": before_it},
        "roofline": {
            "bound": "hbm",
            "kernel": args.timed_kernel,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": committed_traffic(args.timed_kernel, args.workload),
            "launches_per_step": launches_per_step,
            "avg_launch_ms": avg_launch_s * 1e3 if k_launches else None,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": cpu,
        "plain_vs_learned": learned,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
