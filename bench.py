#!/usr/bin/env python3
"""Benchmark: particle-substeps/s of the adaptive water drop (about 1 M
particles from rest: a coarse pool settled into its adaptive equilibrium and a
fine drop of 100 k particles released above it; split, merge and kernel-width
spreading on, 4 solver iterations, dt = 1/120) on one MI355X, with the mean
density error of the timed substeps, the particles created and deleted in
them, the rooflines of the two largest kernels, and the learned warm start
against the plain solver by iterations to the 0.1 % error tolerance.

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

# Algorithmic bytes of each timeable kernel from the fixed byte model of the
# substep (unpadded elements: xyz 12 B, scalars 4 B, pair entry 8 B; inputs
# read and outputs written once; neighbour-side gathers assumed cache-served):
# (bytes per particle, per pair of the solver's list, per pair of the walk's
# list before the spreading filter).
BYTE_MODEL = {
    # density/lambda pass (36 B + 8 B/pair) plus the pair reads of the
    # position-delta pass (8 B/pair): the row pass carries both
    "solve_rows": (36.0, 16.0, 0.0),
    # position-delta pass, particle part (28 B) + the next box collision (28 B)
    "solve_apply": (56.0, 0.0, 0.0),
    "permute": (100.0, 0.0, 0.0),     # permute 48 B of state: index 4 + gather 48 + write 48
    "fsm12": (38.0, 16.0, 0.0),       # split/merge find, flood fill and nearest neighbour
    # the neighbour walk with the spreading fused: pair build 16 B/particle
    # (position, range) + 8 B per found pair (the reference writes the list),
    # spreading 24 B/particle (radius, target radius, old width, new width ...)
    # + 8 B per found pair read back + 8 B per kept pair written
    "walk_fill": (40.0, 8.0, 16.0),
}
# the neighbour walk of a scene without spreading (basic PBF, or kernel widths
# from boundary distances): the pair build alone, 16 B/particle + 8 B per pair
PLAIN_WALK = (16.0, 8.0, 0.0)


def byte_model(name: str, params) -> tuple:
    s = params["settings"]
    spread = not params["basic_pbf"] and not s["mBaseKernelWidthOnBoundaryDistance"]
    return PLAIN_WALK if (name == "walk_fill" and not spread) else BYTE_MODEL[name]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


# name of each workload in the committed profile summaries
PROFILE_KEY = {"water_drop": "water_drop_settled"}


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


def timed_run(params, state, steps: int, mode: str, model=None, iterations=None, tolerance=None, max_iterations=8):
    """`steps` substeps from `state` (after 2 untimed ones) with the plain solver
    or the learned warm start, with a fixed iteration count or iterating to a
    density-error tolerance; wall time, error and iterations applied."""
    import torch
    from apbf_amd import learned as Lm
    from apbf_amd.scene import Scene
    p = dict(params)
    if iterations is not None:
        p["solver_iterations"] = iterations
    s = Scene(p)
    s.set_state(state)
    s.set_error_measurement(True)
    if tolerance is not None:
        s.set_solver_tolerance(tolerance, max_iterations)
    if mode == "warm":
        Lm.set_mode(s, model, "warm")
    for _ in range(2):
        s.update(p["dt"])
    s.iteration_errors(reset=True)
    s.error_stats(reset=True)
    s.particle_substeps(reset=True)
    s.solver_stats(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.update(p["dt"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    stats = s.error_stats()
    ps, _ = s.particle_substeps()
    iters, subs, fallbacks = s.solver_stats()
    s.close()
    return {"warm_start": mode == "warm", "tolerance": tolerance,
            "fixed_iterations": None if tolerance is not None else p["solver_iterations"],
            "iterations_per_substep": iters / max(subs, 1), "warm_start_undone": fallbacks,
            "ms_per_step": el / steps * 1e3, "particle_substeps_per_s": ps / el,
            "mean_density_error": stats[0], "max_density_error": stats[1], "mean_abs_density_error": stats[3]}


def learned_comparison(scene, params, args, log):
    """Harvest (feature, total-lambda) pairs from plain substeps, train the
    warm-start network on the GPU, then run the plain solver and the warm
    start from the same state, each iterating until the mean density error of
    the positions an iteration starts from is within the tolerance: the
    iterations each needs per substep are the comparison."""
    from apbf_amd import learned as Lm
    model, rep = Lm.fit_warm_start(scene, params["dt"], harvest_substeps=args.learned_harvest,
                                   epochs=args.learned_epochs, damping=args.learned_damping)
    log(f"warm-start network trained on {rep['samples']} pairs in {rep['train_s']:.2f} s "
        f"(relative rms error {rep['relative_rms_error']:.3f})")
    st = scene.get_state(pairs=False)
    st["fluid_index"] = st["fluid_index"][: len(st["radius"])]
    steps = max(10, min(args.steps, 30))
    out = {"steps": steps, "training": rep, "by_tolerance": {}}
    for tol in (1e-3, 1e-4):
        plain = timed_run(params, st, steps, "off", tolerance=tol)
        warm = timed_run(params, st, steps, "warm", model, tolerance=tol)
        ratio = plain["iterations_per_substep"] / warm["iterations_per_substep"] \
            if warm["iterations_per_substep"] > 0 else None
        log(f"tolerance {tol:g}: plain {plain['iterations_per_substep']:.2f} iterations/substep "
            f"({plain['ms_per_step']:.2f} ms, err {plain['mean_density_error']:.2e}); learned "
            f"{warm['iterations_per_substep']:.2f} ({warm['ms_per_step']:.2f} ms, err "
            f"{warm['mean_density_error']:.2e}, undone {warm['warm_start_undone']})")
        out["by_tolerance"][f"{tol:g}"] = {"plain": plain, "learned": warm, "iteration_ratio": ratio}
    model.close()
    head = out["by_tolerance"]["0.001"]
    out.update(tolerance=1e-3, plain_iterations_to_tolerance=head["plain"]["iterations_per_substep"],
               learned_iterations_to_tolerance=head["learned"]["iterations_per_substep"],
               iteration_ratio=head["iteration_ratio"])
    return out


def measure_rooflines(scenes, steps: int, names, profile_key):
    """Roofline entries of the kernels `names`, timed with HIP events on
    every launch of the timed region in each of `scenes` ((scene, its
    substep totals) pairs, stepped on one stream). Every launch of these
    kernels covers its scene's whole substep, so the algorithmic bytes of the
    region are the byte model over the substep totals times the launches per
    substep; achieved = those bytes / the launches' summed duration."""
    out = {}
    for name in names:
        total_bytes, total_ms, launches = 0.0, 0.0, 0
        for sc, tot in scenes:
            per_p, per_pair, per_raw = byte_model(name, sc.params)
            k_ms, k_launches = sc.kernel_time_of(name)
            if not k_launches:
                continue
            substep_bytes = per_p * tot["particles"] + per_pair * tot["pairs"] + per_raw * tot["raw_pairs"]
            total_bytes += substep_bytes * k_launches / steps
            total_ms += k_ms
            launches += k_launches
        if not launches:
            continue
        avg_s = total_ms / 1e3 / launches
        ach = total_bytes / (total_ms / 1e3) / 1e9
        out[name] = {"bound": "hbm", "kernel": name, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": committed_traffic(name, profile_key),
                     "launches_per_step": launches / steps, "avg_launch_ms": avg_s * 1e3,
                     "ms_per_step": total_ms / steps, "algorithmic_bytes_per_launch": total_bytes / launches}
    return out


def run_batch(args, log):
    """64 independent scenes (10k-500k particles, log-uniform) stepped as one
    batch: eager (every scene's substep enqueued on its own stream), as a
    captured two-substep graph of all scenes, and packed (the scenes of each
    settings group in one particle list, one launch per stage)."""
    import torch
    from apbf_amd import config as cfg
    from apbf_amd.scene import Batch, PackedBatch, Scene
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
    batch.close()
    for s in scenes:
        s.close()
    # packed: the scenes of each settings group in one particle list
    packed_batch = PackedBatch(cases)
    for _ in range(args.warmup):
        packed_batch.update(dt)
    packed_batch.synchronize()
    packed_batch.particle_substeps(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        packed_batch.update(dt)
    packed_batch.synchronize()
    packed_s = time.perf_counter() - t0
    packed_ps = packed_batch.particle_substeps(reset=True)
    n_lists = len(packed_batch.groups)
    # kernel rooflines from a few more substeps with the lists on one stream
    # (durations of kernels that overlap on the device are not their own)
    timed = ("walk_fill", "solve_rows")
    lists = [g["scene"] for g in packed_batch.groups]
    one = torch.cuda.current_stream().cuda_stream
    roof_steps = 4
    for sc in lists:
        sc.set_stream(one)
        sc.substep_totals(reset=True)
        sc.time_kernels(timed, max_launches=roof_steps * 16)
    for _ in range(roof_steps):
        packed_batch.update(dt)
    packed_batch.synchronize()
    rooflines = measure_rooflines([(sc, sc.substep_totals()) for sc in lists], roof_steps, timed, "batch64")
    packed_batch.close()
    log(f"eager {eager_s / steps * 1e3:.2f} ms/substep, graph {graph_s / steps * 1e3:.2f} ms/substep, "
        f"packed {packed_s / steps * 1e3:.2f} ms/substep ({n_lists} lists)")
    eager = {"value": eager_ps / eager_s, "ms_per_step": eager_s / steps * 1e3}
    graph = {"value": graph_ps / graph_s, "ms_per_step": graph_s / steps * 1e3}
    packed = {"value": packed_ps / packed_s, "ms_per_step": packed_s / steps * 1e3, "lists": n_lists}
    best, mode = max(((eager, "eager streams"), (graph, "graph replay"), (packed, "packed lists")),
                     key=lambda m: m[0]["value"])
    out = {
        "metric": "particle-substeps/s, 64 independent scenes of 10k-500k particles stepped as one batch",
        "value": best["value"], "unit": "particle-substeps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": best["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+i32", "data": "synthetic",
        "config": {"workload": "batch64", "scenes": len(scenes), "particles": n_total,
                   "kinds": "alternating plain dam break / adaptive water drop", "seed": 0, "mode": mode},
        "eager": eager, "graph": graph, "packed": packed,
    }
    if mode == "packed lists" and rooflines:
        dominant = max(rooflines, key=lambda k: rooflines[k]["ms_per_step"] or 0.0)
        out["roofline"] = rooflines[dominant]
        out["rooflines"] = rooflines
    print(json.dumps(out), flush=True)


def workload(name: str):
    """(params, state, setup report) of a named workload."""
    from apbf_amd import config as cfg
    if name == "water_drop":
        return cfg.settled_water_drop()
    if name == "water_drop_lattice":  # the unsettled lattice start (parity tests use it)
        p, st = cfg.water_drop()
        return p, st, {}
    p, st = None, None
    if name == "dam_break":
        p, st = cfg.dam_break(n_side=32)
    elif name == "small_water_drop":
        p, st = cfg.small_water_drop()
    elif name == "large_pool":
        p, st = cfg.large_pool()
    else:
        raise SystemExit(f"unknown workload {name}")
    return p, st, {}


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


def cpu_baseline_threads(params, state, threads: int):
    """The oracle on `threads` host threads at once: each thread steps its own
    copy of the same scene by one substep (the oracle itself is the sequential
    schedule, so threads multiply scenes, not split one). The pair capacity
    is cut to 250 per particle (above the 212 per particle the walk peaks at)
    so that the copies fit in host memory."""
    import threading
    from oracle import oracle as O
    O.build()
    p = dict(params, cap_pairs=int(len(state["radius"]) * 250))
    scenes = []
    for _ in range(threads):
        sc = O.OracleScene(p)
        sc.set_state(state)
        scenes.append(sc)
    n = scenes[0].counts()[1]
    workers = [threading.Thread(target=sc.update, args=(params["dt"],)) for sc in scenes]
    t0 = time.perf_counter()
    for w in workers:
        w.start()
    for w in workers:
        w.join()
    dt = time.perf_counter() - t0
    for sc in scenes:
        sc.close()
    return {"value": n * threads / dt, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{threads} copies of the post-warmup scene, one substep each on its own thread, "
                      f"{n * threads} particle-substeps in {dt:.1f} s on '{cpu_model()}' (nproc {os.cpu_count()})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="water_drop")
    ap.add_argument("--timed-kernel", default=None, choices=sorted(BYTE_MODEL),
                    help="time only this kernel (default: the neighbour walk and the solver row pass)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="also time the oracle on this many host threads at once (one scene copy per thread)")
    ap.add_argument("--iterations", type=int, default=None,
                    help="override the workload's solver iterations (studies of iterations to tolerance)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start comparison")
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
    log = lambda msg: print(f"[bench] {msg}", file=sys.stderr, flush=True)
    torch.cuda.init()
    t_setup = time.perf_counter()
    params, state, setup = workload(args.workload)
    if args.iterations is not None:
        params["solver_iterations"] = args.iterations
    if args.hash_res is not None:
        params["hash_res_log2"] = args.hash_res
    dt = params["dt"]
    scene = Scene(params)
    scene.set_state(state)
    scene.set_error_measurement(True)
    log(f"scene {params.get('workload')} with {len(state['radius'])} particles ready "
        f"({time.perf_counter() - t_setup:.1f} s of set-up, not timed)")
    shape = ("library defaults: when rows average under 62 entries 4 walk lanes per particle and 8 solver lanes "
             "per row, otherwise a whole wavefront per particle and 16 solver lanes")
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
    n_start = scene.counts()[0]
    scene.iteration_errors(reset=True)
    scene.error_stats(reset=True)
    scene.substep_totals(reset=True)
    timed = (args.timed_kernel,) if args.timed_kernel else ("walk_fill", "solve_rows")
    scene.time_kernels(timed, max_launches=args.steps * (params["solver_iterations"] + 2))

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
    tot = scene.substep_totals()
    rooflines = measure_rooflines([(scene, tot)], args.steps, timed,
                                  PROFILE_KEY.get(args.workload, args.workload))
    scene.time_kernel(None)
    dominant = max(rooflines, key=lambda k: rooflines[k]["ms_per_step"] or 0.0)
    iters = params["solver_iterations"]
    before_it = scene.iteration_errors(max_iters=max(iters, 1)).tolist()[:iters]
    mean_err, max_err, last_err, mean_abs, err_steps = scene.error_stats()
    n_now, nt_now, np_now = scene.counts()
    diag = scene.diagnostics()
    if diag["overflow_flags"]:
        log(f"WARNING: capacity clamp hit (flags {diag['overflow_flags']}, peak pairs "
            f"{diag['peak_pairs_before_clamp']} of {params['cap_pairs']})")
    value = tot["particles"] / elapsed
    mean_n = tot["particles"] / args.steps

    learned = None
    if not args.no_learned:
        learned = learned_comparison(scene, params, args, log)

    cpu = cpu_threads = None
    if not args.no_cpu_baseline or args.cpu_threads:
        st = scene.get_state(pairs=False)
        st["fluid_index"] = st["fluid_index"][: len(st["radius"])]
        if not args.no_cpu_baseline:
            cpu = cpu_baseline(params, st, args.cpu_seconds)
        if args.cpu_threads:
            cpu_threads = cpu_baseline_threads(params, st, args.cpu_threads)

    # one more substep with the candidate counter: cells and candidates the walk tests
    scene.count_walk_candidates(True)
    scene.update(dt)
    cd = scene.diagnostics()
    scene.count_walk_candidates(False)
    walk_candidates = {"per_particle": cd["walk_candidates"] / max(cd["rows"], 1),
                       "per_walk_pair": cd["walk_candidates"] / max(cd["pairs_before_clamp"], 1),
                       "per_solver_pair": cd["walk_candidates"] / max(cd["pairs"], 1),
                       "substep": {"candidates": cd["walk_candidates"], "particles": cd["rows"],
                                   "walk_pairs": cd["pairs_before_clamp"], "solver_pairs": cd["pairs"]}}

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
            "setup": setup,
            "particles_start": int(n_start),
            "particles_end": int(n_now),
            "mean_particles": mean_n,
            "particles_created": tot["created"],
            "particles_deleted": tot["deleted"],
            "mean_pairs": tot["pairs"] / args.steps,
            "mean_walk_pairs_before_spreading": tot["raw_pairs"] / args.steps,
            "transfers_end": int(nt_now),
            "solver_iterations": iters,
            "dt": dt,
            "merge_duration_s": params["settings"]["mMergeDuration"],
            "hash_res_log2": params["hash_res_log2"],
            "solver": "plain (cold-start lambda), fixed iteration count",
            "launch_shape": shape,
            "peak_walk_pairs": diag["peak_pairs_before_clamp"],
            "pair_capacity": params["cap_pairs"],
            "capacity_clamped": bool(diag["overflow_flags"]),
        },
        "density_error": {"mean_of_substep_means": mean_err, "max": max_err,
                          "mean_abs_of_substep_means": mean_abs, "substeps": err_steps,
                          "mean_before_iteration": before_it, "tolerance": 1e-3,
                          "tolerance_met": bool(mean_err <= 1e-3)},
        "roofline": rooflines[dominant],
        "rooflines": rooflines,
        "cpu_baseline": cpu,
        "plain_vs_learned": learned,
        "walk_candidates": walk_candidates,
    }
    if cpu_threads:
        out["cpu_baseline_threads"] = cpu_threads
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
