#!/usr/bin/env python3
"""PBF solver step benchmark on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload dam_break_262k] [--no-cpu]

A step is one pbf_step of the whole solver (predict, hash, counting sort + scan, reorder,
neighbour build, I Jacobi iterations, curl, vorticity + XSPH) over the synthetic dam-break
state, resident in HBM.  W untimed steps, then K steps timed with HIP events recorded on the
solver's stream (hipGraph replay).  Prints ONE JSON line.

roofline: the dominant kernel group of an eagerly launched step (HIP events on the solver's
stream around it), with its algorithmic bytes per launch (fp32 compulsory traffic; neighbour
gathers counted as cache hits); traffic: the HBM bytes per launch of that kernel from the
committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary for this workload, when there is one.
cpu_baseline: the oracle's C++ restatement of the reference step on a bounded sample of the same
workload, all host threads.  learned: the warm start trained and evaluated on the GPU (record,
train, error curves from identical states, graph-replayed step times), for the 262k dam break
and for a 32^3 dam break in a 3.2 x 2.0 x 1.6 m box.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pbf_amd import FLAG_FAST_MATH, FLAG_GRAPH, DeviceBuffer, Solver, driver, learned, lib, scenes  # noqa: E402

METRIC = "ms/step + particle-steps/s @262k; time to 0.5% density err, learned vs plain"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # dense BF16 MFMA
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per 4 cycles each, at 2.4 GHz
VALU_PEAK = 256 * 4 * 2.4e9 / 4
# Reference's published figure: 262144 particles, 0.012-0.015 s per cudaPBFSolve on an RTX 3090
# (README.md:9); the faster end as particle-steps/s.
PUBLISHED_PSPS = 262144 / 0.012

WORKLOADS = {
    "dam_break_262k": scenes.config_262k,
    "cpu_32k": scenes.config_cpu_32k,
    "double_dam_break_1m": scenes.config_1m,
    "batch_24_scenes": scenes.batch_config,
}


def sequential_ms(scene_list, warmup, steps):
    """Sum over the scenes of the graph-replayed step time of each scene alone."""
    tot = 0.0
    for sc in scene_list:
        s = Solver.from_scene(sc, flags=FLAG_GRAPH)
        s.advance(warmup)
        s.synchronize()
        tot += s.time_steps(steps) / steps
        s.close()
    return tot


def algorithmic_bytes(stage: str, n: int, ncells: int) -> float:
    """Compulsory fp32 bytes per launch group (12 B per xyz stream, 4 B per scalar)."""
    per_particle = {
        "predict_hash": 64,  # R x, v; W x_pred, v, cell
        "scan": 0,  # per cell, below
        "scatter_rank_reorder": 76,  # R cell id, 3 gathered xyz streams; W 3 xyz streams, id
        "neighbors_lambda": 16,  # R x; W lambda
        "lambda": 16,  # R x; W lambda
        "delta_apply": 28,  # R x, lambda; W x
        "curl": 36,  # R x, v; W curl
        "vort_apply": 48,  # R x, v, curl; W v
        "vort_apply_hash": 56,  # R x, v, curl; W v, next step's cell and bucket place
    }
    b = per_particle.get(stage, 0) * n
    if stage == "scan":
        b += 12 * ncells  # R count, W start, clear count
    return float(b)


def step_bytes(n: int, ncells: int, iters: int) -> float:
    """Whole-step algorithmic bytes: N*(252 + 44*I) + 12*C."""
    return float(n * (252 + 44 * iters) + 12 * ncells)


# HIP kernel behind each timed stage group (traffic lookup)
STAGE_KERNEL = {
    "predict_hash": "k_predict_hash",
    "neighbors_lambda": "k_neighbors_lambda",
    "lambda": "k_lambda",
    "delta_apply": "k_delta_apply",
    "curl": "k_curl",
    "vort_apply": "k_vort_apply_nohash",  # the stage replayed alone
    "vort_apply_hash": "k_vort_apply",  # a step's vorticity pass, with the next step's hash
    # the same passes in the fast-arithmetic mode
    "lambda_fast": "k_lambda_fast",
    "delta_apply_fast": "k_delta_apply_fast",
    "curl_fast": "k_curl_fast",
    "vort_apply_hash_fast": "k_vort_apply_fast",
}


def measured_traffic(workload: str, stage: str):
    """HBM bytes per launch of the stage's kernel from profiles/traffic_<workload>.md (lines
    'traffic_bytes_per_launch: <kernel> <bytes>'), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
    kern = STAGE_KERNEL.get(stage)
    if kern is None or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch:" and parts[1] == kern:
                return float(parts[2])
    return None


def measured_valu(workload: str, stage: str):
    """wave64 VALU instructions per launch of the stage's kernel from the counter summary
    profiles/valu_dam_break_262k_1m.md (lines 'valu_instr_per_launch: <workload> <kernel> <n>'),
    or None."""
    path = os.path.join(ROOT, "profiles", "valu_dam_break_262k_1m.md")
    kern = STAGE_KERNEL.get(stage)
    if kern is None or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 4 and parts[0] == "valu_instr_per_launch:" and parts[1] == workload and parts[2] == kern:
                return float(parts[3])
    return None


def progress(msg: str) -> None:
    """One progress line on stderr (stdout carries only the JSON result line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def learned_leg(n_side, box, train_seeds, eval_seeds, rec_steps, stride, warm, gap):
    progress(f"learned leg {n_side}: recording {len(train_seeds)} seeds x {rec_steps} steps")
    ds, rows = learned.record_dataset(train_seeds, rec_steps, n_side, box, stride=stride, log=lambda *a: None)
    mean, inv, ym, ys, nvalid = learned.normalise(ds, rows)
    progress(f"learned leg {n_side}: training on {nvalid} samples")
    m, losses = learned.train(ds, nvalid, mean, inv, ym, ys, steps=3000, log=lambda *a: None)
    ds.close()
    progress(f"learned leg {n_side}: evaluating on seeds {list(eval_seeds)}")
    _, summ = learned.evaluate(m, list(eval_seeds), warm, nstates=4, gap=gap, n_side=n_side, box=box, time_steps=200,
                               log=lambda *a: None)
    summ["samples"] = nvalid
    summ["loss"] = [float(losses[:20].mean()), float(losses[-100:].mean())]
    summ["train_seeds"] = list(train_seeds)
    summ["eval_seeds"] = list(eval_seeds)
    # the stated target: wall time to 0.5% mean compression, per state, by the adaptive step
    t = summ["targets"]["abs_0.005"]
    summ["time_to_0.005_ms"] = {
        "plain": t.get("ms_plain_mean"), "learned": t.get("ms_learned_mean"),
        "plain_reached_fraction": t["plain_reached"], "learned_reached_fraction": t["learned_reached"],
        "max_passes": t["max_passes"],
        "note": "null: not reached within max_passes passes on the evaluated states",
    }
    # the network's inference against the BF16 MFMA peak
    sc = scenes.dam_break(n_side, box, seed=eval_seeds[0])
    ms = learned.mlp_infer_ms(m, sc, warm)
    flop = learned.MLP_FLOP_PER_SAMPLE * sc.n
    summ["mlp_roofline"] = {"bound": "mfma", "kernel": "k_mlp_infer", "achieved": flop / (ms * 1e-3) / 1e12,
                            "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": flop / (ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, "flop_per_launch": flop,
                            "avg_launch_ms": ms}
    m.close()
    return summ


def reference_app_leg(warmup: int, steps: int, frames: int = 100):
    """The workload behind the reference's one published number (12-15 ms per frame for 262,144
    particles on an RTX 3090, its README): the app's scene 4 (src/renderapp.cpp:137-234, the
    default scene of src/main.cpp) with the app's 13 parameters (h = 1, rest density 6,
    4 iterations, dt = 0.016, renderapp.cpp:116-129), gravity -49 and the fixed 64^3 grid.
    Graph-replayed steps as in the headline, and the app's own frame loop: one step per frame
    with the host clock around the step and its synchronisation (renderapp.cpp:497-502)."""
    sc = driver.reference_scene(4)
    s = Solver.from_scene(sc, flags=FLAG_GRAPH)
    s.advance(warmup)
    s.synchronize()
    ms = s.time_steps(steps) / max(steps, 1)
    t = []
    for _ in range(frames):
        t0 = time.perf_counter()
        s.step()
        s.synchronize()
        t.append(time.perf_counter() - t0)
    s.close()
    frame_ms = 1000.0 * float(np.median(t))
    return {"scene": "reference app scene 4 (renderapp.cpp:137-234), 64^3 lattice at h = 1, 64^3 grid",
            "particles": sc.n, "parameters": {k: (list(v) if isinstance(v, tuple) else v) for k, v in sc.params.items()},
            "ms_per_step": ms, "value": sc.n / (ms * 1e-3), "unit": "particle-steps/s",
            "frame_ms_host_clock_median": frame_ms, "frames": frames,
            "published_ms_per_frame": [12.0, 15.0], "published_hardware": "GeForce RTX 3090 (reference README)",
            "speedup_vs_published_12ms": 12.0 / ms, "frame_speedup_vs_published_12ms": 12.0 / frame_ms}


def copy_bandwidth(mb: int = 512, reps: int = 5) -> float:
    """GB/s (read + write) of a plain streaming copy kernel: the counter-calibration reference and
    the achievable-bandwidth context for the roofline."""
    import ctypes as C

    nbytes = mb << 20
    a = DeviceBuffer(nbytes)
    b = DeviceBuffer(nbytes)
    best = None
    ms = np.zeros(1, np.float32)
    for _ in range(reps):
        rc = lib().pbf_copy_kernel(b.ptr, a.ptr, nbytes, ms.ctypes.data_as(C.POINTER(C.c_float)))
        if rc != 0:
            raise RuntimeError(lib().pbf_last_error().decode())
        best = ms[0] if best is None else min(best, ms[0])
    a.close()
    b.close()
    return float(2.0 * nbytes / (float(best) * 1e-3) / 1e9)


def reference_cpu_step(scene, budget_s: float = 8.0):
    """The reference's own CPU step, PBFSolver::solve (src/solver.cpp:60-157), built from its
    sources into oracle/_ref/libpbfref.so, on the same scene: a bounded sample, or None when the
    library is absent or the scene does not fit its fixed 64^3 grid.  Each call of the driver
    constructs a solver, so the per-step time is (T(k steps) - T(1 step)) / (k - 1).  Its
    physics is the CPU variant's (gravity -9.8, its own vorticity and jitter), and it prints
    every particle's density in every iteration (sent to /dev/null here): timed as written."""
    import ctypes as C

    from oracle_lib import orc, ref, ref_available  # noqa: E402

    h = float(scene.params["h"])
    if not ref_available() or any(d > 64 for d in scene.grid_dims):
        return None
    L = ref()
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.ref_run.restype = C.c_int
    L.ref_run.argtypes = [fp, fp, fp, fp, C.c_int, C.c_int]
    p = scene.params
    params = np.array([h, p["restRho"], p["epsilon"], p["k"], p["n"], p["deltaq"], p["radius"], p["c"], p["dt"],
                       p["iterations"], p["vorticityConfinementEpilon"]], np.float32)
    box = np.ascontiguousarray(scene.box, np.float32)

    def run(k):
        pos = np.ascontiguousarray(scene.pos, np.float32).copy()
        vel = np.ascontiguousarray(scene.vel, np.float32).copy()
        t0 = time.perf_counter()
        L.ref_run(params, box, pos, vel, scene.n, k)
        return time.perf_counter() - t0

    progress(f"cpu baseline: the reference's CPU solver on {scene.name}")
    t1 = run(1)
    k = int(max(3, min(20, budget_s / max(t1, 1e-3))))
    tk = run(k)
    per = max(tk - t1, 1e-9) / (k - 1)
    return {
        "value": scene.n / per,
        "unit": "particle-steps/s",
        "cores": int(orc().orc_max_threads()),
        "kind": "reference",
        "sample": f"{k} steps (and 1 for the construction) of {scene.name} ({scene.n} particles), {t1 + tk:.1f} s",
        "ms_per_step": 1000.0 * per,
        "note": "src/solver.cpp PBFSolver::solve as written: its CPU physics variant, per-particle printing",
    }


def cpu_baseline(scene, budget_s: float = 12.0):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, orc  # noqa: E402

    threads = orc().orc_max_threads()
    progress(f"cpu baseline: the oracle on {scene.name}, {threads} threads")
    o = Oracle.from_scene(scene)
    o.step(1)  # warm-up
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    psps = scene.n * steps / el
    return {
        "value": psps,
        "unit": "particle-steps/s",
        "cores": int(threads),
        "kind": "port",
        "sample": f"{steps} steps of {scene.name} ({scene.n} particles) after 1 warm-up step, {el:.1f} s",
        "ms_per_step": 1000.0 * el / steps,
        "reference_solver": reference_cpu_step(scene),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="dam_break_262k", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--copy-mb", type=int, default=512, help="size of the calibration copy kernel (0: skip)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start leg")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--no-app", action="store_true", help="skip the reference app's scene-4 leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py runs on one GPU only: the solver path shards by spatial slabs in design, "
            f"but no multi-GPU run is built; got --gpus {args.gpus}",
            file=sys.stderr,
        )
        sys.exit(2)

    copy_gbs = copy_bandwidth(args.copy_mb) if args.copy_mb > 0 else None
    flags = 0 if args.no_graph else FLAG_GRAPH
    batch = None
    if args.workload == "batch_24_scenes":
        batch = WORKLOADS[args.workload]()
        s = Solver.batch(batch, flags=flags)
        scene = max(batch, key=lambda sc: sc.n)  # the CPU baseline's sample
        n, ncells = s.n, s.ncells
        name = f"batch of {len(batch)} dam breaks ({n} particles, {min(sc.n for sc in batch)}-{max(sc.n for sc in batch)} per scene)"
    else:
        scene = WORKLOADS[args.workload]()
        s = Solver.from_scene(scene, flags=flags)
        n = scene.n
        ncells = int(np.prod(scene.grid_dims))
        name = scene.name
    iters = int(scene.params["iterations"])

    if args.warmup > 0:
        s.advance(args.warmup)
        s.synchronize()
    tuning, tune_ms = s.get_tuning(), None
    if not args.no_autotune:
        tuning, tune_ms = s.autotune(10)
    row_order = ("z-y", "morton-yz")[s.get_ordering()]
    ms_total = s.time_steps(args.steps)
    ms_step = ms_total / max(args.steps, 1)
    progress(f"{name}: {args.steps} timed steps, {ms_step:.4f} ms per step")
    value = n / (ms_step * 1e-3)

    # per-kernel-group times of eagerly launched steps (HIP events on the solver stream)
    acc = {}
    cnt = {}
    for _ in range(max(1, args.profile_steps)):
        for stage, ms in s.profile_step():
            acc[stage] = acc.get(stage, 0.0) + ms
            cnt[stage] = cnt.get(stage, 0) + 1
    launches_per_step = {k: cnt[k] / max(1, args.profile_steps) for k in cnt}
    avg_ms = {k: acc[k] / cnt[k] for k in acc}  # per launch
    per_step_ms = {k: avg_ms[k] * launches_per_step[k] for k in avg_ms}
    dom = max(per_step_ms, key=per_step_ms.get)
    dom_bytes = algorithmic_bytes(dom, n, ncells)
    # launch duration: back-to-back replays between two events (no per-launch event overhead)
    # where the stage can be replayed, else the per-launch events of the profile steps
    launch_ms = s.time_stage(dom, 50) if dom in Solver.STAGES else avg_ms[dom]
    achieved = dom_bytes / (launch_ms * 1e-3) / 1e9
    sort_stage = ["predict_hash", "scan", "scatter_rank_reorder"]
    sort_ms_events = sum(per_step_ms.get(k, 0.0) for k in sort_stage)
    # The step's sort stage: the previous step's vorticity pass hashes and counts for it, then the
    # scan, scatter and reorder; timed as back-to-back replays of (vorticity pass with the hash,
    # scan, scatter, reorder) less those of the vorticity pass alone.
    vort_hash_sort_ms = s.time_stage("vort_hash_sort", 50)
    vort_ms = s.time_stage("vort_apply", 50)
    sort_ms = vort_hash_sort_ms - vort_ms
    # a step after a state change hashes for itself: predict + hash, scan, scatter, reorder
    sort_unfused_ms = s.time_stage("sort_reorder", 50)
    sort_bytes = float(n * 96 + 12 * ncells)
    # scan (2 launches up to 2048 tiles of 1024 cell counters, else 3), scatter, reorder; the hash
    # rides in the previous step's vorticity pass
    sort_launches = 2 + (2 if ncells + 1 <= 2048 * 1024 else 3)
    floor_ms = s.time_stage("launch_floor", 50)  # an empty kernel of the same grid, replayed
    # the constraint passes are bound by VALU issue: the dominant kernel's instruction rate
    # against the nominal issue peak and against the probe's rate measured here
    valu_instr = measured_valu(args.workload, dom)
    probe_rate = Solver.VALU_PROBE_INSTR / (s.time_stage("valu_probe", 5) * 1e-3)
    valu = None
    if valu_instr is not None:
        rate = valu_instr / (launch_ms * 1e-3)
        valu = {"instr_per_launch": valu_instr, "achieved": rate, "peak": VALU_PEAK, "unit": "wave64 VALU instr/s",
                "frac": rate / VALU_PEAK, "probe_rate": probe_rate, "frac_of_probe": rate / probe_rate}

    pos, vel = s.get_state()
    finite = bool(np.isfinite(pos).all() and np.isfinite(vel).all())
    s.close()

    # the fast-arithmetic mode (hardware rsqrt in the neighbour terms): same
    # workload, warm-up and launch shape
    progress("fast-arithmetic leg")
    if batch is not None:
        f = Solver.batch(batch, flags=flags | FLAG_FAST_MATH)
    else:
        f = Solver.from_scene(scene, flags=flags | FLAG_FAST_MATH)
    f.set_tuning(*tuning)
    if args.warmup > 0:
        f.advance(args.warmup)
    f.synchronize()
    fast_ms = f.time_steps(args.steps) / max(args.steps, 1)
    fast = {"ms_per_step": fast_ms, "value": n / (fast_ms * 1e-3), "speedup": ms_step / fast_ms,
            "delta_apply_launch_ms": f.time_stage("delta_apply", 50), "lambda_launch_ms": f.time_stage("lambda", 50),
            "parity": "tolerances: one step within 1e-5 of the oracle's positions for particles 2 h or more from the "
                      "walls, within (iterations + 1) x the clamp jitter for all; over 1000 steps the run means of "
                      "the mean density error and the compression within 1% of the oracle's, of the maximum within 5%"}
    # wave64 VALU instructions per launch, both modes, from the committed counter summary
    fast["valu_instr_per_launch"] = {
        st: {"exact": measured_valu(args.workload, st), "fast": measured_valu(args.workload, st + "_fast")}
        for st in ("lambda", "delta_apply", "curl", "vort_apply_hash")
    }
    f.close()
    seq = None
    if batch is not None:
        seq_ms = sequential_ms(batch, args.warmup, max(20, args.steps // 5))
        seq = {"ms_per_step_sum_of_scenes_alone": seq_ms, "batched_speedup": seq_ms / ms_step}

    learned_res = None
    if not args.no_learned and args.workload == "dam_break_262k":
        learned_res = {
            "dam_break_262k": learned_leg((64, 64, 64), (0, 0, 0, 6.4, 4.0, 3.2), range(100, 108), (200, 201, 202, 203),
                                          300, 64, 60, 12),
            "dam_break_32k_box_3.2x2.0x1.6": learned_leg((32, 32, 32), (0, 0, 0, 3.2, 2.0, 1.6), range(100, 104),
                                                         (200, 201), 150, 8, 40, 8),
        }

    app = None
    if args.workload == "dam_break_262k" and not args.no_app:
        progress("the reference app's scene 4")
        app = reference_app_leg(args.warmup, args.steps)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / PUBLISHED_PSPS,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 dam-break lattice)",
        "config": {
            "workload": name,
            "particles": n,
            "grid": list(scene.grid_dims) if batch is None else "one per scene",
            "box": [float(x) for x in scene.box] if batch is None else "one per scene (25% fill)",
            "iterations": iters,
            "dt": scene.params["dt"],
            "physical_parameters": {k: v for k, v in scenes.PHYSICAL.items() if k not in ("iterations", "dt")},
            "solver_parameters": {k: (list(v) if isinstance(v, tuple) else v) for k, v in scene.params.items()},
            "parameter_mapping": "restRho = rho0 / m; c and vorticity epsilon times m / rho0; k times (h / 1 m)^2; "
                                 "spiky cut-off and clamp jitter 0.01 h (pbf_amd.scenes.reference_units)",
            "graph": not args.no_graph,
            "parallelism": "single GPU",
            "tuning": {"block": tuning[0], "grid": tuning[1], "xcd_map": tuning[2], "row_order": row_order,
                       "autotuned": not args.no_autotune},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.workload, dom),
            "algorithmic_bytes_per_launch": dom_bytes,
            "avg_launch_ms": launch_ms,
            "avg_launch_ms_events_per_launch": avg_ms[dom],
            "issue_bound": "valu",
            "valu": valu,
        },
        "stages_ms": {k: round(v, 5) for k, v in per_step_ms.items()},
        "step_roofline_frac": step_bytes(n, ncells, iters) / (ms_step * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "sort_reorder": {"ms": sort_ms, "frac": sort_bytes / (sort_ms * 1e-3) / (HBM_PEAK_GBS * 1e9) if sort_ms else None,
                         "algorithmic_bytes": sort_bytes, "launches": sort_launches,
                         "hash": "fused into the previous step's vorticity pass",
                         "ms_vort_hash_sort": vort_hash_sort_ms, "ms_vort_alone": vort_ms,
                         "unfused_ms": sort_unfused_ms, "unfused_launches": sort_launches + 1,
                         "unfused_frac": sort_bytes / (sort_unfused_ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
                         "ms_events_eager": sort_ms_events,
                         "launch_floor_ms": floor_ms,
                         "launch_floor_share": sort_launches * floor_ms / sort_ms if sort_ms > 0 else None},
        "fast_math": fast,
        "copy_kernel_gbs": copy_gbs,
        "state_finite": finite,
        "cpu_baseline": cpu,
        "learned": learned_res,
        "reference_app": app,
        "batch_vs_sequential": seq,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
