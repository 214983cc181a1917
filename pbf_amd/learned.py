"""Learned warm start: dataset from the plain solver, on-device training, evaluation.

    python -m pbf_amd.learned [--quick] [--out results.json]

1. Record: for each training seed, run the dam break with plain steps of ``record_iters``
   Jacobi iterations; every step records, for a rotating 1/stride subset of particles, the 16
   raw features at the predicted positions and the multiplier the step accumulated over its
   iterations (divided by h^2).  Everything stays in device memory.
2. Normalise: per-feature mean / inverse std and target mean / std from the recorded data.
3. Train: Adam on random batches, BF16 MFMA forward / backward kernels, FP32 master weights.
4. Evaluate on held-out seeds: after a common warm-up, step with a residual tolerance (mean
   compression error max(C,0) <= tol) plain and learned; report iterations and wall time per
   step.  The plain iterations remain the fallback of every learned step.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import DeviceDataset, FLAG_DENSITY_STATS, Mlp, Solver, scenes


def record_dataset(seeds, steps, n_side=(64, 64, 64), box=(0, 0, 0, 6.4, 4.0, 3.2), record_iters=8, stride=64, log=print):
    per_step = None
    ds = None
    rows = 0
    t0 = time.perf_counter()
    for si, seed in enumerate(seeds):
        sc = scenes.dam_break(n_side, box, seed=seed)
        s = Solver.from_scene(sc)
        if per_step is None:
            per_step = -(-sc.n // stride)
            ds = DeviceDataset(per_step * steps * len(seeds))
        for k in range(steps):
            s.record_step(record_iters, ds, base=rows, stride=stride, phase=k % stride)
            rows += per_step
        s.synchronize()
        s.close()
        log(f"  recorded seed {seed}: {steps} steps ({time.perf_counter() - t0:.1f} s)")
    return ds, rows


def normalise(ds: DeviceDataset, rows: int):
    x, t = ds.host(rows)
    valid = np.isfinite(x).all(1) & np.isfinite(t)
    x, t = x[valid], t[valid]
    mean = x.mean(0)
    std = x.std(0)
    inv = np.where(std > 1e-12, 1.0 / np.maximum(std, 1e-12), 0.0).astype(np.float32)
    ym, ys = float(t.mean()), float(t.std() if t.std() > 0 else 1.0)
    tn = ((t - ym) / ys).astype(np.float32)
    ds.upload(x, tn)
    return mean.astype(np.float32), inv, ym, ys, int(valid.sum())


def train(ds, nsamples, mean, inv, ym, ys, steps=3000, batch=4096, lr=2e-3, seed=0, log=print):
    m = Mlp(seed=seed)
    m.set_normalization(mean, inv, ys, ym)
    t0 = time.perf_counter()
    losses = m.train(ds, nsamples, batch=batch, steps=steps, lr=lr, lr_final=lr * 0.02, seed=seed + 1)
    log(f"  trained {steps} steps in {time.perf_counter() - t0:.2f} s: loss {losses[:20].mean():.4f} -> {losses[-100:].mean():.4f}")
    return m, losses


def evaluate(m, seeds, warmup=60, steps=40, tol=0.005, max_iters=64, n_side=(64, 64, 64), box=(0, 0, 0, 6.4, 4.0, 3.2), log=print):
    res = {"plain": [], "learned": []}
    for seed in seeds:
        sc = scenes.dam_break(n_side, box, seed=seed)
        for mode in ("plain", "learned"):
            s = Solver.from_scene(sc, flags=FLAG_DENSITY_STATS)
            # common warm-up: plain fixed-iteration steps with the multiplier accumulated
            for _ in range(warmup):
                s.step_adaptive(False, 0.0, 4, 4)
            if mode == "learned":
                s.attach_mlp(m)
            s.synchronize()
            its, resid, kept, final = [], [], 0, []
            t0 = time.perf_counter()
            for _ in range(steps):
                it, r = s.step_adaptive(mode == "learned", tol, 1, max_iters)
                its.append(it)
                resid.append(r)
                if mode == "learned":
                    kept += int(s.last_guess_info()[2])
            s.synchronize()
            wall = time.perf_counter() - t0
            st = s.density_stats()
            res[mode].append(
                dict(
                    seed=seed,
                    mean_iters=float(np.mean(its)),
                    reached=float(np.mean(np.array(resid) <= tol)),
                    ms_per_step=1000.0 * wall / steps,
                    guess_kept=kept / steps if mode == "learned" else None,
                    final_mean_abs_err=st[0],
                    final_mean_compression=st[1],
                )
            )
            s.close()
            log(f"  seed {seed} {mode:7s}: {np.mean(its):5.1f} iters/step, reached tol {np.mean(np.array(resid) <= tol):.2f}, {1000.0 * wall / steps:.2f} ms/step")
    summ = {}
    for mode in res:
        summ[mode] = {k: float(np.mean([r[k] for r in res[mode]])) for k in ("mean_iters", "reached", "ms_per_step", "final_mean_compression")}
    summ["time_saving"] = 1.0 - summ["learned"]["ms_per_step"] / summ["plain"]["ms_per_step"]
    return res, summ


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="small scenes and short runs")
    ap.add_argument("--tol", type=float, default=0.005)
    ap.add_argument("--out", default=None)
    ap.add_argument("--train-steps", type=int, default=3000)
    args = ap.parse_args()
    if args.quick:
        n_side, box = (32, 32, 32), (0, 0, 0, 3.2, 2.0, 1.6)
        train_seeds, eval_seeds, rec_steps, warm, ev = range(100, 104), range(200, 202), 100, 30, 20
    else:
        n_side, box = (64, 64, 64), (0, 0, 0, 6.4, 4.0, 3.2)
        train_seeds, eval_seeds, rec_steps, warm, ev = range(100, 108), range(200, 204), 300, 60, 40
    print(f"record: {len(train_seeds)} seeds x {rec_steps} steps, {n_side} lattice", flush=True)
    ds, rows = record_dataset(train_seeds, rec_steps, n_side, box)
    mean, inv, ym, ys, nvalid = normalise(ds, rows)
    print(f"dataset: {nvalid} samples; target mean {ym:.4g} std {ys:.4g}", flush=True)
    m, losses = train(ds, nvalid, mean, inv, ym, ys, steps=args.train_steps)
    res, summ = evaluate(m, eval_seeds, warm, ev, args.tol, 64, n_side, box)
    out = dict(tol=args.tol, samples=nvalid, loss_start=float(losses[:20].mean()), loss_end=float(losses[-100:].mean()), summary=summ, runs=res)
    print(json.dumps(summ), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
