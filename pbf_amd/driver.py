"""Headless driver: the reference app's scenes and frame loop without a window.

    python -m pbf_amd.driver --scene 4 --frames 200 [--save out.pbfs] [--load in.pbfs]

Scenes 1-5 follow the reference's initial conditions (src/renderapp.cpp:137-234): a 64^3
particle lattice (nParticles = 64^3, renderapp.cpp:115-117) filling an initial region with
spacing region/64 plus a jitter of 0.2 per coordinate, in the box [-30,0,-30]..[30,40,30]
(renderapp.cpp:130) with its boundary shrunk by the radius (renderapp.cpp:131), and the app's
parameters (renderapp.cpp:116-129: h = 1, rest density 6, epsilon 1, k = 0.01, dq = 0.1 h,
n = 4, c = 0.1, dt = 0.016, 4 iterations, vorticity 0.1, radius 0.3).  The jitter comes from
SplitMix64 instead of rand() seeded from the clock (renderapp.cpp:136).  Gravity is the GPU
kernel's (0, -49, 0) (kernels/cudaPBFSolver.cu:9), the grid the reference's 64^3 anchored at the
box minimum.  Each frame prints the wall time of the step, as renderapp.cpp:498-502 does.
"""
from __future__ import annotations

import argparse
import time

import numpy as np

from . import Solver, statefile
from .scenes import Scene, uniform24

APP_PARAMS = dict(
    h=1.0,
    restRho=6.0,
    epsilon=1.0,
    k=0.01,
    n=4.0,
    deltaq=0.1,
    c=0.1,
    dt=0.016,
    iterations=4,
    vorticityConfinementEpilon=0.1,
    radius=0.3,
    gravity=(0.0, -49.0, 0.0),
    jitter_coeff=0.01,
    grad_rmin=0.01,
    max_neighbors=63,
)

BOX = np.array([-30.0, 0.0, -30.0, 30.0, 40.0, 30.0], np.float32)


def _lattice(region, count, dims_xy, seed, offset=0):
    """Points i = 0..count-1 at region.min + (i % dx, i / dx % dy, i / (dx*dy)) * gap + 0.2 u."""
    pmin = np.asarray(region[:3], np.float64)
    gap = (np.asarray(region[3:], np.float64) - pmin) / 64.0
    i = np.arange(count)
    dx, dy = dims_xy
    d = np.stack([i % dx, (i // dx) % dy, i // (dx * dy)], 1).astype(np.float64)
    u = uniform24(seed + offset, 3 * count).reshape(count, 3)
    return (pmin + d * gap + 0.2 * u).astype(np.float32)


def reference_scene(index: int, seed: int = 0, n_side: int = 64) -> Scene:
    """Scene ``index`` (1-5) of the reference app (renderapp.cpp:137-234)."""
    n = n_side**3
    if index == 1:
        pos = _lattice((-16.0, 0.2, -16.0, 16.0, 32.2, 16.0), n, (n_side, n_side), seed)
        vel = np.tile(np.array([0.0, 20.0, 0.0], np.float32), (n, 1))
    elif index == 2:
        pos = _lattice((-28.0, 0.3, -4.0, 4.0, 32.3, 28.0), n, (n_side, n_side), seed)
        vel = np.tile(np.array([0.0, 0.0, -20.0], np.float32), (n, 1))
    elif index == 3:
        half = n // 2
        a = _lattice((-29.7, 0.2, -29.7, -2.0, 32.2, -2.0), half, (n_side, n_side // 2), seed)
        # the reference keeps the first region's spacing for the second block (renderapp.cpp:186-197)
        b = _lattice((2.0, 0.2, 2.0, 29.7, 32.2, 29.7), half, (n_side, n_side // 2), seed, offset=1)
        pos = np.concatenate([a, b])
        vel = np.concatenate([np.tile(np.array([50.0, 0.0, 50.0], np.float32), (half, 1)), np.tile(np.array([-50.0, 0.0, -50.0], np.float32), (half, 1))])
    elif index == 4:
        pos = _lattice((-24.0, 0.2, -16.0, 8.0, 32.2, 16.0), n, (n_side, n_side), seed)
        vel = np.tile(np.array([-10.0, 15.0, 0.0], np.float32), (n, 1))
    elif index == 5:
        pos = _lattice((-16.0, 0.2, -16.0, 16.0, 32.2, 16.0), n, (n_side, n_side), seed)
        t = np.stack([pos[:, 2], np.zeros(n, np.float32), -pos[:, 0]], 1).astype(np.float64)
        norm = np.linalg.norm(t, axis=1, keepdims=True)
        vel = (50.0 * t / np.where(norm > 0, norm, 1.0)).astype(np.float32)
    else:
        raise ValueError("scene index must be 1..5")
    r = np.float32(APP_PARAMS["radius"])
    boundary = np.concatenate([BOX[:3] + r, BOX[3:] - r]).astype(np.float32)
    return Scene(f"reference_scene_{index}", pos, vel.astype(np.float32), BOX.copy(), boundary, (64, 64, 64), dict(APP_PARAMS))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--scene", type=int, default=4, help="reference scene 1-5 (default 4, src/main.cpp:6)")
    ap.add_argument("--frames", type=int, default=100)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--save", default=None, help="write the final state to this file")
    ap.add_argument("--load", default=None, help="start from this state file instead of a scene")
    ap.add_argument("--quiet", action="store_true")
    args = ap.parse_args()
    if args.load:
        para, gs, bd, pos, vel = statefile.load(args.load)
        s = Solver(para, gs, bd, pos, vel)
    else:
        sc = reference_scene(args.scene, args.seed)
        s = Solver.from_scene(sc)
        gs, bd = sc.box, sc.boundary
    total = 0.0
    for f in range(args.frames):
        t0 = time.perf_counter()
        s.step()
        s.synchronize()
        dt = time.perf_counter() - t0
        total += dt
        if not args.quiet:
            print(f"{dt:.6f}", flush=True)
    print(f"{args.frames} frames, {s.n} particles, mean {1000.0 * total / max(args.frames, 1):.3f} ms/frame", flush=True)
    if args.save:
        pos, vel = s.get_state()
        statefile.save(args.save, s.para, gs, bd, pos, vel)


if __name__ == "__main__":
    main()
