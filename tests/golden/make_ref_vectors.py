"""Writes golden vectors from the reference's own CPU code (oracle/_ref/libpbfref.so, built by
oracle/Makefile from src/solver.cpp, src/AABB.cpp and test/testPrefixSum.cpp where the reference
tree lies).  Run from the repository root:

    make -C oracle && python tests/golden/make_ref_vectors.py

Outputs (data only):
  ref_probe_3k.npz  one PBFSolver::solve with dt = 0 and one iteration on a seeded 3375-particle
                    block (h = 0.1, 64^3 grid anchored at the box minimum): cell id per particle,
                    cell start/end, sorted cell members, neighbour lists, lambda and delta-p.
  ref_scan.npz      the reference scan test's serial and Blelloch exclusive scans
                    (test/testPrefixSum.cpp:3-53) of seeded counts in [0, 16).
  ref_step_probe.npz  one whole PBFSolver::solve (3 iterations, dt = 1/120, XSPH c = 1e-3) on a
                    seeded 1000-particle block with random velocities, all inside the boundary:
                    positions after the constraint loop, the loop's velocity (x - x_old) * invDt,
                    curl, XSPH viscosity and the last iteration's lambda.

    python tests/golden/make_ref_vectors.py [probe] [scan] [step]   (default: all)
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle_lib import fp, ip, ref, ref_step_probe, up  # noqa: E402
from pbf_amd import scenes  # noqa: E402

PROBE_PARAMS = dict(h=0.1, restRho=8000.0, epsilon=100.0, k=1e-3, n=4.0, deltaq=0.02, radius=0.03)


def probe_scene():
    """15^3 block at spacing 0.05 (jitter 0.005) in the middle of a 6.4 m box."""
    pos = scenes.lattice((15, 15, 15), 0.05, (1.0, 1.0, 1.0), seed=123)
    box = np.array([0, 0, 0, 6.4, 6.4, 6.4], np.float32)
    return pos, box


STEP_PARAMS = dict(h=0.1, restRho=8000.0, epsilon=100.0, k=1e-3, n=4.0, deltaq=0.02, radius=0.03, c=1e-3,
                   dt=float(np.float32(1.0 / 120.0)), iterations=3, vorticityConfinementEpilon=0.0)


def step_probe_scene():
    """10^3 block at spacing 0.05 (jitter 0.005) with velocities U(-0.3, 0.3) in a 6.4 m box."""
    pos = scenes.lattice((10, 10, 10), 0.05, (1.0, 1.0, 1.0), seed=321)
    vel = (scenes.uniform24(654, pos.size).reshape(pos.shape) * 0.6 - 0.3).astype(np.float32)
    box = np.array([0, 0, 0, 6.4, 6.4, 6.4], np.float32)
    return pos, vel, box


def write_step_probe():
    pos, vel, box = step_probe_scene()
    out = ref_step_probe(STEP_PARAMS, box, pos, vel)
    prm = np.array([STEP_PARAMS[k] for k in ("h", "restRho", "epsilon", "k", "n", "deltaq", "radius", "c", "dt",
                                             "iterations", "vorticityConfinementEpilon")], np.float32)
    np.savez_compressed(os.path.join(HERE, "ref_step_probe.npz"), pos=pos, vel=vel, box=box, params=prm,
                        pos_out=out["pos"], vel_loop=out["vel_loop"], curl=out["curl"], visc=out["visc"],
                        lam=out["lambda"])


def main():
    which = set(sys.argv[1:]) or {"probe", "scan", "step"}
    if "step" in which:
        write_step_probe()
        print("wrote ref_step_probe.npz")
    if not which & {"probe", "scan"}:
        return
    R = ref()
    pos, box = probe_scene()
    n = pos.shape[0]
    cells = R.ref_grid_cells()
    p = PROBE_PARAMS
    prm = np.array([p["h"], p["restRho"], p["epsilon"], p["k"], p["n"], p["deltaq"], p["radius"]], np.float32)
    gid = np.empty(n, np.int32)
    cs = np.empty(cells, np.int32)
    ce = np.empty(cells, np.int32)
    srt = np.empty(n, np.int32)
    nbr = np.empty((n, 64), np.int32)
    lam = np.empty(n, np.float32)
    off = np.empty((n, 3), np.float32)
    nin = R.ref_probe(fp(prm), fp(box), fp(pos), n, ip(gid), ip(cs), ip(ce), ip(srt), ip(nbr), fp(lam), fp(off))
    # entries past the count (slot 63) are uninitialised in the reference (solver.h:105-108)
    count = nbr[:, 63].copy()
    nbr[np.arange(64)[None, :] >= count[:, None]] = -1
    nbr[:, 63] = count  # slot 63 holds the count
    used = np.unique(gid[gid >= 0])
    np.savez_compressed(
        os.path.join(HERE, "ref_probe_3k.npz"),
        pos=pos,
        box=box,
        params=prm,
        n_in_grid=np.int32(nin),
        gid=gid,
        cells=used.astype(np.int32),
        cell_start=cs[used],
        cell_end=ce[used],
        sorted_ids=srt,
        nbr=nbr,
        lam=lam,
        offset=off,
    )
    rng = np.random.default_rng(2024)
    data = rng.integers(0, 16, size=4096).astype(np.uint32)
    a = np.empty_like(data)
    b = np.empty_like(data)
    R.ref_scan_serial(up(data), up(a), data.size)
    R.ref_scan_blelloch(up(data), up(b), data.size)
    np.savez_compressed(os.path.join(HERE, "ref_scan.npz"), data=data, serial=a, blelloch=b)
    print("wrote ref_probe_3k.npz, ref_scan.npz")


if __name__ == "__main__":
    main()
