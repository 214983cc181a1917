"""Seeded initial conditions and parameter sets for the PBF workloads.

The reference builds its scenes in its app (src/renderapp.cpp:115-234): a particle lattice
plus ``rand()`` jitter seeded from the clock, ``gridSpace = box`` and ``boundary = box`` shrunk
by the particle radius (renderapp.cpp:130-131, 251).  Here the jitter comes from SplitMix64
(Steele, Lea & Flood 2014) so that every scene is reproducible bit for bit.

Parameters follow the dam-break workloads: h = 0.1 m, rest density 1000 kg/m^3, particle mass
0.125 kg, CFM epsilon 100, s_corr k = 0.1, n = 4, dq = 0.2 h, XSPH c = 0.01, vorticity
epsilon 5e-4, dt = 1/120 s, gravity -9.8 m/s^2.  The reference's formulas carry no particle
mass or SPH volume, so :func:`reference_units` maps these physical values onto its constants:

* ``restRho`` = rho0 / m = 8000: its density is sum W (kernels/cudaPBFSolver.cu:170);
* ``c`` and ``vorticityConfinementEpilon`` times the particle volume m / rho0: its XSPH and curl
  sums omit the m/rho factor of SPH (cu:212, cu:231);
* ``k`` times (h / 1 m)^2: s_corr is dimensionless but is added to lambda, which carries
  length^2 (cu:195); the reference's own scene runs at h = 1 where the two agree;
* the spiky-gradient cut-off and clamp jitter at 0.01 h: the reference hard-codes 0.01 for its
  h = 1 (cuh:107, solverBase.h:5).

Taken literally (:func:`literal_params`), the physical values make both the reference's CPU
solver and the oracle diverge within a few steps; the literal set is kept for parity tests.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> np.ndarray:
    """``count`` successive SplitMix64 outputs for ``seed`` (uint64)."""
    idx = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK64) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform24(seed: int, count: int) -> np.ndarray:
    """``count`` uniform floats in [0, 1) with 24 random bits each (float64 holding exact values)."""
    return (splitmix64(seed, count) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


@dataclass
class Scene:
    """Initial state plus the parameters and boxes the solver needs."""

    name: str
    pos: np.ndarray  # (N, 3) float32
    vel: np.ndarray  # (N, 3) float32
    box: np.ndarray  # (6,) float32  pmin xyz, pmax xyz  (the reference's gridSpace)
    boundary: np.ndarray  # (6,) float32 box shrunk by the radius
    grid_dims: tuple
    params: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])


PHYSICAL = dict(
    h=0.1,
    rho0=1000.0,
    mass=0.125,
    epsilon=100.0,
    k=0.1,
    n=4.0,
    deltaq=0.2 * 0.1,
    c=0.01,
    dt=1.0 / 120.0,
    iterations=4,
    vorticity=5e-4,
    radius=0.03,
    gravity=(0.0, -9.8, 0.0),
    max_neighbors=63,
)


def reference_units(phys: dict) -> dict:
    """Solver parameters (reference formula constants) for physical parameters ``phys``."""
    h = phys["h"]
    vol = phys["mass"] / phys["rho0"]
    return dict(
        h=h,
        restRho=phys["rho0"] / phys["mass"],
        epsilon=phys["epsilon"],
        k=phys["k"] * h * h,
        n=phys["n"],
        deltaq=phys["deltaq"],
        c=phys["c"] * vol,
        dt=phys["dt"],
        iterations=phys["iterations"],
        vorticityConfinementEpilon=phys["vorticity"] * vol,
        radius=phys["radius"],
        gravity=tuple(phys["gravity"]),
        jitter_coeff=0.01 * h,
        grad_rmin=0.01 * h,
        max_neighbors=phys["max_neighbors"],
    )


def literal_params(**over) -> dict:
    """The physical values passed to the reference formula unchanged (diverges; parity tests only)."""
    ph = dict(PHYSICAL)
    p = dict(
        h=ph["h"],
        restRho=ph["rho0"] / ph["mass"],
        epsilon=ph["epsilon"],
        k=ph["k"],
        n=ph["n"],
        deltaq=ph["deltaq"],
        c=ph["c"],
        dt=ph["dt"],
        iterations=ph["iterations"],
        vorticityConfinementEpilon=ph["vorticity"],
        radius=ph["radius"],
        gravity=ph["gravity"],
        jitter_coeff=0.01,
        grad_rmin=0.01,
        max_neighbors=ph["max_neighbors"],
    )
    p.update(over)
    return p


def default_params(**over) -> dict:
    """The dam-break parameter set in reference units (all workloads share it)."""
    phys = dict(PHYSICAL)
    for key in list(over):
        if key in ("iterations", "dt", "gravity", "max_neighbors") and key in phys:
            phys[key] = over[key]
    p = reference_units(phys)
    p.update(over)
    return p


def lattice(dims, spacing, origin, seed, jitter=0.005, seed_offset=0) -> np.ndarray:
    """Lattice of dims = (nx, ny, nz) points, x fastest, with U(-jitter, jitter) per coordinate."""
    nx, ny, nz = dims
    n = nx * ny * nz
    i = np.arange(n)
    ix, iy, iz = i % nx, (i // nx) % ny, i // (nx * ny)
    u = uniform24(seed + seed_offset, 3 * n).reshape(n, 3)
    base = np.stack([ix, iy, iz], axis=1).astype(np.float64) * spacing + np.asarray(origin, np.float64)
    return (base + (2.0 * u - 1.0) * jitter).astype(np.float32)


def _make(name, pos, box, params, vel=None) -> Scene:
    box = np.asarray(box, np.float32)
    r = np.float32(params["radius"])
    boundary = np.concatenate([box[:3] + r, box[3:] - r]).astype(np.float32)
    h = params["h"]
    dims = tuple(int(round((float(box[3 + a]) - float(box[a])) / h)) for a in range(3))
    if vel is None:
        vel = np.zeros_like(pos)
    return Scene(name, pos, vel.astype(np.float32), box, boundary, dims, dict(params))


def dam_break(n_side=(32, 32, 32), box=(0, 0, 0, 1.6, 1.6, 1.6), seed=0, spacing=0.05, **over) -> Scene:
    """A block of n_side lattice points at ``spacing`` in the low corner of ``box``."""
    p = default_params(**over)
    origin = [box[a] + 0.5 * spacing for a in range(3)]
    pos = lattice(n_side, spacing, origin, seed)
    return _make(f"dam_break_{n_side[0]}x{n_side[1]}x{n_side[2]}", pos, box, p)


def config_cpu_32k(seed=0) -> Scene:
    """32768 particles, 32^3 lattice in a 1.6 m box (16^3 grid)."""
    sc = dam_break((32, 32, 32), (0, 0, 0, 1.6, 1.6, 1.6), seed)
    sc.name = "dam break 32768 particles: 32^3 lattice at 0.05 m in a 1.6 m cube, 16^3 cells of h = 0.1 m"
    return sc


def config_262k(seed=1) -> Scene:
    """262144 particles, 64^3 lattice in a 6.4 x 4.0 x 3.2 m box (64 x 40 x 32 grid)."""
    sc = dam_break((64, 64, 64), (0, 0, 0, 6.4, 4.0, 3.2), seed)
    sc.name = "dam break 262144 particles: 64^3 lattice at 0.05 m in a 6.4 x 4.0 x 3.2 m box, 64 x 40 x 32 cells of h = 0.1 m"
    return sc


def config_1m(seed=2) -> Scene:
    """1048576 particles in two 64x128x64 blocks at the two x-ends of a 12.8 x 6.4 x 6.4 m box."""
    p = default_params(iterations=8)
    box = (0.0, 0.0, 0.0, 12.8, 6.4, 6.4)
    sp = 0.05
    a = lattice((64, 128, 64), sp, (0.5 * sp, 0.5 * sp, 0.5 * sp), seed)
    x0 = 12.8 - 64 * sp + 0.5 * sp
    b = lattice((64, 128, 64), sp, (x0, 0.5 * sp, 0.5 * sp), seed, seed_offset=1)
    return _make("double dam break 1048576 particles: two 64 x 128 x 64 blocks at the x-ends of a 12.8 x 6.4 x 6.4 m box, "
                 "128 x 64 x 64 cells of h = 0.1 m, 8 iterations", np.concatenate([a, b]), box, p)


def batch_sizes(count=24, lo=8192, hi=262144, seed=3) -> list:
    """Log-uniform particle counts in [lo, hi] for ``count`` scenes."""
    u = uniform24(seed, count)
    return [int(round(math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * x))) for x in u]


def scaled_dam_break(n: int, seed: int, fill=0.25, spacing=0.05, **over) -> Scene:
    """A dam break of ``n`` particles in a box whose volume is n*spacing^3/fill.

    The block is a near-cube lattice (x-fastest) with exactly ``n`` points; the box is a cube of
    side rounded up to whole cells of h.
    """
    p = default_params(**over)
    side = max(1, int(round(n ** (1.0 / 3.0))))
    nx = side
    nz = side
    ny = -(-n // (nx * nz))
    full = lattice((nx, ny, nz), spacing, (0.5 * spacing,) * 3, seed)[:n]
    vol = n * spacing ** 3 / fill
    h = p["h"]
    L = math.ceil((vol ** (1.0 / 3.0)) / h) * h
    ext = [max(L, (d + 1) * spacing) for d in (nx, ny, nz)]
    ext = [math.ceil(e / h - 1e-9) * h for e in ext]
    box = (0.0, 0.0, 0.0, ext[0], ext[1], ext[2])
    return _make(f"scaled_dam_break_{n}", full, box, p)


def batch_config(seed=3, count=24, lo=8192, hi=262144) -> list:
    """``count`` dam breaks with log-uniform particle counts in [lo, hi], each in its own box at
    25% fill (seeds seed*1000 + k)."""
    return [scaled_dam_break(nk, seed * 1000 + k, fill=0.25) for k, nk in enumerate(batch_sizes(count, lo, hi, seed))]


CONFIGS = {
    "cpu_32k": config_cpu_32k,
    "dam_break_262k": config_262k,
    "double_dam_break_1m": config_1m,
}
