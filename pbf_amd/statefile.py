"""Flat binary solver state: checkpoint / resume and identical-state fixtures.

Layout (little endian):
    8 B   magic b"PBFSTATE"
    4 B   u32 version (1)
    4 B   u32 particle count n
    4 B   u32 sizeof(pbf_params)
    P B   pbf_params (include/pbf.h, as stored in memory)
    24 B  grid space  (pmin xyz, pmax xyz) f32
    24 B  boundary    (pmin xyz, pmax xyz) f32
    12n B positions, xyz f32 per particle (caller's particle order)
    12n B velocities, xyz f32 per particle

The reference keeps no state beyond positions and velocities between frames
(include/solverBase.h:45-47; lambda, neighbours and the grid are rebuilt every step), so this is
a complete checkpoint of a plain solve.
"""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np

from . import SolverParameter

MAGIC = b"PBFSTATE"
VERSION = 1


def save(path, para: SolverParameter, grid_space, boundary, pos, vel):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    raw = bytes(C.string_at(C.addressof(para), C.sizeof(para)))
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<III", VERSION, n, len(raw)))
        f.write(raw)
        f.write(np.ascontiguousarray(grid_space, "<f4").tobytes())
        f.write(np.ascontiguousarray(boundary, "<f4").tobytes())
        f.write(pos.astype("<f4").tobytes())
        f.write(vel.astype("<f4").tobytes())


def load(path):
    """Returns (params, grid_space[6], boundary[6], pos[n,3], vel[n,3])."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != MAGIC:
        raise ValueError("not a PBF state file")
    version, n, psize = struct.unpack_from("<III", data, 8)
    if version != VERSION:
        raise ValueError(f"unsupported state file version {version}")
    if psize != C.sizeof(SolverParameter):
        raise ValueError(f"parameter block of {psize} bytes, expected {C.sizeof(SolverParameter)}")
    off = 20
    para = SolverParameter.from_buffer_copy(data[off : off + psize])
    off += psize
    gs = np.frombuffer(data, "<f4", 6, off).copy()
    off += 24
    bd = np.frombuffer(data, "<f4", 6, off).copy()
    off += 24
    pos = np.frombuffer(data, "<f4", 3 * n, off).reshape(n, 3).copy()
    off += 12 * n
    vel = np.frombuffer(data, "<f4", 3 * n, off).reshape(n, 3).copy()
    off += 12 * n
    if off != len(data):
        raise ValueError("trailing bytes in state file")
    return para, gs, bd, pos, vel
