"""The reference-shaped C++ facade on the GPU: a caller program built against
include/pbf_solver.hpp steps an AoS Particle array in place with PBFSolver::solve
(include/solver.h:83-111 in the reference), one host round trip per step, and its state after
three steps equals the oracle's bit for bit.  The per-call upload does not change the result:
the slot order is the (cell, id) order of the positions, so a step depends on the state alone."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pbf_amd
from oracle_lib import Oracle
from pbf_amd import SolverParameter, scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CALLER = r"""
#include <cstdio>
#include <vector>
#include "pbf_solver.hpp"
using namespace pbf_ref;
int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    SolverParameter para;
    AABB box;
    int steps = 0;
    bool ok = std::fread(&para, sizeof para, 1, f) == 1 && std::fread(&box, sizeof box, 1, f) == 1 &&
              std::fread(&steps, sizeof steps, 1, f) == 1;
    std::vector<Particle> p(ok ? para.nParticles : 0);
    ok = ok && std::fread(p.data(), sizeof(Particle), p.size(), f) == p.size();
    std::fclose(f);
    if (!ok) return 4;
    try {
        PBFSolver solver(para, box);
        for (int k = 0; k < steps; ++k) solver.solve(p.data());
    } catch (const Error& e) {
        std::printf("error %d %s\n", e.code, e.what());
        return 5;
    }
    FILE* g = std::fopen(argv[2], "wb");
    if (!g) return 6;
    std::fwrite(p.data(), sizeof(Particle), p.size(), g);
    std::fclose(g);
    std::printf("stepped %d\n", steps);
    return 0;
}
"""


def test_facade_solve_matches_oracle():
    sc = scenes.config_cpu_32k(seed=5)
    steps = 3
    print(f"[case] facade PBFSolver::solve {sc.name} n={sc.n} steps={steps}", flush=True)
    para = SolverParameter.from_dict(dict(sc.params), sc.n, sc.grid_dims, 0)
    d = tempfile.mkdtemp()
    src, exe = os.path.join(d, "caller.cpp"), os.path.join(d, "caller")
    with open(src, "w") as f:
        f.write(CALLER)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src, "-o", exe, pbf_amd.LIB_PATH,
                    f"-Wl,-rpath,{os.path.dirname(pbf_amd.LIB_PATH)}"], check=True, capture_output=True)
    aos = np.concatenate([sc.pos, sc.vel], axis=1).astype(np.float32)  # Particle{position, velocity}
    fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
    with open(fin, "wb") as f:
        f.write(bytes(para))
        f.write(np.asarray(sc.box, np.float32).tobytes())
        f.write(np.int32(steps).tobytes())
        f.write(np.ascontiguousarray(aos).tobytes())
    r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(fout, np.float32).reshape(sc.n, 6)
    o = Oracle.from_scene(sc)
    o.step(steps)
    op, ov = o.get_state()
    assert np.array_equal(got[:, :3].view(np.uint32), np.ascontiguousarray(op, np.float32).view(np.uint32)), "positions"
    assert np.array_equal(got[:, 3:].view(np.uint32), np.ascontiguousarray(ov, np.float32).view(np.uint32)), "velocities"
