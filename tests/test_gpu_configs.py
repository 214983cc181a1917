"""GPU runs of the workloads at their stated sizes, bit for bit against the oracle:

  * the reference app's own workload: scene 4 of src/renderapp.cpp:137-234 (64^3 = 262,144
    particles, h = 1, rest density 6, g = -49 on the fixed 64^3 grid), stepped by a C++ caller
    built against include/pbf_solver.hpp that sets only the 13 fields of the reference's
    SolverParameter (renderapp.cpp:116-129) and reads positions from its own device buffer
    (exportPositions, the reference's VBO ownership, renderapp.cpp:236-249);
  * the app's other four scenes at full size, graph-replayed, against the oracle;
  * the 24-scene batch (about 1.74M particles, one launch per kernel for all scenes): every
    scene equals its own oracle;
  * checkpoint / resume through the state file, and the device export.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pbf_amd
from oracle_lib import Oracle
from pbf_amd import FLAG_GRAPH, DeviceBuffer, Solver, driver, scenes, statefile

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


APP_CALLER = r"""
#include <cstdio>
#include <vector>
#include "pbf_solver.hpp"
using namespace pbf_ref;
// The reference app's set-up and frame loop (src/renderapp.cpp:115-131, 251, 497-502), headless.
int main(int argc, char** argv) {
    if (argc != 4) return 2;
    const int frames = std::atoi(argv[3]);
    SolverParameter para;  // only the reference's 13 fields are set, as renderapp.cpp:116-129
    para.h = 1.0;
    para.nParticles = 64 * 64 * 64;
    para.restRho = 6;
    para.epsilon = 1.0f;
    para.k = 0.01;
    para.deltaq = 0.1 * para.h;
    para.n = 4;
    para.c = 0.1f;
    para.dt = 0.016f;
    para.invDt = 1.0f / para.dt;
    para.iterations = 4;
    para.vorticityConfinementEpilon = 0.1;
    para.radius = 0.3;
    AABB box{{-30.0f, 0.0f, -30.0f}, {30.0f, 40.0f, 30.0f}};
    AABB boundary{{box.pMin.x + para.radius, box.pMin.y + para.radius, box.pMin.z + para.radius},
                  {box.pMax.x - para.radius, box.pMax.y - para.radius, box.pMax.z - para.radius}};
    std::vector<vec3> position(para.nParticles), velocity(para.nParticles);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    bool ok = std::fread(position.data(), sizeof(vec3), position.size(), f) == position.size() &&
              std::fread(velocity.data(), sizeof(vec3), velocity.size(), f) == velocity.size();
    std::fclose(f);
    if (!ok) return 4;
    std::vector<vec3> host(para.nParticles);
    try {
        void* vbo = nullptr;  // the app's own device buffer (a mapped VBO in the reference)
        check(pbf_device_alloc(sizeof(vec3) * para.nParticles, &vbo));
        SolverData data;
        initCUDASolverDataBuffer(position, velocity, data, para, box, boundary);
        for (int frame = 0; frame < frames; ++frame) cudaPBFSolve(data, para);
        exportPositions(data, static_cast<float*>(vbo));
        check(pbf_memcpy_to_host(host.data(), vbo, sizeof(vec3) * host.size()));
        readState(data, position, velocity);
        freeCUDASolverDataBuffer(data);
        check(pbf_device_free(vbo));
    } catch (const Error& e) {
        std::printf("error %d %s\n", e.code, e.what());
        return 5;
    }
    FILE* g = std::fopen(argv[2], "wb");
    if (!g) return 6;
    std::fwrite(host.data(), sizeof(vec3), host.size(), g);
    std::fwrite(position.data(), sizeof(vec3), position.size(), g);
    std::fwrite(velocity.data(), sizeof(vec3), velocity.size(), g);
    std::fclose(g);
    std::printf("frames %d\n", frames);
    return 0;
}
"""


def test_reference_app_scene4_through_the_facade():
    sc = driver.reference_scene(4)
    frames = 3
    print(f"[case] reference app scene 4 via the C++ facade, n={sc.n}, {frames} frames", flush=True)
    d = tempfile.mkdtemp()
    src, exe = os.path.join(d, "app.cpp"), os.path.join(d, "app")
    with open(src, "w") as f:
        f.write(APP_CALLER)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src, "-o", exe, pbf_amd.LIB_PATH,
                    f"-Wl,-rpath,{os.path.dirname(pbf_amd.LIB_PATH)}"], check=True, capture_output=True)
    fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
    with open(fin, "wb") as f:
        f.write(np.ascontiguousarray(sc.pos, np.float32).tobytes())
        f.write(np.ascontiguousarray(sc.vel, np.float32).tobytes())
    r = subprocess.run([exe, fin, fout, str(frames)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(fout, np.float32).reshape(3, sc.n, 3)
    o = Oracle.from_scene(sc)
    o.step(frames)
    op, ov = o.get_state()
    assert same_bits(out[0], op), "positions from the app's device buffer"
    assert same_bits(out[1], op), "positions (readState)"
    assert same_bits(out[2], ov), "velocities (readState)"
    assert np.abs(op - sc.pos).max() > 0.1, "the scene moves"


@pytest.mark.parametrize("index", [1, 2, 3, 5])
def test_reference_app_scenes_equal_the_oracle(index):
    """The reference app's other scenes (src/renderapp.cpp:137-234) at their full 64^3 lattices
    with the app's parameters and the fixed 64^3 grid: a column thrown up at 20 m/s (1), a block
    thrown sideways (2), two blocks colliding at 50 m/s (3), a 50 m/s swirl (5).  Ten
    graph-replayed steps, bit for bit against the oracle; the scenes reach the walls and the
    grid's edges."""
    sc = driver.reference_scene(index)
    print(f"[case] reference app scene {index}, n={sc.n}, 10 steps", flush=True)
    g = Solver.from_scene(sc, flags=FLAG_GRAPH)
    g.advance(10)
    gp, gv = g.get_state()
    g.close()
    o = Oracle.from_scene(sc)
    o.step(10)
    op, ov = o.get_state()
    assert same_bits(gp, op), "positions"
    assert same_bits(gv, ov), "velocities"
    assert np.abs(op - sc.pos).max() > 0.5, "the scene moves"


def test_batch_of_24_scenes_each_equals_its_oracle():
    batch = scenes.batch_config()
    n = sum(sc.n for sc in batch)
    print(f"[case] batch of {len(batch)} scenes, {n} particles, 2 steps", flush=True)
    assert len(batch) == 24 and n > 1_000_000
    b = Solver.batch(batch, flags=FLAG_GRAPH)
    b.advance(2)
    pos, vel = b.get_state()
    b.close()
    off = 0
    for k, sc in enumerate(batch):
        o = Oracle.from_scene(sc)
        o.step(2)
        op, ov = o.get_state()
        assert same_bits(pos[off:off + sc.n], op), f"scene {k} positions"
        assert same_bits(vel[off:off + sc.n], ov), f"scene {k} velocities"
        off += sc.n
        del o


def test_statefile_resume_continues_bit_for_bit(tmp_path):
    """A run checkpointed to the state file and resumed in a new solver equals the uninterrupted
    run: the state file holds everything a step depends on."""
    sc = scenes.config_262k(seed=3)
    print(f"[case] state file resume {sc.name} n={sc.n}", flush=True)
    a = Solver.from_scene(sc, flags=FLAG_GRAPH)
    a.advance(10)
    pos, vel = a.get_state()
    path = tmp_path / "mid.pbfs"
    statefile.save(path, a.para, sc.box, sc.boundary, pos, vel)
    a.advance(10)
    pa, va = a.get_state()
    a.close()
    para, gs, bd, p0, v0 = statefile.load(path)
    b = Solver(para, gs, bd, p0, v0)
    b.advance(10)
    pb, vb = b.get_state()
    b.close()
    assert same_bits(pa, pb) and same_bits(va, vb)


def test_export_device_equals_get_state():
    sc = scenes.config_cpu_32k(seed=8)
    print(f"[case] export_device {sc.name} n={sc.n}", flush=True)
    s = Solver.from_scene(sc, flags=FLAG_GRAPH)
    s.advance(5)
    dp = DeviceBuffer(sc.n * 12)
    dv = DeviceBuffer(sc.n * 12)
    s.export_device(dp.ptr, dv.ptr)
    s.synchronize()
    ep = dp.to_host(np.float32, sc.n * 3).reshape(sc.n, 3)
    ev = dv.to_host(np.float32, sc.n * 3).reshape(sc.n, 3)
    gp, gv = s.get_state()
    assert same_bits(ep, gp) and same_bits(ev, gv)
    # positions only (velocity pointer NULL)
    dp2 = DeviceBuffer(sc.n * 12)
    s.export_device(dp2.ptr, None)
    s.synchronize()
    assert same_bits(dp2.to_host(np.float32, sc.n * 3).reshape(sc.n, 3), gp)
    s.close()
    for buf in (dp, dv, dp2):
        buf.close()


def test_device_view_is_the_sorted_state():
    """pbf_device_state: the device pointers of the cell-sorted state (the reference's SolverData
    buffers, include/solverBase.h:45-47).  Slot k holds particle ids[k]: its position and
    velocity equal pbf_get_state's row ids[k] bit for bit, ids is a permutation, w of the
    velocities is 0 and w of the positions is the step's |curl|."""
    sc = scenes.config_cpu_32k(seed=14)
    print(f"[case] device view {sc.name}", flush=True)
    s = Solver.from_scene(sc, flags=FLAG_GRAPH | pbf_amd.FLAG_DEBUG)
    s.advance(5)
    s.synchronize()
    v = s.device_view()
    assert v.n == sc.n and v.ncells >= int(np.prod(sc.grid_dims))
    p4 = np.empty((sc.n, 4), np.float32)
    v4 = np.empty((sc.n, 4), np.float32)
    ids = np.empty(sc.n, np.uint32)
    L = pbf_amd.lib()
    for dst, src in ((p4, v.pos4), (v4, v.vel4), (ids, v.ids)):
        assert L.pbf_memcpy_to_host(dst.ctypes.data, src, dst.nbytes) == 0
    gp, gv = s.get_state()
    assert np.array_equal(np.sort(ids), np.arange(sc.n, dtype=np.uint32))
    assert same_bits(p4[:, :3], gp[ids]) and same_bits(v4[:, :3], gv[ids])
    assert not v4[:, 3].any()
    curl = s.fields()[3][ids].astype(np.float32)
    mag = np.sqrt((curl[:, 0] * curl[:, 0] + curl[:, 1] * curl[:, 1]) + curl[:, 2] * curl[:, 2])
    np.testing.assert_allclose(p4[:, 3], mag, rtol=1e-6, atol=0)
    s.close()


def test_stage_times_show_the_fused_hash():
    """pbf_stage_times: seven stage times of one step.  After a step, the next one's predict and
    hash ran in its vorticity pass, so the first stage holds only the events' own interval; after
    pbf_set_state the step hashes for itself and that stage takes longer by the hash kernel."""
    sc = scenes.config_262k(seed=15)
    print(f"[case] stage times {sc.name}", flush=True)
    s = Solver.from_scene(sc)
    s.advance(3)
    t = s.stage_times()
    assert (t >= 0).all() and t[3] > 0.01 and t[4] > 0.05
    p, v = s.get_state()
    s.set_state(p, v)
    u = s.stage_times()
    print(f"  after a step: {np.round(t, 4).tolist()}; after set_state: {np.round(u, 4).tolist()}", flush=True)
    # two events with nothing between them already read ~5 us apart; the hash adds ~4 us at 262k
    assert u[0] > t[0] + 0.0015
    s.close()


def test_graphs_survive_set_state():
    """pbf_set_state keeps the captured step graphs: steps after an upload replay them and equal
    eager steps from the same upload."""
    sc = scenes.config_cpu_32k(seed=9)
    print(f"[case] graphs across set_state {sc.name}", flush=True)
    g = Solver.from_scene(sc, flags=FLAG_GRAPH)
    e = Solver.from_scene(sc)
    g.advance(9)
    e.advance(9)
    for k in range(3):
        p, v = e.get_state()
        v = v * np.float32(0.5)
        g.set_state(p, v)
        e.set_state(p, v)
        g.advance(9)
        e.advance(9)
        gp, gv = g.get_state()
        ep, ev = e.get_state()
        assert same_bits(gp, ep) and same_bits(gv, ev), f"round {k}"
    g.close()
    e.close()


def test_dt_changing_every_step_matches_eager():
    """A caller that adapts dt on every call (pbf_step): with PBF_FLAG_GRAPH the steps from the
    second change in a row on launch eagerly, then graphs again once dt holds; every step equals
    an eager solver's bit for bit, and so do the graph-replayed steps of a following advance."""
    sc = scenes.config_cpu_32k(seed=16)
    print("[case] dt changing every step", flush=True)
    g = Solver.from_scene(sc, flags=FLAG_GRAPH)
    e = Solver.from_scene(sc)
    dts = [0.008, 0.0075, 0.007, 0.0072, 0.0072, 0.0072, 0.0068, 0.0081, 0.0081]
    for k, dt in enumerate(dts):
        g.step(dt)
        e.step(dt)
        gp, gv = g.get_state()
        ep, ev = e.get_state()
        assert same_bits(gp, ep) and same_bits(gv, ev), f"step {k} (dt {dt})"
    g.advance(9)
    e.advance(9)
    gp, gv = g.get_state()
    ep, ev = e.get_state()
    assert same_bits(gp, ep) and same_bits(gv, ev)
    g.close()
    e.close()


@pytest.mark.parametrize("change", ["set_state", "dt", "ordering", "parameters"])
def test_changes_between_steps_rehash(change):
    """A step predicts, hashes and counts the next step's cells with the step's own parameters,
    state and row order.  After any of those changes, the next step hashes again: the solver
    equals a fresh one created from the same state with the same settings, bit for bit."""
    sc = scenes.config_cpu_32k(seed=10)
    print(f"[case] rehash after {change}", flush=True)
    a = Solver.from_scene(sc, flags=FLAG_GRAPH)
    a.advance(5)
    p, v = a.get_state()
    para = a.para.copy()
    if change == "set_state":
        p = p + np.float32(0.013)
        a.set_state(p, v)
    b = Solver(para, sc.box, sc.boundary, p, v)
    b.set_tuning(*a.get_tuning())
    if change == "dt":
        a.step(0.005)
        b.step(0.005)
    elif change == "ordering":
        a.set_ordering(1)
        b.set_ordering(1)
    elif change == "parameters":
        q = para.copy()
        q.gravity[0] = 2.0
        a.solve(q)
        b.solve(q)
    a.advance(3)
    b.advance(3)
    pa, va = a.get_state()
    pb, vb = b.get_state()
    a.close()
    b.close()
    assert same_bits(pa, pb) and same_bits(va, vb)
