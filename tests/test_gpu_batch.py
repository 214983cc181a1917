"""Batched scenes and launch-shape tuning on the GPU.

A batch advances every scene in each launch through one counting sort over the concatenated
cells of all scenes; each scene must evolve bit for bit as the oracle evolves it alone.  Launch
shapes (workgroup size, persistent grid, XCD mapping) must not change any result, and the
autotuner must leave the state where it found it."""
import numpy as np
import pytest

from oracle_lib import Oracle
from pbf_amd import FLAG_GRAPH, Solver, scenes

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def small_batch():
    sizes = [1000, 4913, 2197, 8000, 512]
    return [scenes.scaled_dam_break(n, 500 + k, fill=0.25) for k, n in enumerate(sizes)]


@pytest.mark.parametrize("flags", [0, FLAG_GRAPH])
def test_batch_matches_each_scene_alone(flags):
    sl = small_batch()
    print(f"[case] batch of {len(sl)} scenes, {sum(s.n for s in sl)} particles, flags={flags}", flush=True)
    b = Solver.batch(sl, flags=flags)
    orcs = [Oracle.from_scene(sc) for sc in sl]
    for _ in range(3):
        b.advance(7)
        for o in orcs:
            o.step(7)
        p, v = b.get_state()
        off = 0
        for sc, o in zip(sl, orcs):
            op, ov = o.get_state()
            assert same_bits(p[off : off + sc.n], op), sc.name
            assert same_bits(v[off : off + sc.n], ov), sc.name
            off += sc.n


@pytest.mark.parametrize("flags", [0, FLAG_GRAPH])
def test_morton_row_order_batch_matches_each_scene_alone(flags):
    """Each scene's rows in Morton order (its own table, at its offset in the batch): every scene
    still evolves bit for bit as the oracle evolves it alone."""
    sl = small_batch()
    print(f"[case] batch with Morton rows, flags={flags}", flush=True)
    b = Solver.batch(sl, flags=flags)
    b.set_ordering(1)
    orcs = [Oracle.from_scene(sc) for sc in sl]
    for _ in range(2):
        b.advance(7)
        for o in orcs:
            o.step(7)
        p, v = b.get_state()
        off = 0
        for sc, o in zip(sl, orcs):
            op, ov = o.get_state()
            assert same_bits(p[off : off + sc.n], op), sc.name
            assert same_bits(v[off : off + sc.n], ov), sc.name
            off += sc.n


@pytest.mark.parametrize("block,grid,xcd", [(64, 0, True), (128, 0, False), (512, 0, True), (256, 64, True), (128, 37, False)])
def test_launch_shape_does_not_change_results(block, grid, xcd):
    print(f"[case] tuning block={block} grid={grid} xcd={xcd}", flush=True)
    sc = scenes.config_cpu_32k()
    ref = Solver.from_scene(sc)
    t = Solver.from_scene(sc)
    t.set_tuning(block, grid, xcd)
    ref.advance(5)
    t.advance(5)
    pr, vr = ref.get_state()
    pt, vt = t.get_state()
    assert same_bits(pr, pt) and same_bits(vr, vt)


def test_autotune_keeps_state_and_picks_a_shape():
    print("[case] autotune", flush=True)
    sc = scenes.config_cpu_32k()
    a = Solver.from_scene(sc)
    b = Solver.from_scene(sc)
    a.advance(3)
    b.advance(3)
    best, ms = b.autotune(steps=3)
    assert best[0] in (64, 128, 256, 512) and ms > 0
    assert b.get_tuning() == best
    assert b.get_ordering() in (0, 1)
    pa, va = a.get_state()
    pb, vb = b.get_state()
    assert same_bits(pa, pb) and same_bits(va, vb)
    a.advance(4)
    b.advance(4)
    pa, va = a.get_state()
    pb, vb = b.get_state()
    assert same_bits(pa, pb) and same_bits(va, vb)


def test_batch_learned_step_with_rejected_network_is_plain():
    """Learned steps on a batch (features use each scene's own walls): with a network whose pass
    scatters the fluid, the device gate rejects it and the batch evolves exactly as with plain
    steps of the same iteration count."""
    import numpy as _np

    from pbf_amd import Mlp

    sl = small_batch()
    print(f"[case] batched learned step, {len(sl)} scenes", flush=True)
    plain = Solver.batch(sl, flags=FLAG_GRAPH, iterations=3)
    lrn = Solver.batch(sl, flags=FLAG_GRAPH, iterations=3)
    m = Mlp(seed=13)
    rng = _np.random.default_rng(13)
    m.set_params((rng.standard_normal(9473) * 3.0).astype(_np.float32))
    m.set_normalization(_np.zeros(16), _np.ones(16), 1e3, 0.0)
    lrn.attach_mlp(m)
    lrn.set_learned(3)
    plain.advance(5)
    lrn.advance(5)
    assert not lrn.last_guess_info()[2]
    pa, va = plain.get_state()
    pb, vb = lrn.get_state()
    assert same_bits(pa, pb) and same_bits(va, vb)
