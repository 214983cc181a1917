"""Host-side pieces without a GPU: scene generators, the reference app's scenes, the state
file, and the batch workload."""

import numpy as np
import pytest

from pbf_amd import SolverParameter, scenes, statefile
from pbf_amd.driver import APP_PARAMS, BOX, reference_scene


def test_splitmix64_known_values():
    # SplitMix64 (Steele, Lea & Flood 2014) from state 0: published first outputs
    z = scenes.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_lattice_shape_and_jitter():
    p = scenes.lattice((4, 3, 2), 0.05, (1.0, 2.0, 3.0), seed=1, jitter=0.005)
    assert p.shape == (24, 3) and p.dtype == np.float32
    i = np.arange(24)
    base = np.stack([i % 4, (i // 4) % 3, i // 12], 1) * 0.05 + np.array([1.0, 2.0, 3.0])
    assert np.abs(p - base).max() <= 0.005 + 1e-6
    assert np.array_equal(p, scenes.lattice((4, 3, 2), 0.05, (1.0, 2.0, 3.0), seed=1, jitter=0.005))


@pytest.mark.parametrize("fn,n,dims", [(scenes.config_cpu_32k, 32768, (16, 16, 16)), (scenes.config_262k, 262144, (64, 40, 32)), (scenes.config_1m, 1048576, (128, 64, 64))])
def test_workload_configs(fn, n, dims):
    sc = fn()
    assert sc.n == n and sc.grid_dims == dims
    assert (sc.pos >= sc.box[:3] - 0.01).all() and (sc.pos <= sc.box[3:] + 0.01).all()
    assert np.allclose(sc.boundary, np.concatenate([sc.box[:3] + 0.03, sc.box[3:] - 0.03]))


def test_reference_units_mapping():
    p = scenes.default_params()
    assert p["restRho"] == pytest.approx(8000.0)
    assert p["k"] == pytest.approx(0.1 * 0.1 * 0.1)
    assert p["c"] == pytest.approx(0.01 * 0.125 / 1000.0)
    assert p["vorticityConfinementEpilon"] == pytest.approx(5e-4 * 0.125 / 1000.0)
    lit = scenes.literal_params()
    assert lit["k"] == 0.1 and lit["c"] == 0.01 and lit["restRho"] == 8000.0


def test_batch_config():
    b = scenes.batch_config()
    assert len(b) == 24
    counts = np.array([s.n for s in b])
    assert counts.min() >= 8192 and counts.max() <= 262144
    assert 1.2e6 < counts.sum() < 2.2e6
    for s in b:
        vol = np.prod(s.box[3:] - s.box[:3])
        fill = s.n * 0.05**3 / vol
        assert 0.15 < fill <= 0.26
        assert (s.pos > s.box[:3]).all() and (s.pos < s.box[3:]).all()


@pytest.mark.parametrize("idx", [1, 2, 3, 4, 5])
def test_reference_scenes(idx):
    sc = reference_scene(idx, n_side=16)
    assert sc.n == 16**3
    assert sc.grid_dims == (64, 64, 64)
    assert np.array_equal(sc.box, BOX)
    assert (sc.pos > BOX[:3]).all() and (sc.pos < BOX[3:]).all()
    assert sc.params == APP_PARAMS
    if idx == 4:
        assert np.allclose(sc.vel, [-10.0, 15.0, 0.0])
    if idx == 3:
        assert np.allclose(sc.vel[: sc.n // 2], [50.0, 0.0, 50.0]) and np.allclose(sc.vel[sc.n // 2 :], [-50.0, 0.0, -50.0])
    if idx == 5:
        speed = np.linalg.norm(sc.vel, axis=1)
        assert np.allclose(speed, 50.0, rtol=1e-5)
        # swirl about the y axis: velocity tangential in the xz plane (renderapp.cpp:230)
        radial = np.stack([sc.pos[:, 0], np.zeros(sc.n), sc.pos[:, 2]], 1)
        assert np.abs((sc.vel * radial).sum(1)).max() < 1e-3 * 50 * np.linalg.norm(radial, axis=1).max()
        assert np.allclose(sc.vel[:, 1], 0.0)


def test_statefile_roundtrip(tmp_path):
    sc = scenes.dam_break((5, 6, 7))
    para = SolverParameter.from_dict(sc.params, sc.n, sc.grid_dims, 3)
    f = tmp_path / "s.pbfs"
    statefile.save(f, para, sc.box, sc.boundary, sc.pos, sc.vel + 1.5)
    p2, gs, bd, pos, vel = statefile.load(f)
    assert bytes(p2) == bytes(para)
    assert np.array_equal(gs, sc.box) and np.array_equal(bd, sc.boundary)
    assert np.array_equal(pos, sc.pos) and np.array_equal(vel, sc.vel + 1.5)
    raw = f.read_bytes()
    f.write_bytes(raw + b"x")
    with pytest.raises(ValueError):
        statefile.load(f)
    f.write_bytes(b"NOTSTATE" + raw[8:])
    with pytest.raises(ValueError):
        statefile.load(f)


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys as _sys
    import os as _os

    root = _os.path.dirname(_os.path.dirname(_os.path.abspath(__file__)))
    r = subprocess.run([_sys.executable, _os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bench_reads_the_committed_counter_summaries():
    """bench.py's roofline takes the dominant kernel's measured traffic and VALU instruction count
    from the committed rocprofv3 summaries; both stay machine-readable for the kernels it can name,
    at 262k and (VALU) 1M, and an unknown kernel or workload gives None."""
    import importlib.util
    import os as _os

    root = _os.path.dirname(_os.path.dirname(_os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_under_test", _os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for stage in ("delta_apply", "lambda", "curl", "vort_apply"):
        assert bench.measured_traffic("dam_break_262k", stage) > 1e6
        for wl in ("dam_break_262k", "double_dam_break_1m"):
            assert bench.measured_valu(wl, stage) > 1e6
    # a step's vorticity pass (with the next step's hash) and the fast-arithmetic passes, at 262k
    for stage in ("vort_apply_hash", "lambda_fast", "delta_apply_fast", "curl_fast", "vort_apply_hash_fast"):
        assert bench.measured_traffic("dam_break_262k", stage) > 1e6
        assert bench.measured_valu("dam_break_262k", stage) > 1e6
    # the fast mode issues fewer VALU instructions per launch (the hardware rsqrt in place of three
    # Newton steps, about a fifth of a neighbour term)
    for stage in ("delta_apply", "lambda"):
        assert bench.measured_valu("dam_break_262k", stage + "_fast") < 0.85 * bench.measured_valu("dam_break_262k", stage)
    assert bench.measured_valu("dam_break_262k", "no_such_stage") is None
    assert bench.measured_valu("no_such_workload", "delta_apply") is None
