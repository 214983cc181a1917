"""The oracle's pins catch plausible mistakes.

Each case plants one mistake in a copy of the oracle's sources (a sign, an operand order, a
dropped normalisation or factor, a transposed weight), builds that copy on the CPU, and runs
the pin tests that must catch it against the planted build (PBF_ORACLE_LIB): they have to fail.
A planted mistake that every pin lets through would mean the oracle is unpinned at that spot.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
VEL = "tests/test_oracle_velocity.py::"
MLP = "tests/test_mlp_oracle.py::"

# (name, file, original text, planted text, pin tests that must fail)
MUTANTS = [
    ("xsph_sign", "pbf_oracle.cpp", "pv.x = std::fma(-vij.x, w, pv.x);", "pv.x = std::fma(vij.x, w, pv.x);",
     [VEL + "test_curl_and_xsph_against_reference_cpu_solver", VEL + "test_xsph_and_apply_l_shape_closed_form"]),
    ("xsph_own_velocity", "pbf_oracle.cpp", "V3 vij = sub(s->v[i], s->v[nid]);", "V3 vij = s->v[i];",
     [VEL + "test_xsph_conserves_momentum_on_a_lattice", VEL + "test_curl_and_xsph_against_reference_cpu_solver"]),
    ("xsph_no_c", "pbf_oracle.cpp", "s->visc[i] = {P.c * pv.x, P.c * pv.y, P.c * pv.z};",
     "s->visc[i] = {pv.x, pv.y, pv.z};", [VEL + "test_curl_and_xsph_against_reference_cpu_solver"]),
    ("curl_sign", "pbf_oracle.cpp", "V3 vij = sub(s->v[nid], s->v[i]);", "V3 vij = sub(s->v[i], s->v[nid]);",
     [VEL + "test_curl_and_xsph_against_reference_cpu_solver", VEL + "test_vorticity_l_shape_closed_form"]),
    ("vorticity_omega_cross_n", "pbf_oracle.cpp", "V3 cr = cross3(eta, s->curl[i]);",
     "V3 cr = cross3(s->curl[i], eta);", [VEL + "test_vorticity_l_shape_closed_form"]),
    ("eta_not_normalised", "pbf_oracle.cpp", "if ((double)std::fabs(el) > 1e-3) eta",
     "if (false && (double)std::fabs(el) > 1e-3) eta", [VEL + "test_vorticity_l_shape_closed_form"]),
    ("eta_threshold", "pbf_oracle.cpp", "if ((double)std::fabs(el) > 1e-3) eta",
     "if ((double)std::fabs(el) > 1e-2) eta", [VEL + "test_vorticity_l_shape_closed_form"]),
    ("eta_own_omega", "pbf_oracle.cpp", "float wl = std::sqrt(dot3(s->curl[nid], s->curl[nid]));",
     "float wl = std::sqrt(dot3(s->curl[i], s->curl[i]));", [VEL + "test_vorticity_l_shape_closed_form"]),
    ("velocity_times_dt", "pbf_oracle.cpp", "s->v[i] = {(x.x - s->xo[i].x) / dt,", "s->v[i] = {(x.x - s->xo[i].x) * dt,",
     [VEL + "test_loop_positions_and_velocity_against_reference_cpu_solver"]),
    ("apply_without_dt", "pbf_oracle.cpp", "v.x = v.x + (dt * s->vort[i].x + s->visc[i].x);",
     "v.x = v.x + (s->vort[i].x + s->visc[i].x);", [VEL + "test_xsph_and_apply_l_shape_closed_form"]),
    ("hswish_shift", "mlp_oracle.cpp", "return z * std::min(std::max(z + 3.0, 0.0), 6.0) / 6.0;",
     "return z * std::min(std::max(z + 2.0, 0.0), 6.0) / 6.0;",
     [MLP + "test_forward_matches_torch_float64", MLP + "test_forward_closed_forms"]),
    ("hswish_slope", "mlp_oracle.cpp", "(z < 3.0 ? z / 3.0 + 0.5 : 1.0)", "(z < 3.0 ? z / 6.0 + 0.5 : 1.0)",
     [MLP + "test_gradient_matches_torch_autograd", MLP + "test_gradient_matches_finite_differences"]),
    ("w2_transposed", "mlp_oracle.cpp", "z += w(P.W2[o * H + i]) * a1[i];", "z += w(P.W2[i * H + o]) * a1[i];",
     [MLP + "test_forward_matches_torch_float64", MLP + "test_forward_closed_forms"]),
]


def build_mutant(tmp, name, fname, old, new):
    d = os.path.join(tmp, name)
    os.makedirs(d)
    srcs = []
    for f in ("pbf_oracle.cpp", "mlp_oracle.cpp"):
        with open(os.path.join(ORACLE_DIR, f)) as fh:
            s = fh.read()
        if f == fname:
            assert old in s, f"{name}: the text to plant into is gone from {f}"
            s = s.replace(old, new, 1)
        p = os.path.join(d, f)
        with open(p, "w") as fh:
            fh.write(s)
        srcs.append(p)
    so = os.path.join(d, "liboracle.so")
    cmd = ["g++", "-std=c++17", "-O1", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-o", so, *srcs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return so


def run_pins(so, tests):
    env = dict(os.environ, PBF_ORACLE_LIB=so, OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", *tests], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout[-2000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_planted_mistakes_fail_their_pins(tmp_path):
    def one(m):
        name, fname, old, new, tests = m
        so = build_mutant(str(tmp_path), name, fname, old, new)
        return name, tests, run_pins(so, tests)

    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(one, MUTANTS))
    missed = []
    for name, tests, (rc, out) in results:
        # 1: tests ran and failed.  0 would be a mistake every pin let through; other codes are
        # collection or usage errors, which would not show that the pins work.
        if rc != 1:
            missed.append(f"{name} (pytest exit {rc}): {out}")
    assert not missed, "\n".join(missed)


def test_unplanted_copy_passes_its_pins(tmp_path):
    """The same build recipe without a planted mistake passes every pin the cases use."""
    so = build_mutant(str(tmp_path), "none", "", "", "")
    tests = sorted({t for m in MUTANTS for t in m[4]})
    rc, out = run_pins(so, tests)
    assert rc == 0, out
