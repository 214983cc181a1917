"""Pins the velocity end of the oracle's step: the velocity update after the constraint loop
(kernels/cudaPBFSolver.cu:21), the curl (cu:201-215), the XSPH viscosity and vorticity
confinement (cu:217-239, cuh:146-148) and their application (cu:24-29).

  * Against the reference's own CPU solver (src/solver.cpp:60-157), which computes the XSPH sum
    with the same formula (:138-150), the loop velocity as (x - x_old) * invDt (:117-120) and the
    curl with v_i - v_j (:128), the negation of the GPU's v_j - v_i (cu:211): golden vectors in
    tests/golden/ref_step_probe.npz and, where oracle/_ref was built, a live run.  The CPU
    solver's vorticity differs from the GPU's (a scalar Wpoly6 sum for eta and a fixed 0.5 for
    epsilon, solver.h:141-143, solver.cpp:143), so it pins nothing there.
  * Vorticity confinement by closed forms: an L-shaped triple of particles whose curls, eta,
    N = eta / |eta| (only when |eta| > 1e-3, cu:234) and forces eps N x omega are worked out by
    hand from the kernel gradient's magnitude G(d) = 45 / (pi h^6) (h - d)^2, on both sides of
    the threshold.
  * Invariants: XSPH conserves momentum on a lattice (sum of an antisymmetric pair term).
"""
import math
import os

import numpy as np
import pytest

from oracle_lib import Oracle, ref_available, ref_step_probe
from pbf_amd import SolverParameter

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
F32 = np.float32
PROBE_KEYS = ("h", "restRho", "epsilon", "k", "n", "deltaq", "radius", "c", "dt", "iterations",
              "vorticityConfinementEpilon")


def oracle_for(params: dict, box, pos, vel, dims, **over):
    """An oracle solver with the reference CPU solver's fixed constants (gravity -9.8 on y,
    spiky cut-off 0.01, cap 63) unless overridden."""
    p = dict(params)
    p.setdefault("gravity", (0.0, -9.8, 0.0))
    p.setdefault("jitter_coeff", 0.02)
    p.setdefault("grad_rmin", 0.01)
    p.setdefault("max_neighbors", 63)
    p.update(over)
    pos = np.ascontiguousarray(pos, F32)
    box = np.ascontiguousarray(box, F32)
    r = F32(p["radius"])
    bd = np.concatenate([box[:3] + r, box[3:] - r]).astype(F32)
    para = SolverParameter.from_dict(p, len(pos), dims, 0)
    o = Oracle(para, box, bd, pos, vel)
    o.step(1)
    return o


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def reference_step_vectors():
    """The golden step probe; the live reference run of the same inputs where it was built."""
    z = np.load(os.path.join(GOLDEN, "ref_step_probe.npz"))
    params = {k: float(v) for k, v in zip(PROBE_KEYS, z["params"])}
    params["iterations"] = int(params["iterations"])
    gold = dict(pos=z["pos_out"], vel_loop=z["vel_loop"], curl=z["curl"], visc=z["visc"], lam=z["lam"])
    out = [("golden", gold)]
    if ref_available():
        live = ref_step_probe(params, z["box"], z["pos"], z["vel"])
        live["lam"] = live.pop("lambda")
        out.append(("live", live))
    return params, z["box"], z["pos"], z["vel"], out


def test_step_probe_inputs_stay_inside_the_boundary():
    """The comparison needs no clamp (the CPU solver's collsionProcess jitter differs)."""
    params, box, pos, vel, outs = reference_step_vectors()
    r = params["radius"]
    for _, ref in outs:
        assert (ref["pos"] > box[:3] + r).all() and (ref["pos"] < box[3:] - r).all()


def test_loop_positions_and_velocity_against_reference_cpu_solver():
    """Three Jacobi iterations, then v = (x - x_old) / dt (cu:21; the CPU solver multiplies by
    invDt, solver.cpp:119, one rounding apart).  With c = 0 and eps = 0 the oracle's final
    velocity is exactly its loop velocity."""
    params, box, pos, vel, outs = reference_step_vectors()
    o = oracle_for(params, box, pos, vel, (64, 64, 64), c=0.0, vorticityConfinementEpilon=0.0)
    op, ov = o.get_state()
    lf, ll, er, cu = o.fields()
    for name, ref in outs:
        disp = np.abs(ref["pos"] - pos).max()
        assert disp > 0.01, "the probe must move"
        assert np.abs(op.astype(np.float64) - ref["pos"]).max() < 2e-6, name
        assert rel_err(ov, ref["vel_loop"]) < 1e-4, name
        assert rel_err(ll, ref["lam"]) < 2e-5, name


def test_curl_and_xsph_against_reference_cpu_solver():
    """curl: the oracle's sum_j (v_j - v_i) x gradW (cu:211) is minus the CPU solver's
    sum_j (v_i - v_j) x gradW (solver.cpp:128).  XSPH: c sum_j -(v_i - v_j) W(r) in both
    (cu:229-238, solver.cpp:140-150)."""
    params, box, pos, vel, outs = reference_step_vectors()
    o = oracle_for(params, box, pos, vel, (64, 64, 64))
    lf, ll, er, cu = o.fields()
    vort, visc = o.vort_visc()
    for name, ref in outs:
        assert np.abs(ref["visc"]).max() > 1.0 and np.abs(ref["curl"]).max() > 1.0
        assert rel_err(cu, -ref["curl"]) < 1e-4, name
        assert rel_err(visc, ref["visc"]) < 1e-4, name
    # eps = 0: no vorticity force
    assert not vort.any()


# ---- vorticity confinement by hand ---------------------------------------------------------
H1 = 1.0
CS = 45.0 / (math.pi * H1**6)  # |spiky gradient constant| (cuh:108)
C6 = 315.0 / (64.0 * math.pi * H1**9)  # poly6 constant (cuh:101)


def G(d):
    """|gradW| at distance d (cuh:105-110)."""
    return CS * (H1 - d) ** 2


def Wp(d):
    return C6 * (H1 * H1 - d * d) ** 3


D1, D2 = 0.7, 0.8  # P-Q1 along +x, P-Q2 along +y; |Q1 - Q2| = 1.063 > h: Q1, Q2 not neighbours


def l_shape(a, b, c, eps, dt):
    """P at rest, Q1 = P + D1 x moving (0, a, 0), Q2 = P + D2 y moving (b, 0, 0); h = 1, no
    gravity, no iteration (the velocities after the loop are the given ones)."""
    P0 = np.array([4.0, 4.0, 4.0])
    pos = np.array([P0, P0 + [D1, 0, 0], P0 + [0, D2, 0]], F32)
    vel = np.array([[0, 0, 0], [0, a, 0], [b, 0, 0]], F32)
    params = dict(h=H1, restRho=6.0, epsilon=1.0, k=0.0, n=4.0, deltaq=0.2, radius=0.1, c=c, dt=dt, iterations=0,
                  vorticityConfinementEpilon=eps)
    o = oracle_for(params, (0, 0, 0, 8, 8, 8), pos, vel, (8, 8, 8), gravity=(0.0, 0.0, 0.0))
    return o, vel


def l_shape_closed_form(a, b, eps):
    """Curls, eta and forces of the L-shaped triple, worked out by hand.

    gradW(x_P - x_Q1) = G1 x, gradW(x_P - x_Q2) = G2 y (the kernel's constant is negative), so
      omega_P  = (v_Q1 - v_P) x G1 x + (v_Q2 - v_P) x G2 y = (0, 0, b G2 - a G1)
      omega_Q1 = (v_P - v_Q1) x (-G1 x)                     = (0, 0, -a G1)
      omega_Q2 = (v_P - v_Q2) x (-G2 y)                     = (0, 0, b G2)
      eta_P  = |omega_Q1| G1 x + |omega_Q2| G2 y,  eta_Q1 = -|omega_P| G1 x,
      eta_Q2 = -|omega_P| G2 y,
    and N = eta / |eta| only when |eta| > 1e-3; force = eps N x omega."""
    G1, G2 = G(D1), G(D2)
    wP = np.array([0.0, 0.0, b * G2 - a * G1])
    wQ1 = np.array([0.0, 0.0, -a * G1])
    wQ2 = np.array([0.0, 0.0, b * G2])
    etas = [
        np.array([abs(a * G1) * G1, abs(b * G2) * G2, 0.0]),
        np.array([-abs(wP[2]) * G1, 0.0, 0.0]),
        np.array([0.0, -abs(wP[2]) * G2, 0.0]),
    ]
    curls = [wP, wQ1, wQ2]
    forces = []
    for e, w in zip(etas, curls):
        n_ = e / np.linalg.norm(e) if np.linalg.norm(e) > 1e-3 else e
        forces.append(eps * np.cross(n_, w))
    return np.array(curls), np.array(etas), np.array(forces)


@pytest.mark.parametrize("scale", [1.0, 4e-3, 4e-4])
def test_vorticity_l_shape_closed_form(scale):
    """scale 1: every |eta| far above 1e-3 (normalised); 4e-3: every |eta| in (1e-3, 1e-2)
    (normalised); 4e-4: every |eta| below 1e-3 (left unnormalised, cu:234)."""
    a, b, eps = 0.5 * scale, -0.3 * scale, 0.7
    o, _ = l_shape(a, b, c=0.0, eps=eps, dt=1e-6)
    curls, etas, forces = l_shape_closed_form(a, b, eps)
    norms = np.linalg.norm(etas, axis=1)
    if scale == 1.0:
        assert norms.min() > 0.05
    elif scale == 4e-3:
        assert norms.min() > 1e-3 and norms.max() < 1e-2
    else:
        assert norms.max() < 1e-3
    nb, cnt = o.neighbors()
    assert cnt.tolist() == [2, 1, 1]
    cu = o.fields()[3]
    vort, visc = o.vort_visc()
    np.testing.assert_allclose(cu, curls, rtol=2e-5, atol=2e-5 * np.abs(curls).max())
    np.testing.assert_allclose(vort, forces, rtol=2e-5, atol=2e-5 * np.abs(forces).max())
    assert not visc.any()


def test_xsph_and_apply_l_shape_closed_form():
    """XSPH: visc_i = c sum_j (v_j - v_i) W(d_ij) pulls each velocity toward its neighbours'
    (visc_P = c (b W2, a W1, 0), visc_Q1 = -c a W1 y, visc_Q2 = -c b W2 x), and the update is
    v += dt * (eps N x omega) + visc (cu:24-29): the particle at rest ends at exactly that."""
    a, b, eps, c, dt = 0.5, -0.3, 0.7, 0.05, 1e-3
    o, vel = l_shape(a, b, c=c, eps=eps, dt=dt)
    W1, W2 = Wp(D1), Wp(D2)
    want_visc = np.array([[c * b * W2, c * a * W1, 0.0], [0.0, -c * a * W1, 0.0], [-c * b * W2, 0.0, 0.0]])
    vort, visc = o.vort_visc()
    # the pair moves by dt * v (~5e-4 of d): first-order changes of W and G at that relative size
    np.testing.assert_allclose(visc, want_visc, rtol=4e-3, atol=4e-3 * np.abs(want_visc).max())
    _, _, forces = l_shape_closed_form(a, b, eps)
    np.testing.assert_allclose(vort, forces, rtol=4e-3, atol=4e-3 * np.abs(forces).max())
    _, v = o.get_state()
    want_vP = dt * forces[0] + want_visc[0]
    np.testing.assert_allclose(v[0], want_vP, rtol=4e-3, atol=4e-3 * np.abs(want_vP).max())
    assert np.abs(v[0] - (dt * vort[0] + visc[0])).max() <= 1e-7 * np.abs(v[0]).max()
    # momentum: the pair terms cancel
    assert np.abs(visc.astype(np.float64).sum(0)).max() <= 1e-6 * np.abs(visc).max()


def test_xsph_conserves_momentum_on_a_lattice():
    """sum_i sum_j (v_j - v_i) W(r_ij) = 0 when the neighbour relation is symmetric (no cap hit):
    a term with v_i in place of v_j, or a missing pair, breaks it."""
    from pbf_amd import scenes

    pos = scenes.lattice((9, 9, 9), 0.05, (1.0, 1.0, 1.0), seed=9, jitter=0.006)
    vel = (scenes.uniform24(10, pos.size).reshape(pos.shape) * 2.0 - 1.0).astype(F32)
    params = dict(h=0.1, restRho=8000.0, epsilon=100.0, k=1e-3, n=4.0, deltaq=0.02, radius=0.03, c=1e-4,
                  dt=float(F32(1.0 / 120.0)), iterations=2, vorticityConfinementEpilon=0.0)
    o = oracle_for(params, (0, 0, 0, 3.2, 3.2, 3.2), pos, vel, (32, 32, 32), gravity=(0.0, 0.0, 0.0))
    nb, cnt = o.neighbors()
    assert cnt.max() < 63
    vort, visc = o.vort_visc()
    v = visc.astype(np.float64)
    assert np.abs(v).sum(0).min() > 0
    assert np.abs(v.sum(0)).max() <= 1e-5 * np.abs(v).sum(0).max()
