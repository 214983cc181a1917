"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was built, to the
reference's own CPU solver (oracle/_ref/libpbfref.so).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from pbf_amd import SolverParameter

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# PBF_ORACLE_LIB: another build of the oracle sources (the planted-mistake builds of
# tests/test_oracle_mutants.py)
ORACLE = os.environ.get("PBF_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libpbfref.so")

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)
_U = C.POINTER(C.c_uint32)
_VP = C.c_void_p

_orc = None
_ref = None


def fp(a):
    return None if a is None else a.ctypes.data_as(_F)


def ip(a):
    return None if a is None else a.ctypes.data_as(_I)


def up(a):
    return None if a is None else a.ctypes.data_as(_U)


def orc() -> C.CDLL:
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = C.CDLL(ORACLE)
        sig = {
            "orc_sin": (C.c_float, [C.c_float]),
            "orc_rsqrt": (None, [_F, _F, C.c_int64]),
            "orc_random_offset": (C.c_float, [C.c_float, C.c_float]),
            "orc_bound": (None, [_F, _F, C.c_float]),
            "orc_consts": (None, [C.POINTER(SolverParameter), _F]),
            "orc_poly6": (C.c_float, [C.POINTER(SolverParameter), _F]),
            "orc_spiky_grad": (None, [C.POINTER(SolverParameter), _F, _F]),
            "orc_exclusive_scan": (None, [_I, C.c_int64, _I]),
            "orc_cell_ids": (None, [_F, C.c_int32, _F, C.c_float, _I, _I]),
            "orc_create": (_VP, [C.POINTER(SolverParameter), _F, _F, _F, _F, C.c_int32]),
            "orc_destroy": (None, [_VP]),
            "orc_set_threads": (None, [C.c_int]),
            "orc_max_threads": (C.c_int, []),
            "orc_step": (None, [_VP, C.c_int32]),
            "orc_get_state": (None, [_VP, _F, _F]),
            "orc_set_state": (None, [_VP, _F, _F]),
            "orc_get_grid": (C.c_int32, [_VP, _I, _I, _I, _I]),
            "orc_get_neighbors": (None, [_VP, _I, _I]),
            "orc_get_fields": (None, [_VP, _F, _F, _F, _F]),
            "orc_get_dp": (None, [_VP, _F]),
            "orc_get_vort_visc": (None, [_VP, _F, _F]),
            "mlp_param_count": (C.c_int, []),
            "mlp_bf16": (C.c_float, [C.c_float]),
            "mlp_forward": (None, [_F, _F, C.c_int, C.c_int, C.POINTER(C.c_double)]),
            "mlp_loss_grad": (C.c_double, [_F, _F, _F, C.c_int, C.c_int, C.POINTER(C.c_double)]),
            "mlp_adam": (None, [_F, _F, _F, _F, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _orc = L
    return _orc


def ref_available() -> bool:
    return os.path.exists(REF)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        L = C.CDLL(REF)
        L.ref_scan_serial.argtypes = [_U, _U, C.c_int]
        L.ref_scan_blelloch.argtypes = [_U, _U, C.c_int]
        L.ref_wpoly6.restype = C.c_float
        L.ref_wpoly6.argtypes = [_F, C.c_float]
        L.ref_spiky_grad.argtypes = [_F, C.c_float, _F]
        L.ref_probe.restype = C.c_int
        L.ref_probe.argtypes = [_F, _F, _F, C.c_int, _I, _I, _I, _I, _I, _F, _F]
        L.ref_grid_cells.restype = C.c_int
        L.ref_step_probe.restype = C.c_int
        L.ref_step_probe.argtypes = [_F, _F, _F, _F, C.c_int, _F, _F, _F, _F, _F, _F]
        _ref = L
    return _ref


class Oracle:
    """One oracle solver state (caller's particle order kept throughout)."""

    def __init__(self, para: SolverParameter, grid_space, boundary, pos, vel=None):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        self.n = pos.shape[0]
        vel = np.zeros_like(pos) if vel is None else np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        self.para = para.copy()
        self.para.nParticles = self.n
        self.ncells = int(para.grid_dims[0] * para.grid_dims[1] * para.grid_dims[2])
        gs = np.ascontiguousarray(grid_space, np.float32)
        bd = np.ascontiguousarray(boundary, np.float32)
        self._h = orc().orc_create(C.byref(self.para), fp(gs), fp(bd), fp(pos), fp(vel), self.n)

    @classmethod
    def from_scene(cls, scene, **over):
        p = dict(scene.params)
        p.update(over)
        para = SolverParameter.from_dict(p, scene.n, scene.grid_dims, 0)
        return cls(para, scene.box, scene.boundary, scene.pos, scene.vel)

    def __del__(self):
        try:
            if self._h:
                orc().orc_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def step(self, nsteps: int = 1):
        orc().orc_step(self._h, int(nsteps))

    def get_state(self):
        p = np.empty((self.n, 3), np.float32)
        v = np.empty((self.n, 3), np.float32)
        orc().orc_get_state(self._h, fp(p), fp(v))
        return p, v

    def set_state(self, pos, vel):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        orc().orc_set_state(self._h, fp(pos), fp(vel))

    def grid(self):
        gid = np.empty(self.n, np.int32)
        cnt = np.empty(self.ncells, np.int32)
        cs = np.empty(self.ncells + 1, np.int32)
        srt = np.empty(self.n, np.int32)
        nin = orc().orc_get_grid(self._h, ip(gid), ip(cnt), ip(cs), ip(srt))
        return gid, cnt, cs, srt, nin

    def neighbors(self):
        cap = int(self.para.max_neighbors)
        nbr = np.empty((self.n, cap), np.int32)
        cnt = np.empty(self.n, np.int32)
        orc().orc_get_neighbors(self._h, ip(nbr), ip(cnt))
        # entries past the count are unspecified in the oracle; normalise to -1
        mask = np.arange(cap)[None, :] >= cnt[:, None]
        nbr[mask] = -1
        return nbr, cnt

    def dp(self):
        out = np.empty((self.n, 3), np.float32)
        orc().orc_get_dp(self._h, fp(out))
        return out

    def vort_visc(self):
        """(eps N x omega, c sum_j -(v_i - v_j) W) of the last step, by particle id."""
        vo = np.empty((self.n, 3), np.float32)
        vi = np.empty((self.n, 3), np.float32)
        orc().orc_get_vort_visc(self._h, fp(vo), fp(vi))
        return vo, vi

    def fields(self):
        lf = np.empty(self.n, np.float32)
        ll = np.empty(self.n, np.float32)
        er = np.empty(self.n, np.float32)
        cu = np.empty((self.n, 3), np.float32)
        orc().orc_get_fields(self._h, fp(lf), fp(ll), fp(er), fp(cu))
        return lf, ll, er, cu


def ref_step_probe(params: dict, box, pos, vel):
    """One PBFSolver::solve of the reference's CPU solver (src/solver.cpp:60-157) on (pos, vel):
    dict of positions after the constraint loop, the loop's velocity (x - x_old) * invDt, curl,
    XSPH viscosity, final velocity and the last iteration's lambda."""
    prm = np.array([params[k] for k in ("h", "restRho", "epsilon", "k", "n", "deltaq", "radius", "c", "dt",
                                        "iterations", "vorticityConfinementEpilon")], np.float32)
    box = np.ascontiguousarray(box, np.float32)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    out = {k: np.empty((n, 3), np.float32) for k in ("pos", "vel_loop", "curl", "visc", "vel")}
    out["lambda"] = np.empty(n, np.float32)
    ref().ref_step_probe(fp(prm), fp(box), fp(pos), fp(vel), n, fp(out["pos"]), fp(out["vel_loop"]), fp(out["curl"]),
                         fp(out["visc"]), fp(out["vel"]), fp(out["lambda"]))
    return out


def dp_(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def mlp_forward(params, x, bf16=True):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 16)
    y = np.empty(x.shape[0], np.float64)
    orc().mlp_forward(fp(params), fp(x), x.shape[0], int(bf16), dp_(y))
    return y


def mlp_loss_grad(params, x, t, bf16=True):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 16)
    t = np.ascontiguousarray(t, np.float32)
    g = np.empty(orc().mlp_param_count(), np.float64)
    loss = orc().mlp_loss_grad(fp(params), fp(x), fp(t), x.shape[0], int(bf16), dp_(g))
    return loss, g


def mlp_adam(params, m, v, grad, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    params = np.ascontiguousarray(params, np.float32).copy()
    m = np.ascontiguousarray(m, np.float32).copy()
    v = np.ascontiguousarray(v, np.float32).copy()
    grad = np.ascontiguousarray(grad, np.float32)
    orc().mlp_adam(fp(params), fp(m), fp(v), fp(grad), params.size, int(t), lr, b1, b2, eps)
    return params, m, v


MLP_SLICES = {
    "W1": slice(0, 1024),
    "b1": slice(1024, 1088),
    "W2": slice(1088, 5184),
    "b2": slice(5184, 5248),
    "W3": slice(5248, 9344),
    "b3": slice(9344, 9408),
    "W4": slice(9408, 9472),
    "b4": slice(9472, 9473),
}


def density_stats(err: np.ndarray):
    e = err.astype(np.float64)
    return float(np.mean(np.abs(e))), float(np.mean(np.maximum(e, 0.0))), float(np.max(e)) if e.size else 0.0
