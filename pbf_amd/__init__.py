"""pbf_amd -- Position Based Fluids on MI355X (gfx950), host side.

Python mirror of the reference's solver host API (kernels/cudaPBF.h:8-12) over the C ABI of
``libpbf_hip.so`` (include/pbf.h):

=====================================  ==========================================
reference                              here
=====================================  ==========================================
``SolverParameter`` solverBase.h:16    :class:`SolverParameter` (same 13 leading fields)
``SolverData`` solverBase.h:45         :class:`SolverData` (owns one ``pbf_solver``)
``initCUDASolverDataBuffer`` cu:4      :func:`initCUDASolverDataBuffer`
``cudaPBFSolve`` cu:50                 :func:`cudaPBFSolve`
``freeCUDASolverDataBuffer`` cu:32     :func:`freeCUDASolverDataBuffer`
``PBFSolver`` solver.h:83              :class:`PBFSolver` (``solve(particles)`` in place)
=====================================  ==========================================

:class:`Solver` is the direct handle API used by the tests and the bench.  Every call goes to
the HIP library; there is no CPU fallback, and importing this package does not need a GPU,
but creating a solver does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpbf_hip.so")

PBF_OK = 0
PBF_ERR_ARG = -1
PBF_ERR_HIP = -2
PBF_ERR_CAPACITY = -3
PBF_ERR_STATE = -4
PBF_ERR_NOGPU = -5

FLAG_DEBUG = 1
FLAG_DENSITY_STATS = 2
FLAG_GRAPH = 4
FLAG_FAST_MATH = 8


class PBFError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pbf error {code}: {msg}")
        self.code = code


class SolverParameter(C.Structure):
    """``pbf_params``: the reference's SolverParameter (include/solverBase.h:16-40) followed by
    the constants the reference hard-codes (gravity, jitter coefficient, grid dims, neighbour
    cap) and flags."""

    _fields_ = [
        ("h", C.c_float),
        ("nParticles", C.c_int32),
        ("restRho", C.c_float),
        ("epsilon", C.c_float),
        ("k", C.c_float),
        ("n", C.c_float),
        ("deltaq", C.c_float),
        ("c", C.c_float),
        ("dt", C.c_float),
        ("invDt", C.c_float),
        ("iterations", C.c_int32),
        ("vorticityConfinementEpilon", C.c_float),
        ("radius", C.c_float),
        ("gravity", C.c_float * 3),
        ("jitter_coeff", C.c_float),
        ("grad_rmin", C.c_float),
        ("grid_dims", C.c_int32 * 3),
        ("max_neighbors", C.c_int32),
        ("flags", C.c_uint32),
    ]

    @classmethod
    def from_dict(cls, p: dict, n: int, grid_dims, flags: int = 0) -> "SolverParameter":
        s = cls()
        s.h = p["h"]
        s.nParticles = n
        s.restRho = p["restRho"]
        s.epsilon = p["epsilon"]
        s.k = p["k"]
        s.n = p["n"]
        s.deltaq = p["deltaq"]
        s.c = p["c"]
        s.dt = np.float32(p["dt"])
        s.invDt = np.float32(1.0) / np.float32(p["dt"])
        s.iterations = int(p["iterations"])
        s.vorticityConfinementEpilon = p["vorticityConfinementEpilon"]
        s.radius = p["radius"]
        for a in range(3):
            s.gravity[a] = p["gravity"][a]
            s.grid_dims[a] = int(grid_dims[a])
        s.jitter_coeff = p["jitter_coeff"]
        s.grad_rmin = p.get("grad_rmin", 0.01)
        s.max_neighbors = int(p.get("max_neighbors", 63))
        s.flags = int(flags)
        return s

    def copy(self) -> "SolverParameter":
        o = SolverParameter()
        C.memmove(C.byref(o), C.byref(self), C.sizeof(self))
        return o


class DeviceView(C.Structure):
    _fields_ = [
        ("pos4", C.c_void_p),
        ("vel4", C.c_void_p),
        ("ids", C.c_void_p),
        ("n", C.c_int32),
        ("ncells", C.c_int32),
    ]


_lib_handle = None

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)
_VP = C.c_void_p


def lib() -> C.CDLL:
    """The loaded HIP library; raises if it has not been built."""
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m pbf_amd.build`")
    L = C.CDLL(LIB_PATH)
    sig = {
        "pbf_create": (C.c_int, [C.POINTER(SolverParameter), _F, _F, _F, _F, C.c_int32, C.POINTER(_VP)]),
        "pbf_solve": (C.c_int, [_VP, C.POINTER(SolverParameter)]),
        "pbf_step": (C.c_int, [_VP, C.c_float]),
        "pbf_advance": (C.c_int, [_VP, C.c_int32]),
        "pbf_get_state": (C.c_int, [_VP, _F, _F]),
        "pbf_set_state": (C.c_int, [_VP, _F, _F]),
        "pbf_export_device": (C.c_int, [_VP, _VP, _VP]),
        "pbf_destroy": (C.c_int, [_VP]),
        "pbf_last_error": (C.c_char_p, []),
        "pbf_get_params": (C.c_int, [_VP, C.POINTER(SolverParameter)]),
        "pbf_synchronize": (C.c_int, [_VP]),
        "pbf_device_state": (C.c_int, [_VP, C.POINTER(DeviceView)]),
        "pbf_get_grid": (C.c_int, [_VP, _I, _I, _I, _I]),
        "pbf_get_neighbors": (C.c_int, [_VP, _I, _I]),
        "pbf_get_fields": (C.c_int, [_VP, _F, _F, _F, _F]),
        "pbf_density_stats": (C.c_int, [_VP, _F]),
        "pbf_time_steps": (C.c_int, [_VP, C.c_int32, _F]),
        "pbf_stage_times": (C.c_int, [_VP, _F]),
        "pbf_version": (C.c_char_p, []),
        "pbf_profile_step": (C.c_int, [_VP, _F, C.POINTER(C.c_char_p), C.c_int32]),
        "pbf_mlp_param_count": (C.c_int, []),
        "pbf_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(_VP)]),
        "pbf_mlp_destroy": (C.c_int, [_VP]),
        "pbf_mlp_get_params": (C.c_int, [_VP, _F]),
        "pbf_mlp_set_params": (C.c_int, [_VP, _F]),
        "pbf_mlp_set_normalization": (C.c_int, [_VP, _F, _F, C.c_float, C.c_float]),
        "pbf_mlp_forward": (C.c_int, [_VP, _F, C.c_int32, _F]),
        "pbf_mlp_loss_grad": (C.c_int, [_VP, _F, _F, C.c_int32, _F, _F]),
        "pbf_mlp_loss_grad_noalloc": (C.c_int, [_VP, _VP, _VP, C.c_int32, _VP]),
        "pbf_mlp_adam_step": (C.c_int, [_VP, _F, C.c_float, C.c_float, C.c_float, C.c_float]),
        "pbf_mlp_train": (C.c_int, [_VP, _VP, _VP, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_uint64, _F]),
        "pbf_attach_mlp": (C.c_int, [_VP, _VP]),
        "pbf_record_step": (C.c_int, [_VP, C.c_int32, _VP, _VP, C.c_int64, C.c_int64, C.c_int32, C.c_int32]),
        "pbf_step_adaptive": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_float, C.c_int32, C.c_int32, _I, _F]),
        "pbf_last_guess_info": (C.c_int, [_VP, _F]),
        "pbf_set_learned": (C.c_int, [_VP, C.c_int32]),
        "pbf_time_stage": (C.c_int, [_VP, C.c_int32, C.c_int32, _F]),
        "pbf_snapshot": (C.c_int, [_VP, C.c_int32]),
        "pbf_create_batch": (C.c_int, [C.POINTER(SolverParameter), C.c_int32, _I, _F, _F, _I, _F, _F, C.POINTER(_VP)]),
        "pbf_set_tuning": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32]),
        "pbf_get_tuning": (C.c_int, [_VP, _I]),
        "pbf_autotune": (C.c_int, [_VP, C.c_int32, _I, _F]),
        "pbf_set_ordering": (C.c_int, [_VP, C.c_int32]),
        "pbf_get_ordering": (C.c_int, [_VP, _I]),
        "pbf_device_alloc": (C.c_int, [C.c_uint64, C.POINTER(_VP)]),
        "pbf_device_free": (C.c_int, [_VP]),
        "pbf_memcpy_to_host": (C.c_int, [_VP, _VP, C.c_uint64]),
        "pbf_memcpy_to_device": (C.c_int, [_VP, _VP, C.c_uint64]),
        "pbf_copy_kernel": (C.c_int, [_VP, _VP, C.c_uint64, _F]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib_handle = L
    return L


def _check(rc: int) -> int:
    if rc < 0:
        raise PBFError(rc, lib().pbf_last_error().decode())
    return rc


def _fp(a):
    return None if a is None else a.ctypes.data_as(_F)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_I)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Solver:
    """One GPU solver (a ``pbf_solver`` handle)."""

    def __init__(self, para: SolverParameter, grid_space, boundary, pos, vel=None):
        pos = _f32(pos, (-1, 3))
        n = pos.shape[0]
        vel = None if vel is None else _f32(vel, (n, 3))
        gs = _f32(grid_space, (6,))
        bd = _f32(boundary, (6,))
        self.para = para.copy()
        self.para.nParticles = n
        self.n = n
        self.ncells = int(para.grid_dims[0] * para.grid_dims[1] * para.grid_dims[2])
        h = _VP()
        _check(lib().pbf_create(C.byref(self.para), _fp(gs), _fp(bd), _fp(pos), _fp(vel), n, C.byref(h)))
        self._h = h

    @classmethod
    def batch(cls, scene_list, flags: int = 0, **over) -> "Solver":
        """One solver advancing every scene of scene_list in each launch (pbf_create_batch).

        Particles are the scenes' particles one scene after another; parameters are the first
        scene's (with ``over``)."""
        p = dict(scene_list[0].params)
        p.update(over)
        counts = np.array([sc.n for sc in scene_list], np.int32)
        n = int(counts.sum())
        para = SolverParameter.from_dict(p, n, scene_list[0].grid_dims, flags)
        gs = np.ascontiguousarray(np.concatenate([sc.box for sc in scene_list]), np.float32)
        bd = np.ascontiguousarray(np.concatenate([sc.boundary for sc in scene_list]), np.float32)
        dims = np.ascontiguousarray(np.concatenate([np.array(sc.grid_dims, np.int32) for sc in scene_list]), np.int32)
        pos = _f32(np.concatenate([sc.pos for sc in scene_list]), (n, 3))
        vel = _f32(np.concatenate([sc.vel for sc in scene_list]), (n, 3))
        self = cls.__new__(cls)
        self.para = para
        self.n = n
        self.ncells = int(sum(int(np.prod(sc.grid_dims)) for sc in scene_list))
        self.counts = counts
        h = _VP()
        _check(lib().pbf_create_batch(C.byref(self.para), len(scene_list), _ip(counts), _fp(gs), _fp(bd), _ip(dims), _fp(pos), _fp(vel), C.byref(h)))
        self._h = h
        return self

    def set_tuning(self, block: int = 256, grid: int = 0, xcd_map: bool = True):
        _check(lib().pbf_set_tuning(self._h, int(block), int(grid), int(bool(xcd_map))))

    def get_tuning(self):
        out = np.zeros(3, np.int32)
        _check(lib().pbf_get_tuning(self._h, _ip(out)))
        return int(out[0]), int(out[1]), bool(out[2])

    def set_ordering(self, order: int):
        """Cell-row order in memory: 0 = z, then y; 1 = Morton curve over (y, z).  Results do not
        depend on it."""
        _check(lib().pbf_set_ordering(self._h, int(order)))

    def get_ordering(self) -> int:
        out = np.zeros(1, np.int32)
        _check(lib().pbf_get_ordering(self._h, _ip(out)))
        return int(out[0])

    def autotune(self, steps: int = 10):
        """Picks the fastest launch shape, then the faster row order for it (state unchanged);
        returns ((block, grid, xcd_map), ms/step); the order is get_ordering()'s."""
        best = np.zeros(3, np.int32)
        ms = np.zeros(1, np.float32)
        _check(lib().pbf_autotune(self._h, int(steps), _ip(best), _fp(ms)))
        return (int(best[0]), int(best[1]), bool(best[2])), float(ms[0])

    @classmethod
    def from_scene(cls, scene, flags: int = 0, **over) -> "Solver":
        p = dict(scene.params)
        p.update(over)
        para = SolverParameter.from_dict(p, scene.n, scene.grid_dims, flags)
        return cls(para, scene.box, scene.boundary, scene.pos, scene.vel)

    # lifecycle --------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().pbf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # stepping ---------------------------------------------------------------------------
    def solve(self, para: SolverParameter | None = None):
        """One step with ``para`` (the solver's own when None), as cudaPBFSolve."""
        if para is not None:
            self.para = para.copy()
        _check(lib().pbf_solve(self._h, C.byref(self.para)))

    def step(self, dt: float | None = None):
        _check(lib().pbf_step(self._h, float(self.para.dt if dt is None else dt)))

    def advance(self, nsteps: int):
        _check(lib().pbf_advance(self._h, int(nsteps)))

    def synchronize(self):
        _check(lib().pbf_synchronize(self._h))

    # state ------------------------------------------------------------------------------
    def get_state(self):
        pos = np.empty((self.n, 3), np.float32)
        vel = np.empty((self.n, 3), np.float32)
        _check(lib().pbf_get_state(self._h, _fp(pos), _fp(vel)))
        return pos, vel

    def set_state(self, pos, vel=None):
        pos = _f32(pos, (self.n, 3))
        vel = None if vel is None else _f32(vel, (self.n, 3))
        _check(lib().pbf_set_state(self._h, _fp(pos), _fp(vel)))

    def export_device(self, dev_pos: int, dev_vel: int | None = None):
        _check(lib().pbf_export_device(self._h, dev_pos, dev_vel))

    def device_view(self) -> DeviceView:
        v = DeviceView()
        _check(lib().pbf_device_state(self._h, C.byref(v)))
        return v

    # diagnostics ------------------------------------------------------------------------
    def grid(self):
        """(gid[n], count[ncells], cell_start[ncells+1], sorted_ids[n], n_in_grid) of the last step."""
        gid = np.empty(self.n, np.int32)
        cnt = np.empty(self.ncells, np.int32)
        cs = np.empty(self.ncells + 1, np.int32)
        srt = np.empty(self.n, np.int32)
        nin = _check(lib().pbf_get_grid(self._h, _ip(gid), _ip(cnt), _ip(cs), _ip(srt)))
        return gid, cnt, cs, srt, nin

    def neighbors(self):
        cap = int(self.para.max_neighbors)
        nbr = np.empty((self.n, cap), np.int32)
        cnt = np.empty(self.n, np.int32)
        _check(lib().pbf_get_neighbors(self._h, _ip(nbr), _ip(cnt)))
        return nbr, cnt

    def fields(self):
        """(lambda_first, lambda_last, density_error, curl) of the last step, by particle id."""
        lf = np.empty(self.n, np.float32)
        ll = np.empty(self.n, np.float32)
        er = np.empty(self.n, np.float32)
        cu = np.empty((self.n, 3), np.float32)
        _check(lib().pbf_get_fields(self._h, _fp(lf), _fp(ll), _fp(er), _fp(cu)))
        return lf, ll, er, cu

    def density_stats(self):
        out = np.empty(3, np.float32)
        _check(lib().pbf_density_stats(self._h, _fp(out)))
        return float(out[0]), float(out[1]), float(out[2])

    # measurement ------------------------------------------------------------------------
    def time_steps(self, nsteps: int) -> float:
        ms = np.zeros(1, np.float32)
        _check(lib().pbf_time_steps(self._h, int(nsteps), _fp(ms)))
        return float(ms[0])

    STAGES = {"neighbors_lambda": 0, "lambda": 1, "delta_apply": 2, "sort_reorder": 3, "launch_floor": 4,
              "valu_probe": 5, "vort_hash_sort": 6, "vort_apply": 7, "curl": 8}
    VALU_PROBE_INSTR = 2048 * 4 * 16 * 4096  # wave64 VALU instructions of one probe launch (loop body)

    def time_stage(self, stage: str, reps: int = 50) -> float:
        """ms per launch of one step stage ("neighbors_lambda", "lambda", "delta_apply",
        "sort_reorder" = predict + hash, scan, scatter + rank + reorder, "launch_floor" = an
        empty kernel of the particle grid's shape, "valu_probe" = VALU_PROBE_INSTR multiply and
        add instructions, "vort_hash_sort" = the vorticity pass with the next step's hash, scan,
        scatter + rank + reorder, "vort_apply" = the vorticity pass alone), replayed back to back
        on the last step's state (state unchanged)."""
        ms = np.zeros(1, np.float32)
        _check(lib().pbf_time_stage(self._h, self.STAGES[stage], int(reps), _fp(ms)))
        return float(ms[0])

    def profile_step(self):
        """[(stage name, ms)] of one eagerly launched step, HIP events between launch groups."""
        ms = np.zeros(64, np.float32)
        names = (C.c_char_p * 64)()
        cnt = _check(lib().pbf_profile_step(self._h, _fp(ms), names, 64))
        return [(names[i].decode(), float(ms[i])) for i in range(cnt)]

    # warm start / adaptive iteration ------------------------------------------------------
    def attach_mlp(self, mlp: "Mlp | None"):
        _check(lib().pbf_attach_mlp(self._h, None if mlp is None else mlp._h))
        self._mlp = mlp

    def step_adaptive(self, use_mlp: bool, tol: float, min_iters: int, max_iters: int, metric: str = "compression"):
        """One step stopping at residual <= tol (metric "compression": mean max(C,0); "abs":
        mean |C|); tol < 0 runs exactly max_iters passes with no read-back.  Returns (passes,
        last residual read, -1 if none)."""
        it = np.zeros(1, np.int32)
        r = np.zeros(1, np.float32)
        m = {"compression": 0, "abs": 1}[metric]
        _check(lib().pbf_step_adaptive(self._h, int(bool(use_mlp)), m, float(tol), int(min_iters), int(max_iters), _ip(it), _fp(r)))
        return int(it[0]), float(r[0])

    def set_learned(self, passes: int):
        """passes >= 2: following steps run a plain pass, the learned pass and passes - 2 plain
        iterations (device-side keep / reject, graph-captured); 0: plain steps again."""
        _check(lib().pbf_set_learned(self._h, int(passes)))

    def snapshot(self):
        """Save the state in a device-side snapshot."""
        _check(lib().pbf_snapshot(self._h, 1))

    def restore(self):
        """Restore the state saved by snapshot()."""
        _check(lib().pbf_snapshot(self._h, 0))

    def last_guess_info(self):
        """(compression after the plain pass 1 / after the network's pass, pass kept, mean |C|
        before / after) of the last learned pass."""
        out = np.zeros(5, np.float32)
        _check(lib().pbf_last_guess_info(self._h, _fp(out)))
        return float(out[0]), float(out[1]), bool(out[2]), float(out[3]), float(out[4])

    def record_step(self, iterations: int, ds: "DeviceDataset", base: int, stride: int, phase: int):
        _check(lib().pbf_record_step(self._h, int(iterations), ds.x, ds.t, int(base), int(ds.cap), int(stride), int(phase)))

    def stage_times(self):
        out = np.zeros(7, np.float32)
        _check(lib().pbf_stage_times(self._h, _fp(out)))
        return out


class DeviceBuffer:
    """Device allocation owned from Python (datasets, export targets)."""

    def __init__(self, nbytes: int):
        h = _VP()
        _check(lib().pbf_device_alloc(int(nbytes), C.byref(h)))
        self.ptr = h.value
        self.nbytes = int(nbytes)

    def to_host(self, dtype, count):
        out = np.empty(count, dtype)
        _check(lib().pbf_memcpy_to_host(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def from_host(self, arr):
        arr = np.ascontiguousarray(arr)
        _check(lib().pbf_memcpy_to_device(self.ptr, arr.ctypes.data, arr.nbytes))

    def close(self):
        if getattr(self, "ptr", None):
            lib().pbf_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceDataset:
    """(features [cap][16], targets [cap]) in device memory, filled by Solver.record_step."""

    def __init__(self, cap: int):
        self.cap = int(cap)
        self._x = DeviceBuffer(self.cap * 16 * 4)
        self._t = DeviceBuffer(self.cap * 4)
        self.x = self._x.ptr
        self.t = self._t.ptr

    def host(self, count=None):
        count = self.cap if count is None else int(count)
        x = self._x.to_host(np.float32, self.cap * 16).reshape(self.cap, 16)[:count]
        t = self._t.to_host(np.float32, self.cap)[:count]
        return x, t

    def upload(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(t, np.float32).reshape(-1)
        assert x.shape[0] <= self.cap and t.shape[0] == x.shape[0]
        self._x.from_host(x)
        self._t.from_host(t)

    def close(self):
        self._x.close()
        self._t.close()


class Mlp:
    """The warm-start network (pbf_mlp): 16 -> 64 -> 64 -> 64 -> 1, BF16 MFMA kernels."""

    def __init__(self, seed: int = 0):
        h = _VP()
        _check(lib().pbf_mlp_create(int(seed), C.byref(h)))
        self._h = h
        self.nparams = lib().pbf_mlp_param_count()

    def close(self):
        if getattr(self, "_h", None):
            lib().pbf_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.empty(self.nparams, np.float32)
        _check(lib().pbf_mlp_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p):
        p = _f32(p, (self.nparams,))
        _check(lib().pbf_mlp_set_params(self._h, _fp(p)))

    def set_normalization(self, mean, inv_std, yscale=1.0, yshift=0.0):
        _check(lib().pbf_mlp_set_normalization(self._h, _fp(_f32(mean, (16,))), _fp(_f32(inv_std, (16,))), float(yscale), float(yshift)))

    def forward(self, x):
        x = _f32(x, (-1, 16))
        y = np.empty(x.shape[0], np.float32)
        _check(lib().pbf_mlp_forward(self._h, _fp(x), x.shape[0], _fp(y)))
        return y

    def loss_grad(self, x, t):
        x = _f32(x, (-1, 16))
        t = _f32(t, (x.shape[0],))
        g = np.empty(self.nparams, np.float32)
        loss = np.zeros(1, np.float32)
        _check(lib().pbf_mlp_loss_grad(self._h, _fp(x), _fp(t), x.shape[0], _fp(g), _fp(loss)))
        return float(loss[0]), g

    def loss_grad_noalloc(self, dev_x: int, dev_t: int, n: int, dev_grad: int):
        """pbf_mlp_loss_grad_noalloc: device arrays in and out (dev_grad: nparams + 1 floats, the
        last the sum of squared errors); allocates nothing."""
        _check(lib().pbf_mlp_loss_grad_noalloc(self._h, dev_x, dev_t, int(n), dev_grad))

    def adam_step(self, grad=None, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        g = None if grad is None else _f32(grad, (self.nparams,))
        _check(lib().pbf_mlp_adam_step(self._h, _fp(g), float(lr), float(beta1), float(beta2), float(eps)))

    def train(self, ds: DeviceDataset, nsamples: int, batch: int, steps: int, lr=2e-3, lr_final=1e-4, seed=0):
        losses = np.zeros(max(steps, 1), np.float32)
        _check(lib().pbf_mlp_train(self._h, ds.x, ds.t, int(nsamples), int(batch), int(steps), float(lr), float(lr_final), int(seed), _fp(losses)))
        return losses[:steps]


# ---- the reference's host API names (kernels/cudaPBF.h:8-12) ------------------------------
class SolverData:
    """Holds the device solver that initCUDASolverDataBuffer creates (solverBase.h:45-64)."""

    def __init__(self):
        self.solver: Solver | None = None


def initCUDASolverDataBuffer(position, velocity, data: SolverData, para: SolverParameter, gridSpace, boundary):
    """kernels/cudaPBF.cu:4-30: allocate the device state and upload x, v."""
    pos = _f32(position, (-1, 3))
    if int(para.nParticles) != pos.shape[0]:
        raise PBFError(PBF_ERR_ARG, "para.nParticles differs from len(position)")
    data.solver = Solver(para, gridSpace, boundary, pos, velocity)


def cudaPBFSolve(data: SolverData, para: SolverParameter):
    """kernels/cudaPBF.cu:50-114: one step with para (dt, iterations, ...)."""
    if data.solver is None:
        raise PBFError(PBF_ERR_STATE, "solver data not initialised")
    data.solver.solve(para)


def freeCUDASolverDataBuffer(data: SolverData):
    """kernels/cudaPBF.cu:32-48."""
    if data.solver is not None:
        data.solver.close()
        data.solver = None


init_solver_data_buffer = initCUDASolverDataBuffer
pbf_solve = cudaPBFSolve
free_solver_data_buffer = freeCUDASolverDataBuffer


class PBFSolver:
    """The reference's CPU-side solver shape (include/solver.h:83-111) backed by the GPU solver.

    ``PBFSolver(para, box)`` -- the clamp boundary is ``box`` shrunk by ``para.radius``
    (solver.h:97) and the grid is anchored at ``box.pMin`` (solver.h:109).  ``solve(particles)``
    advances one step in place on a structured array with ``position`` and ``velocity``
    fields (include/particle.h:9-15).  ``fext`` is accepted as in solver.h:86-88; like the
    reference (solver.h:201) the step does not call it.
    """

    PARTICLE_DTYPE = np.dtype([("position", np.float32, 3), ("velocity", np.float32, 3)])

    def __init__(self, para: SolverParameter, box, fext=None):
        self.para = para.copy()
        self.box = _f32(box, (6,))
        r = np.float32(para.radius)
        self.boundary = np.concatenate([self.box[:3] + r, self.box[3:] - r]).astype(np.float32)
        self.fext = fext
        self._solver: Solver | None = None

    def solve(self, particles: np.ndarray):
        pos = np.ascontiguousarray(particles["position"], np.float32)
        vel = np.ascontiguousarray(particles["velocity"], np.float32)
        if self._solver is None or self._solver.n != pos.shape[0]:
            if self._solver is not None:
                self._solver.close()
            self.para.nParticles = pos.shape[0]
            self._solver = Solver(self.para, self.box, self.boundary, pos, vel)
        else:
            self._solver.set_state(pos, vel)
        self._solver.solve(self.para)
        p, v = self._solver.get_state()
        particles["position"] = p
        particles["velocity"] = v

    def close(self):
        if self._solver is not None:
            self._solver.close()
            self._solver = None
