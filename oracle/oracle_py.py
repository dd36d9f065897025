"""ctypes binding of the CPU oracle (oracle/lib/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# BP_ORACLE_LIB selects another build of the same source, such as the
# sanitized one of `make -C oracle asan`.
LIB_PATH = os.environ.get("BP_ORACLE_LIB") or os.path.join(_HERE, "lib", "liboracle.so")
_lib = None

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)


def load():
    global _lib
    if _lib is None:
        lib = C.CDLL(LIB_PATH)
        sig = {
            "or_create": (C.c_void_p, [C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
            "or_destroy": (None, [C.c_void_p]),
            "or_num_particles": (C.c_int, [C.c_void_p]),
            "or_step": (C.c_int, [C.c_void_p, C.c_int, C.c_float, C.c_int, _F, _F]),
            "or_set_params": (None, [C.c_void_p, C.c_void_p]),
            "or_get_state": (None, [C.c_void_p, _D, _D, _I]),
            "or_set_state": (None, [C.c_void_p, _D, _D, _I]),
            "or_get_rigid": (None, [C.c_void_p, _D, _D]),
            "or_set_rigid": (None, [C.c_void_p, _D]),
            "or_debug_fetch": (None, [C.c_void_p, C.c_int, _D]),
            "or_debug_fetch_int": (None, [C.c_void_p, C.c_int, _I]),
            "or_neighbor_grid": (None, [_F, C.c_int, _F, _I, _I, _I, _I, _I]),
            "or_set_threads": (C.c_int, [C.c_int]),
            "or_cell_address": (C.c_int, [_F, _F, _I]),
            "or_poly6": (C.c_double, [C.c_double, C.c_double, C.c_int]),
            "or_grad_spiky": (None, [_D, C.c_double, C.c_double, C.c_int, _D]),
            "or_akinci_c": (C.c_double, [C.c_double, C.c_double, C.c_int]),
            "or_extract_rotation": (None, [_D, _D, C.c_int, C.c_int]),
            "or_rollout_loss": (C.c_double, [C.c_void_p, C.c_void_p, C.c_int, C.c_float, _D, C.c_double, _D, _D]),
            "or_rollout_tangent": (C.c_double, [C.c_void_p, C.c_void_p, C.c_int, C.c_float, _D, C.c_double, _D,
                                                C.c_int, _D]),
        }
        for k, (r, a) in sig.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _lib = lib
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_D)


def _f(a):
    return None if a is None else a.ctypes.data_as(_F)


def _i(a):
    return None if a is None else a.ctypes.data_as(_I)


SCATTER = 0  # fluidPosition's pair loop as written (reference order)
GATHER = 1   # per-particle sums, the GPU kernel's summation order


class Oracle:
    """FP32 (fp64=False) or FP64 host restatement of ParticleSolver::update."""

    def __init__(self, scene_arrays, params, fp64=False, mode=SCATTER):
        self.lib = load()
        self.arrays = scene_arrays  # keeps the numpy buffers alive
        self.h = self.lib.or_create(C.byref(scene_arrays.desc), C.byref(params), int(fp64), int(mode))
        if not self.h:
            raise RuntimeError("or_create failed")
        self.n = self.lib.or_num_particles(self.h)
        self.num_rigid = len(scene_arrays.rranges)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_destroy(self.h)
            self.h = None

    def set_params(self, params):
        self.lib.or_set_params(self.h, C.byref(params))

    def step(self, substeps, dt, picked=-1, pos=None, vel=None):
        pp = np.asarray(pos if pos is not None else (0, 0, 0), np.float32)
        pv = np.asarray(vel if vel is not None else (0, 0, 0), np.float32)
        self.lib.or_step(self.h, int(substeps), float(dt), int(picked), _f(pp), _f(pv))

    def get_state(self):
        pos = np.empty((self.n, 3))
        vel = np.empty((self.n, 3))
        mp = np.empty(self.n, np.int32)
        self.lib.or_get_state(self.h, _d(pos), _d(vel), _i(mp))
        return pos, vel, mp

    def set_state(self, pos, vel, mp):
        pos = np.ascontiguousarray(pos, np.float64)
        vel = np.ascontiguousarray(vel, np.float64)
        mp = np.ascontiguousarray(mp, np.int32)
        self.lib.or_set_state(self.h, _d(pos), _d(vel), _i(mp))

    def state_original_order(self):
        pos, vel, mp = self.get_state()
        po = np.empty_like(pos)
        vo = np.empty_like(vel)
        po[mp] = pos
        vo[mp] = vel
        return po, vo

    def rigid_state(self):
        q = np.zeros((self.num_rigid, 4))
        cm = np.zeros((self.num_rigid, 3))
        if self.num_rigid:
            self.lib.or_get_rigid(self.h, _d(q), _d(cm))
        return q, cm

    def set_rigid_state(self, q):
        q = np.ascontiguousarray(q, np.float64)
        self.lib.or_set_rigid(self.h, _d(q))

    def debug(self, which):
        if which in (2, 3):
            out = np.empty(self.n, np.int32)
            self.lib.or_debug_fetch_int(self.h, int(which), _i(out))
            return out
        out = np.empty(self.n if which in (0, 1) else (self.n, 3))
        self.lib.or_debug_fetch(self.h, int(which), _d(out))
        return out


def set_threads(n=0):
    """OpenMP threads of the oracle's per-particle loops; n <= 0 only queries."""
    return int(load().or_set_threads(int(n)))


def neighbor_grid(xs, cell_size, grid_dims):
    lib = load()
    xs = np.ascontiguousarray(xs, np.float32)
    n = len(xs)
    cs = np.asarray(cell_size, np.float32)
    dims = np.asarray(grid_dims, np.int32)
    cells = int(np.prod(dims))
    keys = np.empty(n, np.int32)
    perm = np.empty(n, np.int32)
    st = np.empty(cells, np.int32)
    en = np.empty(cells, np.int32)
    lib.or_neighbor_grid(_f(xs), n, _f(cs), _i(dims), _i(keys), _i(perm), _i(st), _i(en))
    return keys, perm, st, en


def cell_address(p, cell_size, grid_dims):
    return load().or_cell_address(_f(np.asarray(p, np.float32)), _f(np.asarray(cell_size, np.float32)),
                                  _i(np.asarray(grid_dims, np.int32)))


def poly6(r2, h, fp64=True):
    return load().or_poly6(float(r2), float(h), int(fp64))


def grad_spiky(v, r2, h, fp64=True):
    out = np.zeros(3)
    load().or_grad_spiky(_d(np.asarray(v, np.float64)), float(r2), float(h), int(fp64), _d(out))
    return out


def akinci_c(r, h, fp64=True):
    return load().or_akinci_c(float(r), float(h), int(fp64))


def extract_rotation(A_cols, q, iterations=20, fp64=True):
    A = np.ascontiguousarray(np.asarray(A_cols, np.float64).reshape(9))
    qq = np.ascontiguousarray(np.asarray(q, np.float64).copy())
    load().or_extract_rotation(_d(A), _d(qq), int(iterations), int(fp64))
    return qq


def rollout_loss(scene_arrays, params, substeps, subdt, target, norm, theta=None, v0=None):
    """FP64 loss sum |x_T - target|^2 / norm of a reference-order rollout;
    theta (6) / v0 (N x 3, original order) override the inputs."""
    t = np.ascontiguousarray(target, np.float64)
    th = None if theta is None else np.ascontiguousarray(theta, np.float64)
    v = None if v0 is None else np.ascontiguousarray(v0, np.float64)
    return load().or_rollout_loss(C.byref(scene_arrays.desc), C.byref(params), int(substeps), float(subdt), _d(t),
                                  float(norm), _d(th), _d(v))


def rollout_tangent(scene_arrays, params, substeps, subdt, target, norm, vdirs=()):
    """FP64 forward-mode derivative of rollout_loss: returns (L, dL) with
    dL[0:6] = dL/d(wind xyz, drag, lift, xsph_c) and dL[6 + k] = dL/dv0 . vdirs[k]."""
    t = np.ascontiguousarray(target, np.float64)
    nd = len(vdirs)
    vd = np.ascontiguousarray(np.stack(vdirs), np.float64) if nd else None
    out = np.zeros(8, np.float64)
    L = load().or_rollout_tangent(C.byref(scene_arrays.desc), C.byref(params), int(substeps), float(subdt), _d(t),
                                  float(norm), _d(vd), nd, _d(out))
    return L, out[:6 + nd]
