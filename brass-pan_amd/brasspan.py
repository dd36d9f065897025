"""ctypes binding of the gfx950 particle solver C ABI (include/brasspan.h).

Python mirror of brass-pan's ``ParticleSolver`` (brasstower/particlesolver.cuh)
for tests, benchmarks and the gradient harness.  The product is the native
library ``lib/libbrasspan.so``; this module only marshals arguments.  It never
falls back to a CPU path: if the library or a gfx950 device is missing, calls
raise ``BrassPanError``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbrasspan.so")

BP_OK = 0
STATUS_NAMES = {0: "BP_OK", 1: "BP_ERR_INVALID_ARG", 2: "BP_ERR_CAPACITY", 3: "BP_ERR_HIP",
                4: "BP_ERR_NOT_FINITE", 5: "BP_ERR_NO_DEVICE", 6: "BP_ERR_UNSUPPORTED"}


class BrassPanError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class bp_params(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("stabilization_passes", C.c_int), ("gravity", C.c_float * 3),
        ("fluid_rest_density", C.c_float), ("solid_density_scale", C.c_float), ("lambda_mode", C.c_int),
        ("lambda_denominator", C.c_float), ("lambda_relaxation", C.c_float), ("akinci", C.c_int),
        ("scorr_k", C.c_float), ("scorr_n", C.c_int), ("scorr_dq", C.c_float), ("delta_clamp", C.c_float),
        ("vorticity_eps", C.c_float), ("tension_gamma", C.c_float), ("xsph_c", C.c_float),
        ("friction_static", C.c_float), ("friction_dynamic", C.c_float), ("mass_scaling_k", C.c_float),
        ("sleep_epsilon", C.c_float), ("wind_enabled", C.c_int), ("wind", C.c_float * 3),
        ("drag_coeff", C.c_float), ("lift_coeff", C.c_float), ("bend_stiffness", C.c_float),
        ("shape_matching_iterations", C.c_int), ("nan_guard", C.c_int),
    ]


class bp_scene_desc(C.Structure):
    _fields_ = [
        ("num_particles", C.c_int), ("positions", C.POINTER(C.c_float)), ("velocities", C.POINTER(C.c_float)),
        ("masses", C.POINTER(C.c_float)), ("phases", C.POINTER(C.c_int)), ("group_ids", C.POINTER(C.c_int)),
        ("particle_radius", C.c_float), ("kernel_radius", C.c_float), ("cell_size", C.c_float * 3),
        ("grid_dims", C.c_int * 3), ("num_planes", C.c_int), ("plane_origins", C.POINTER(C.c_float)),
        ("plane_normals", C.POINTER(C.c_float)), ("num_distance", C.c_int),
        ("distance_pairs", C.POINTER(C.c_int)), ("distance_params", C.POINTER(C.c_float)),
        ("num_bendings", C.c_int), ("bendings", C.POINTER(C.c_int)), ("num_faces", C.c_int),
        ("faces", C.POINTER(C.c_int)), ("num_rigid", C.c_int), ("rigid_ranges", C.POINTER(C.c_int)),
        ("rigid_rest_len", C.c_int), ("rigid_rest_positions", C.POINTER(C.c_float)),
        ("num_immovables", C.c_int), ("immovables", C.POINTER(C.c_int)),
    ]


class bp_device_view(C.Structure):
    _fields_ = [
        ("num_particles", C.c_int), ("new_positions", C.c_void_p), ("positions", C.c_void_p),
        ("velocities", C.c_void_p), ("phases", C.c_void_p), ("map_to_original", C.c_void_p),
        ("map_to_new", C.c_void_p), ("rigid_rotations", C.c_void_p), ("rigid_cms", C.c_void_p),
    ]


# every entry point include/brasspan.h declares: name -> (restype, argtypes)
_S = C.c_int
_P = C.c_void_p
SIGNATURES = {
    "bp_abi_version": (C.c_int, []),
    "bp_last_error_string": (C.c_char_p, []),
    "bp_default_params": (_S, [C.POINTER(bp_params)]),
    "bp_solver_create": (_S, [C.POINTER(bp_scene_desc), C.POINTER(bp_params), _P, C.POINTER(_P)]),
    "bp_solver_destroy": (_S, [_P]),
    "bp_solver_step": (_S, [_P, C.c_int, C.c_float, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "bp_solver_sync": (_S, [_P]),
    "bp_solver_set_params": (_S, [_P, C.POINTER(bp_params)]),
    "bp_query_original_id": (_S, [_P, C.c_int, C.POINTER(C.c_int)]),
    "bp_query_slot": (_S, [_P, C.c_int, C.POINTER(C.c_int)]),
    "bp_get_particle": (_S, [_P, C.c_int, C.POINTER(C.c_float)]),
    "bp_set_particle": (_S, [_P, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "bp_get_device_arrays": (_S, [_P, C.POINTER(bp_device_view)]),
    "bp_copy_state_original_order": (_S, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "bp_get_state": (_S, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int)]),
    "bp_set_state": (_S, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int)]),
    "bp_get_rigid_state": (_S, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "bp_set_rigid_state": (_S, [_P, C.POINTER(C.c_float)]),
    "bp_debug_fetch": (_S, [_P, C.c_int, _P]),
    "bp_neighbor_grid": (_S, [C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int),
                              C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "bp_neighbor_grid_noalloc": (_S, [_P, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int), _P, _P, _P, _P, _P]),
    "bp_batch_create": (_S, [C.POINTER(bp_scene_desc), C.POINTER(bp_params), C.c_int, _P, C.POINTER(_P)]),
    "bp_batch_scene_range": (_S, [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "bp_solver_set_graph": (_S, [_P, C.c_int]),
    "bp_solver_autotune": (_S, [_P, C.c_float, C.c_int, C.c_char_p, C.c_int]),
    "bp_solver_set_fusion": (_S, [_P, C.c_int]),
    "bp_solver_set_walk": (_S, [_P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bp_solver_set_partner_fetch": (_S, [_P, C.c_int]),
    "bp_solver_set_stage_screen": (_S, [_P, C.c_int, C.c_int]),
    "bp_solver_set_block_sizes": (_S, [_P, C.POINTER(C.c_int)]),
    "bp_stream_copy_bandwidth": (_S, [C.c_size_t, C.c_int, C.POINTER(C.c_float)]),
    "bp_forward_checkpointed": (_S, [_P, C.c_int, C.c_float, C.c_int]),
    "bp_backward": (_S, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                         C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "bp_solver_timed_substeps": (_S, [_P, C.c_int, C.c_float, C.POINTER(C.c_float)]),
    "bp_export_frame": (_S, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "bp_map_particle_infos": (_S, [_P, C.c_void_p]),
    "bp_solver_enable_kernel_timing": (_S, [_P, C.c_int]),
    "bp_solver_kernel_times": (_S, [_P, C.c_int, C.c_char_p, C.POINTER(C.c_float), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]),
}

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libbrasspan.so and binds every declared entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise BrassPanError(6, f"{path} is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def hip_runtime() -> C.CDLL:
    """The HIP runtime libbrasspan.so is bound to, for callers that hand it
    raw device buffers or streams. Opened by its soname after the library is
    loaded, so the dynamic linker returns the copy already in the process
    (torch ships its own with the same soname; opening another file by path
    would start a second runtime that does not share the library's device
    state)."""
    load_library()
    return C.CDLL("libamdhip64.so.7")


def device_name(device: int = 0):
    """The HIP runtime's name for a device (hipDeviceGetName), or None."""
    hip = hip_runtime()
    buf = C.create_string_buffer(256)
    if hip.hipDeviceGetName(buf, 256, int(device)) != 0:
        return None
    return buf.value.decode(errors="replace") or None


def _check(status: int):
    if status != BP_OK:
        msg = load_library().bp_last_error_string().decode(errors="replace")
        raise BrassPanError(status, msg)


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def default_params(**overrides) -> bp_params:
    """bp_default_params (the reference's literals) with keyword overrides."""
    p = bp_params()
    _check(load_library().bp_default_params(C.byref(p)))
    for k, v in overrides.items():
        if k in ("gravity", "wind"):
            arr = getattr(p, k)
            for a in range(3):
                arr[a] = float(v[a])
        else:
            setattr(p, k, v)
    return p


def params_to_dict(p: bp_params) -> dict:
    out = {}
    for name, _ in p._fields_:
        v = getattr(p, name)
        out[name] = list(v) if name in ("gravity", "wind") else v
    return out


class SceneArrays:
    """Owns contiguous numpy arrays backing a bp_scene_desc (see scenes.Scene)."""

    def __init__(self, scene):
        self.n = int(scene.num_particles())
        f32 = lambda a, w: np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1, w)) if len(a) else np.zeros((0, w), np.float32)
        i32 = lambda a, w: np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1, w)) if len(a) else np.zeros((0, w), np.int32)
        self.positions = f32(scene.positions, 3)
        self.velocities = f32(scene.velocities, 3) if scene.velocities is not None else None
        self.masses = np.ascontiguousarray(np.asarray(scene.masses, np.float32))
        self.phases = np.ascontiguousarray(np.asarray(scene.phases, np.int32))
        self.group_ids = np.ascontiguousarray(np.asarray(scene.group_ids, np.int32))
        self.plane_o = f32([p[0] for p in scene.planes], 3)
        self.plane_n = f32([p[1] for p in scene.planes], 3)
        self.dpairs = i32(scene.distance_pairs, 2)
        self.dparams = f32(scene.distance_params, 2)
        self.bends = i32(scene.bending_constraints, 4)
        self.faces = i32(scene.faces, 3)
        self.rranges = i32(scene.rigidbody_particle_id_ranges, 2)
        self.rrest = f32(scene.rigidbody_initial_positions, 3)
        self.immov = np.ascontiguousarray(np.asarray(scene.immovable_constraints, np.int32))
        d = bp_scene_desc()
        d.num_particles = self.n
        d.positions = _fp(self.positions)
        d.velocities = _fp(self.velocities)
        d.masses = _fp(self.masses)
        d.phases = _ip(self.phases)
        d.group_ids = _ip(self.group_ids)
        d.particle_radius = float(scene.particle_radius)
        d.kernel_radius = float(scene.fluid_kernel_radius)
        for a in range(3):
            d.cell_size[a] = float(scene.cell_size[a])
            d.grid_dims[a] = int(scene.grid_dims[a])
        d.num_planes = len(self.plane_o)
        d.plane_origins = _fp(self.plane_o)
        d.plane_normals = _fp(self.plane_n)
        d.num_distance = len(self.dpairs)
        d.distance_pairs = _ip(self.dpairs)
        d.distance_params = _fp(self.dparams)
        d.num_bendings = len(self.bends)
        d.bendings = _ip(self.bends)
        d.num_faces = len(self.faces)
        d.faces = _ip(self.faces)
        d.num_rigid = len(self.rranges)
        d.rigid_ranges = _ip(self.rranges)
        d.rigid_rest_len = len(self.rrest)
        d.rigid_rest_positions = _fp(self.rrest)
        d.num_immovables = len(self.immov)
        d.immovables = _ip(self.immov)
        self.desc = d


class Solver:
    """One solver handle (ParticleSolver, particlesolver.cuh:146) on the null stream."""

    def __init__(self, scene, params: bp_params | None = None, stream: int | None = None):
        self.lib = load_library()
        self.arrays = SceneArrays(scene)
        self.n = self.arrays.n
        self.params = params if params is not None else default_params()
        h = C.c_void_p()
        _check(self.lib.bp_solver_create(C.byref(self.arrays.desc), C.byref(self.params),
                                         C.c_void_p(stream) if stream else None, C.byref(h)))
        self.h = h
        self.num_rigid = len(self.arrays.rranges)

    def set_graph(self, on: bool = True):
        _check(self.lib.bp_solver_set_graph(self.h, 1 if on else 0))

    def set_fusion(self, on: bool = True):
        """Iteration fusion on or off (bp_solver_set_fusion; results are bit-identical)."""
        _check(self.lib.bp_solver_set_fusion(self.h, 1 if on else 0))

    def set_walk(self, lambda_staged=True, position_flat: bool = True, omega_normal_staged=True,
                 vorticity_tension_flat: bool = True):
        """Walk forms of the lambda, position, omega/normal and vorticity/tension
        passes (bp_solver_set_walk; results are bit-identical).  The staged
        arguments take a bool (True: the default LDS region) or the region
        size in records per wave and column offset (128, 256 or 512; 0 = the
        per-lane walk)."""
        _check(self.lib.bp_solver_set_walk(self.h, int(lambda_staged), int(bool(position_flat)),
                                           int(omega_normal_staged), int(bool(vorticity_tension_flat))))

    def set_partner_fetch(self, position_partners: int = 2):
        """Partners the position pass fetches per round trip, 2 or 4
        (bp_solver_set_partner_fetch; results are bit-identical)."""
        _check(self.lib.bp_solver_set_partner_fetch(self.h, int(position_partners)))

    def set_stage_screen(self, lambda_candidates: int = 4, omega_normal_candidates: int = 4):
        """Candidates the staged lambda and omega/normal walks screen per batch,
        4 or 8 (bp_solver_set_stage_screen; results are bit-identical)."""
        _check(self.lib.bp_solver_set_stage_screen(self.h, int(lambda_candidates), int(omega_normal_candidates)))

    def autotune(self, dt: float, substeps_per_trial: int = 2) -> dict:
        import json
        buf = C.create_string_buffer(1024)
        _check(self.lib.bp_solver_autotune(self.h, float(dt), int(substeps_per_trial), buf, 1024))
        return json.loads(buf.value.decode())

    TUNED_PASSES = ("fluid_lambda", "fluid_position", "omega_normal", "vorticity_tension", "xsph",
                    "particle_particle")

    def apply_tuning(self, report: dict):
        """Applies an autotune() report (workgroup sizes, iteration fusion, walk
        forms) without timing anything."""
        sizes = (C.c_int * 6)(*[int(report[k]) for k in self.TUNED_PASSES])
        _check(self.lib.bp_solver_set_block_sizes(self.h, sizes))
        self.set_fusion(bool(report["iteration_fusion"]))
        def region(walk, records):  # the staged region size; reports without one use the default
            if report.get(walk, "global") != "staged":
                return 0
            return int(report.get(records, 1))
        self.set_walk(region("lambda_walk", "lambda_stage_records"), report["position_walk"] == "flat",
                      region("omega_normal_walk", "omega_normal_stage_records"),
                      report["vorticity_tension_walk"] == "flat")
        self.set_partner_fetch(int(report.get("position_partners", 2)))
        self.set_stage_screen(int(report.get("lambda_stage_screen", 8)), int(report.get("omega_normal_stage_screen", 8)))

    # ---------------------------------------------------------- adjoint
    def forward_checkpointed(self, substeps: int, subdt: float, every: int = 8):
        """Differentiable rollout (bp_forward_checkpointed)."""
        _check(self.lib.bp_forward_checkpointed(self.h, int(substeps), float(subdt), int(every)))

    def backward(self, grad_x, grad_v=None):
        """Adjoint of the last rollout (bp_backward).  Returns dL/dx0, dL/dv0
        (N x 3, original order) and dL/dtheta (scenes x 6: wind xyz, drag,
        lift, xsph_c)."""
        gx = None if grad_x is None else np.ascontiguousarray(grad_x, np.float32).reshape(self.n, 3)
        gv = None if grad_v is None else np.ascontiguousarray(grad_v, np.float32).reshape(self.n, 3)
        gx0 = np.zeros((self.n, 3), np.float32)
        gv0 = np.zeros((self.n, 3), np.float32)
        nsc = len(getattr(self, "ranges", [None]))
        gth = np.zeros((nsc, 6), np.float32)
        _check(self.lib.bp_backward(self.h, _fp(gx), _fp(gv), _fp(gx0), _fp(gv0), _fp(gth)))
        return gx0, gv0, gth

    def close(self):
        if getattr(self, "h", None):
            self.lib.bp_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: bp_params):
        self.params = params
        _check(self.lib.bp_solver_set_params(self.h, C.byref(params)))

    def step(self, substeps: int, dt: float, picked: int = -1, pos=None, vel=None):
        pp = np.asarray(pos if pos is not None else (0, 0, 0), np.float32)
        pv = np.asarray(vel if vel is not None else (0, 0, 0), np.float32)
        _check(self.lib.bp_solver_step(self.h, int(substeps), float(dt), int(picked), _fp(pp), _fp(pv)))

    def sync(self):
        _check(self.lib.bp_solver_sync(self.h))

    def get_state(self):
        pos = np.empty((self.n, 3), np.float32)
        vel = np.empty((self.n, 3), np.float32)
        mp = np.empty(self.n, np.int32)
        _check(self.lib.bp_get_state(self.h, _fp(pos), _fp(vel), _ip(mp)))
        return pos, vel, mp

    def set_state(self, pos, vel, map_to_original):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        mp = np.ascontiguousarray(map_to_original, np.int32)
        _check(self.lib.bp_set_state(self.h, _fp(pos), _fp(vel), _ip(mp)))

    def state_original_order(self):
        pos = np.empty((self.n, 3), np.float32)
        vel = np.empty((self.n, 3), np.float32)
        _check(self.lib.bp_copy_state_original_order(self.h, _fp(pos), _fp(vel)))
        return pos, vel

    def rigid_state(self):
        q = np.zeros((self.num_rigid, 4), np.float32)
        cm = np.zeros((self.num_rigid, 3), np.float32)
        if self.num_rigid:
            _check(self.lib.bp_get_rigid_state(self.h, _fp(q), _fp(cm)))
        return q, cm

    def export_frame(self):
        """bp_export_frame: (records [N, 4] float32 in original-id order, x y z
        and the group id's int bits; matrices [bodies, 4, 4], M[b] @ (p, 1)
        maps a body's rest offset p to its position)."""
        rec = np.empty((self.n, 4), np.float32)
        mats = np.zeros((self.num_rigid, 16), np.float32)
        _check(self.lib.bp_export_frame(self.h, _fp(rec), _fp(mats) if self.num_rigid else None))
        return rec, mats.reshape(-1, 4, 4).transpose(0, 2, 1).copy()  # column-major -> M[b][row][col]

    def map_particle_infos(self, device_ptr: int):
        """bp_map_particle_infos into a device buffer of 16 N bytes (e.g. a
        torch CUDA tensor's data_ptr()), ordered on the solver's stream."""
        _check(self.lib.bp_map_particle_infos(self.h, C.c_void_p(device_ptr)))

    def set_rigid_state(self, quats):
        q = np.ascontiguousarray(quats, np.float32)
        _check(self.lib.bp_set_rigid_state(self.h, _fp(q)))

    def debug(self, which: int):
        if which in (0, 1):
            out = np.empty(self.n, np.float32)
        elif which in (2, 3):
            out = np.empty(self.n, np.int32)
        elif which == 6:  # largest |term| the last backward delivered in fixed point
            out = np.zeros(1, np.float32)
        else:
            out = np.empty((self.n, 3), np.float32)
        _check(self.lib.bp_debug_fetch(self.h, int(which), out.ctypes.data_as(C.c_void_p)))
        return out

    def query_original_id(self, slot: int) -> int:
        o = C.c_int()
        _check(self.lib.bp_query_original_id(self.h, int(slot), C.byref(o)))
        return o.value

    def query_slot(self, orig: int) -> int:
        o = C.c_int()
        _check(self.lib.bp_query_slot(self.h, int(orig), C.byref(o)))
        return o.value

    def get_particle(self, slot: int):
        out = np.zeros(3, np.float32)
        _check(self.lib.bp_get_particle(self.h, int(slot), _fp(out)))
        return out

    def set_particle(self, slot: int, pos, vel):
        pp = np.asarray(pos, np.float32)
        pv = np.asarray(vel, np.float32)
        _check(self.lib.bp_set_particle(self.h, int(slot), _fp(pp), _fp(pv)))

    def device_arrays(self) -> bp_device_view:
        v = bp_device_view()
        _check(self.lib.bp_get_device_arrays(self.h, C.byref(v)))
        return v

    def timed_substeps(self, substeps: int, dt: float) -> float:
        ms = C.c_float()
        _check(self.lib.bp_solver_timed_substeps(self.h, int(substeps), float(dt), C.byref(ms)))
        return float(ms.value)

    def enable_kernel_timing(self, on: bool = True):
        _check(self.lib.bp_solver_enable_kernel_timing(self.h, 1 if on else 0))

    def kernel_times(self) -> dict:
        mx = 32
        names = C.create_string_buffer(32 * mx)
        ms = (C.c_float * mx)()
        ln = (C.c_int * mx)()
        cnt = C.c_int()
        _check(self.lib.bp_solver_kernel_times(self.h, mx, names, ms, ln, C.byref(cnt)))
        out = {}
        raw = names.raw
        for i in range(min(cnt.value, mx)):
            nm = raw[32 * i: 32 * i + 32].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(ln[i]))
        return out


def stream_copy_bandwidth(nbytes: int = 1 << 30, reps: int = 5) -> dict:
    """bp_stream_copy_bandwidth: the library's streaming copy kernel over
    `nbytes`, best time and bytes read plus written per second."""
    ms = C.c_float()
    _check(load_library().bp_stream_copy_bandwidth(C.c_size_t(nbytes), int(reps), C.byref(ms)))
    return {"bytes": 2 * nbytes, "ms": round(ms.value, 4), "GBs": round(2 * nbytes / (ms.value * 1e-3) / 1e9, 1)}


def neighbor_grid(xs, cell_size, grid_dims):
    """bp_neighbor_grid: GPU hash + stable sort + cell ranges of positions xs [n,3]."""
    lib = load_library()
    xs = np.ascontiguousarray(xs, np.float32)
    n = len(xs)
    cs = np.asarray(cell_size, np.float32)
    dims = np.asarray(grid_dims, np.int32)
    cells = int(np.prod(dims))
    keys = np.empty(n, np.int32)
    perm = np.empty(n, np.int32)
    st = np.empty(cells, np.int32)
    en = np.empty(cells, np.int32)
    _check(lib.bp_neighbor_grid(_fp(xs), n, _fp(cs), _ip(dims), _ip(keys), _ip(perm), _ip(st), _ip(en)))
    return keys, perm, st, en


class BatchSolver(Solver):
    """Many independent scenes advanced together (bp_batch_create): one pass
    launch serves every scene.  Original ids are global; scene k owns
    ranges[k] = (first, count)."""

    def __init__(self, scene_list, params_list, stream: int | None = None):
        self.lib = load_library()
        self.scene_arrays = [SceneArrays(sc) for sc in scene_list]
        descs = (bp_scene_desc * len(scene_list))(*[a.desc for a in self.scene_arrays])
        prms = (bp_params * len(scene_list))(*params_list)
        self.params = params_list[0]
        h = C.c_void_p()
        _check(self.lib.bp_batch_create(descs, prms, len(scene_list), C.c_void_p(stream) if stream else None,
                                        C.byref(h)))
        self.h = h
        self.n = sum(a.n for a in self.scene_arrays)
        self.num_rigid = sum(len(a.rranges) for a in self.scene_arrays)
        self.ranges = []
        for k in range(len(scene_list)):
            f, c = C.c_int(), C.c_int()
            _check(self.lib.bp_batch_scene_range(self.h, k, C.byref(f), C.byref(c)))
            self.ranges.append((f.value, c.value))
