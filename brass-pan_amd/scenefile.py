"""Scene files: a Scene (scenes.Scene, or the C++ Scene of
cpp/brasspan/scene.h written by writeSceneFile in cpp/brasspan/scenefile.h)
on disk, so a scene built on one side can be replayed on the other.  No
counterpart in the reference, which keeps scenes in memory only.

Layout, little endian:
- header, 48 B: b"BPSC", u32 version 1, f32 particle radius, f32 kernel
  radius, f32 cell size[3], i32 grid dims[3], f32 fluid rest density (NaN
  when the scene leaves it to the solver parameters), u32 array count;
- per array: 16 B name (NUL padded), 1 B type (b"f" float32, b"i" int32),
  u8 width, u16 0, u64 rows, then rows x width values.
"""
from __future__ import annotations

import math
import struct

import numpy as np

import scenes

MAGIC = b"BPSC"
VERSION = 1
_HEADER = struct.Struct("<4sIff3f3ifI")
_ARRAY = struct.Struct("<16scBHQ")

# name -> (type, width, Scene attribute)
ARRAYS = {
    "positions": ("f", 3), "masses": ("f", 1), "phases": ("i", 1), "group_ids": ("i", 1),
    "plane_origins": ("f", 3), "plane_normals": ("f", 3), "distance_pairs": ("i", 2),
    "distance_params": ("f", 2), "bendings": ("i", 4), "faces": ("i", 3), "rigid_ranges": ("i", 2),
    "rigid_rest": ("f", 3), "immovables": ("i", 1),
}


def write_scene_file(path, sc: scenes.Scene, fluid_rest_density: float | None = None):
    """Writes `sc` in the layout above."""
    arrays = {
        "positions": sc.positions, "masses": sc.masses, "phases": sc.phases, "group_ids": sc.group_ids,
        "plane_origins": [p[0] for p in sc.planes], "plane_normals": [p[1] for p in sc.planes],
        "distance_pairs": sc.distance_pairs, "distance_params": sc.distance_params,
        "bendings": sc.bending_constraints, "faces": sc.faces, "rigid_ranges": sc.rigidbody_particle_id_ranges,
        "rigid_rest": sc.rigidbody_initial_positions, "immovables": sc.immovable_constraints,
    }
    rho = float("nan") if fluid_rest_density is None else float(fluid_rest_density)
    with open(path, "wb") as f:
        f.write(_HEADER.pack(MAGIC, VERSION, sc.particle_radius, sc.fluid_kernel_radius, *sc.cell_size,
                             *sc.grid_dims, rho, len(arrays)))
        for name, (typ, width) in ARRAYS.items():
            dt = "<f4" if typ == "f" else "<i4"
            a = np.asarray(arrays[name], dt).reshape(-1, width) if len(arrays[name]) else np.zeros((0, width), dt)
            f.write(_ARRAY.pack(name.encode(), typ.encode(), width, 0, len(a)))
            f.write(np.ascontiguousarray(a).tobytes())


def read_scene_file(path):
    """Returns (scenes.Scene, fluid rest density or None)."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < _HEADER.size:
        raise ValueError(f"{path}: shorter than a scene header")
    magic, version, radius, h, cx, cy, cz, gx, gy, gz, rho, count = _HEADER.unpack_from(data, 0)
    if magic != MAGIC or version != VERSION:
        raise ValueError(f"{path}: not a version {VERSION} scene file")
    off = _HEADER.size
    arrays = {}
    for _ in range(count):
        if off + _ARRAY.size > len(data):
            raise ValueError(f"{path}: truncated array header")
        name, typ, width, _, rows = _ARRAY.unpack_from(data, off)
        off += _ARRAY.size
        nbytes = rows * width * 4
        if off + nbytes > len(data):
            raise ValueError(f"{path}: truncated array {name!r}")
        dt = "<f4" if typ == b"f" else "<i4"
        arrays[name.rstrip(b"\0").decode()] = np.frombuffer(data, dt, rows * width, off).reshape(rows, width).copy()
        off += nbytes
    missing = set(ARRAYS) - set(arrays)
    if missing:
        raise ValueError(f"{path}: missing arrays {sorted(missing)}")
    sc = scenes.Scene(particle_radius=radius, fluid_kernel_radius=h, cell_size=(cx, cy, cz), grid_dims=(gx, gy, gz))
    sc.fluid_kernel_radius = float(np.float32(h))
    pos = arrays["positions"].astype(np.float32)
    sc._pos = [pos]
    sc._vel = [np.zeros_like(pos)]
    sc._mass = [arrays["masses"].reshape(-1).astype(np.float32)]
    sc._phase = [arrays["phases"].reshape(-1)]
    sc._group = [arrays["group_ids"].reshape(-1)]
    sc._n = len(pos)
    sc.planes = [(o.astype(np.float32), n.astype(np.float32))
                 for o, n in zip(arrays["plane_origins"], arrays["plane_normals"])]
    sc.distance_pairs = [tuple(int(v) for v in r) for r in arrays["distance_pairs"]]
    sc.distance_params = [tuple(float(v) for v in r) for r in arrays["distance_params"]]
    sc.bending_constraints = [tuple(int(v) for v in r) for r in arrays["bendings"]]
    sc.faces = [tuple(int(v) for v in r) for r in arrays["faces"]]
    sc.rigidbody_particle_id_ranges = [tuple(int(v) for v in r) for r in arrays["rigid_ranges"]]
    rest = arrays["rigid_rest"]
    sc._rigid_rest = {s: rest[s:e].copy() for s, e in sc.rigidbody_particle_id_ranges}
    sc.immovable_constraints = [int(v) for v in arrays["immovables"].reshape(-1)]
    return sc, (None if math.isnan(rho) else float(rho))
