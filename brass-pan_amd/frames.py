"""Headless frame dumps: the state the reference hands its renderer
(main.cu:273-291), written to a file per frame for inspection and for
diffing a run against the oracle.

File layout (little endian):
- header, 32 B: magic b"BPFR", u32 version (1), u32 particles, u32 rigid
  bodies, u64 frame index, f64 simulated time in seconds;
- particles x 16 B: float32 x, y, z and the int32 group id, in original-id
  order (bp_export_frame);
- bodies x 64 B: float32 column-major 4x4 per rigid body, rotation plus
  centre of mass (mapMatrices, main.cu:284-291).
"""
import struct

import numpy as np

MAGIC = b"BPFR"
VERSION = 1
_HEADER = struct.Struct("<4sIIIQd")


def write_frame(path, records, matrices, frame=0, time=0.0):
    """records: [N, 4] float32 (x, y, z, group id bits); matrices: [B, 4, 4]
    with M[b][row][col] (stored column-major)."""
    rec = np.ascontiguousarray(records, np.float32).reshape(-1, 4)
    mats = np.asarray(matrices, np.float32).reshape(-1, 4, 4)
    with open(path, "wb") as f:
        f.write(_HEADER.pack(MAGIC, VERSION, len(rec), len(mats), int(frame), float(time)))
        f.write(rec.astype("<f4").tobytes())
        f.write(np.ascontiguousarray(mats.transpose(0, 2, 1)).astype("<f4").tobytes())


def read_frame(path):
    """Returns dict(frame, time, positions [N, 3], groups [N], matrices [B, 4, 4])."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < _HEADER.size:
        raise ValueError(f"{path}: shorter than a frame header")
    magic, version, n, nb, frame, time = _HEADER.unpack_from(data, 0)
    if magic != MAGIC or version != VERSION:
        raise ValueError(f"{path}: not a version-{VERSION} frame ({magic!r}, {version})")
    want = _HEADER.size + 16 * n + 64 * nb
    if len(data) != want:
        raise ValueError(f"{path}: {len(data)} bytes, expected {want}")
    rec = np.frombuffer(data, "<f4", 4 * n, _HEADER.size).reshape(n, 4)
    mats = np.frombuffer(data, "<f4", 16 * nb, _HEADER.size + 16 * n).reshape(nb, 4, 4).transpose(0, 2, 1)
    return {"frame": frame, "time": time, "positions": rec[:, :3].copy(),
            "groups": rec[:, 3].copy().view(np.int32), "matrices": mats.copy()}


def rigid_matrices(quats, cms):
    """[B, 4, 4] from unit quaternions (x, y, z, w) [B, 4] and centres of mass
    [B, 3], in float32 with the terms in the order bp_export_frame sums them:
    the frame of a state that came from elsewhere (the oracle, say)."""
    q = np.asarray(quats, np.float32).reshape(-1, 4)
    c = np.asarray(cms, np.float32).reshape(-1, 3)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    two, one = np.float32(2.0), np.float32(1.0)
    xx, yy, zz = x * x, y * y, z * z
    m = np.zeros((len(q), 4, 4), np.float32)
    m[:, 0, 0] = one - two * yy - two * zz
    m[:, 1, 0] = two * x * y + two * w * z
    m[:, 2, 0] = two * x * z - two * w * y
    m[:, 0, 1] = two * x * y - two * w * z
    m[:, 1, 1] = one - two * xx - two * zz
    m[:, 2, 1] = two * y * z + two * w * x
    m[:, 0, 2] = two * x * z + two * w * y
    m[:, 1, 2] = two * y * z - two * w * x
    m[:, 2, 2] = one - two * xx - two * yy
    m[:, :3, 3] = c
    m[:, 3, 3] = one
    return m


def frame_from_state(positions, groups, quats, cms, frame=0, time=0.0):
    """A frame dict (as read_frame returns) from original-order positions,
    group ids and rigid state."""
    return {"frame": frame, "time": time, "positions": np.asarray(positions, np.float32).reshape(-1, 3),
            "groups": np.asarray(groups, np.int32), "matrices": rigid_matrices(quats, cms)}


def frame_diff(a, b):
    """Largest position and matrix differences between two frames of one scene,
    and the number of particles whose group ids differ."""
    if a["positions"].shape != b["positions"].shape or a["matrices"].shape != b["matrices"].shape:
        raise ValueError("frames of different scenes")
    dp = np.abs(a["positions"].astype(np.float64) - b["positions"]).max(initial=0.0)
    dm = np.abs(a["matrices"].astype(np.float64) - b["matrices"]).max(initial=0.0)
    return {"max_position": float(dp), "max_matrix": float(dm),
            "group_mismatches": int(np.count_nonzero(a["groups"] != b["groups"]))}


def save_solver_frame(solver, path, frame=0, time=0.0):
    """Exports the solver's current state (bp_export_frame) to `path`."""
    rec, mats = solver.export_frame()
    write_frame(path, rec, mats, frame, time)
