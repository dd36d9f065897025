"""GPU: the headless renderer boundary (bp_export_frame, bp_map_particle_infos;
the reference's mapParticleInfos / mapMatrices, main.cu:273-291) and the frame
files written from it."""
import os
import tempfile

import numpy as np
import pytest

import brasspan as bp
import frames
import oracle_py as orc
import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0
POS_TOL = 2e-5  # one substep from the same state, oracle in the GPU's summation order


def case(msg):
    print(f"[gpu case] {msg}", flush=True)


def small_mixed():
    return scenes.mixed_phase(fluid_dims=(12, 12, 12), num_boxes=4, cloth_n=12, grid=(64, 64, 64), box=6.4)


def test_export_frame_is_the_state():
    sc = small_mixed()
    prm = scenes.mixed_phase_params()
    case(f"export frame n={sc.num_particles()}")
    s = bp.Solver(sc, prm)
    for _ in range(4):
        s.step(1, DT)
    rec, mats = s.export_frame()
    pos, _ = s.state_original_order()
    assert np.array_equal(rec[:, :3], pos)
    assert np.array_equal(rec[:, 3].copy().view(np.int32), np.asarray(sc.group_ids, np.int32))
    rest = np.asarray(sc.rigidbody_initial_positions, np.float64).reshape(-1, 3)
    assert len(mats) == len(sc.rigidbody_particle_id_ranges) == 4
    for b, (s0, e0) in enumerate(sc.rigidbody_particle_id_ranges):
        m = mats[b].astype(np.float64)
        r = m[:3, :3]
        assert np.abs(r @ r.T - np.eye(3)).max() < 1e-5 and abs(np.linalg.det(r) - 1.0) < 1e-5
        assert np.array_equal(m[3], [0.0, 0.0, 0.0, 1.0])
        # shape matching leaves every body particle at R rest + CM
        assert np.abs(rest[s0:e0] @ r.T + m[:3, 3] - pos[s0:e0]).max() < 2e-5

    case("map particle infos into a device buffer")
    import ctypes as C

    hip = bp.hip_runtime()
    nbytes = C.c_size_t(16 * s.n)
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), nbytes) == 0
    try:
        assert hip.hipMemset(dev, 0xFF, nbytes) == 0  # NaN bits: every record must be written
        s.map_particle_infos(dev.value)
        s.sync()
        out = np.empty((s.n, 4), np.float32)
        assert hip.hipMemcpy(C.c_void_p(out.ctypes.data), dev, nbytes, 2) == 0  # device to host
    finally:
        hip.hipFree(dev)
    assert np.array_equal(out.view(np.uint32), rec.view(np.uint32))
    s.close()


def test_frame_file_against_the_oracle():
    sc = small_mixed()
    prm = scenes.mixed_phase_params()
    og = orc.Oracle(bp.SceneArrays(sc), prm, fp64=False, mode=orc.GATHER)
    for _ in range(3):
        og.step(1, DT)
    case(f"frame file vs oracle n={sc.num_particles()}")
    s = bp.Solver(sc, prm)
    p, v, m = og.get_state()
    s.set_state(p.astype(np.float32), v.astype(np.float32), m)
    q, _ = og.rigid_state()
    s.set_rigid_state(q.astype(np.float32))
    s.step(1, DT)
    og.step(1, DT)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "f.bpfr")
        frames.save_solver_frame(s, path, frame=4, time=4 * DT)
        gpu = frames.read_frame(path)
    xo, _ = og.state_original_order()
    qo, co = og.rigid_state()
    ora = frames.frame_from_state(xo, sc.group_ids, qo, co)
    d = frames.frame_diff(gpu, ora)
    print(d)
    assert gpu["frame"] == 4 and d["group_mismatches"] == 0
    assert d["max_position"] <= POS_TOL and d["max_matrix"] <= 1e-4
    s.close()
