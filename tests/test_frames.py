"""CPU: the headless frame format (brass-pan_amd/frames.py) that replaces the
reference's renderer boundary (main.cu:273-291)."""
import numpy as np
import pytest

import frames


def sample(n=7, nb=2, seed=0):
    rng = np.random.default_rng(seed)
    rec = rng.normal(size=(n, 4)).astype(np.float32)
    rec[:, 3] = np.arange(n, dtype=np.int32).view(np.float32)  # group ids as int bits
    mats = rng.normal(size=(nb, 4, 4)).astype(np.float32)
    return rec, mats


def test_round_trip_and_layout(tmp_path):
    rec, mats = sample()
    p = tmp_path / "f.bpfr"
    frames.write_frame(p, rec, mats, frame=42, time=0.7)
    raw = p.read_bytes()
    assert raw[:4] == b"BPFR" and len(raw) == 32 + 16 * 7 + 64 * 2
    # matrices are stored column-major: the first stored float of body 0 is M[0][0], the second M[1][0]
    first = np.frombuffer(raw, "<f4", 2, 32 + 16 * 7)
    assert first[0] == mats[0, 0, 0] and first[1] == mats[0, 1, 0]
    fr = frames.read_frame(p)
    assert fr["frame"] == 42 and fr["time"] == 0.7
    assert np.array_equal(fr["positions"], rec[:, :3])
    assert np.array_equal(fr["groups"], np.arange(7, dtype=np.int32))
    assert np.array_equal(fr["matrices"], mats)


def test_empty_frame(tmp_path):
    p = tmp_path / "e.bpfr"
    frames.write_frame(p, np.zeros((0, 4), np.float32), np.zeros((0, 4, 4), np.float32))
    fr = frames.read_frame(p)
    assert fr["positions"].shape == (0, 3) and fr["matrices"].shape == (0, 4, 4)


def test_rejects_bad_files(tmp_path):
    rec, mats = sample()
    p = tmp_path / "f.bpfr"
    frames.write_frame(p, rec, mats)
    raw = p.read_bytes()
    (tmp_path / "short.bpfr").write_bytes(raw[:-4])
    with pytest.raises(ValueError):
        frames.read_frame(tmp_path / "short.bpfr")
    (tmp_path / "magic.bpfr").write_bytes(b"XXXX" + raw[4:])
    with pytest.raises(ValueError):
        frames.read_frame(tmp_path / "magic.bpfr")
    (tmp_path / "tiny.bpfr").write_bytes(raw[:10])
    with pytest.raises(ValueError):
        frames.read_frame(tmp_path / "tiny.bpfr")


def test_diff(tmp_path):
    rec, mats = sample()
    frames.write_frame(tmp_path / "a.bpfr", rec, mats)
    rec2, mats2 = rec.copy(), mats.copy()
    rec2[3, 1] += 0.25
    rec2[5, 3] = np.int32(99).view(np.float32)
    mats2[1, 2, 3] -= 0.5
    frames.write_frame(tmp_path / "b.bpfr", rec2, mats2)
    a, b = frames.read_frame(tmp_path / "a.bpfr"), frames.read_frame(tmp_path / "b.bpfr")
    assert frames.frame_diff(a, a) == {"max_position": 0.0, "max_matrix": 0.0, "group_mismatches": 0}
    d = frames.frame_diff(a, b)
    assert d["max_position"] == pytest.approx(0.25, rel=1e-6)
    assert d["max_matrix"] == pytest.approx(0.5, rel=1e-6)
    assert d["group_mismatches"] == 1
    frames.write_frame(tmp_path / "c.bpfr", rec[:3], mats)
    with pytest.raises(ValueError):
        frames.frame_diff(a, frames.read_frame(tmp_path / "c.bpfr"))


def test_rigid_matrices_carry_rest_offsets_to_the_oracle_bodies(have_oracle):
    """frames.rigid_matrices on the oracle's own rigid state after shape
    matching has rotated the bodies: each matrix is a rotation and maps every
    rest offset of its body to that particle's position."""
    import brasspan as bp
    import oracle_py as orc
    import scenes

    sc = scenes.Scene(particle_radius=0.025, fluid_kernel_radius=0.1, cell_size=0.1, grid_dims=(32, 32, 32))
    for k, axis in enumerate(((1, 2, 3), (0, 1, 0), (-2, 1, 0.5))):
        a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
        q = (*(a * np.sin(0.4 + 0.3 * k)), np.cos(0.4 + 0.3 * k))
        sc.add_rigid_box((3, 4, 3), (0.6 + 0.5 * k, 0.4, 0.8), 0.05, 1.0, rotation=q)
    sc.add_box_planes((0, 0, 0), (3.2, 3.2, 3.2))
    prm = bp.default_params(iterations=2, akinci=0, vorticity_eps=0.0)
    o = orc.Oracle(bp.SceneArrays(sc), prm, fp64=False, mode=orc.GATHER)
    for _ in range(6):
        o.step(1, 1.0 / 120.0)
    x, _ = o.state_original_order()
    quats, cms = o.rigid_state()
    mats = frames.rigid_matrices(quats, cms)
    rest = np.asarray(sc.rigidbody_initial_positions, np.float64).reshape(-1, 3)
    for b, (s, e) in enumerate(sc.rigidbody_particle_id_ranges):
        m = mats[b].astype(np.float64)
        r = m[:3, :3]
        assert np.abs(r @ r.T - np.eye(3)).max() < 1e-5 and abs(np.linalg.det(r) - 1.0) < 1e-5
        assert np.array_equal(m[3], [0.0, 0.0, 0.0, 1.0])
        pred = rest[s:e] @ r.T + m[:3, 3]
        assert np.abs(pred - x[s:e]).max() < 1e-5
    # the bodies did rotate and move, so a transposed matrix would not pass
    assert np.abs(mats[:, :3, :3] - np.eye(3, dtype=np.float32)).max() > 0.1
