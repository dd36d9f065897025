"""GPU: the legacy scene format through ParticleSolver(OldSceneFormat)
(particlesolver.cuh:218-342): brass-pan_amd/bin/legacy_scenes builds scenes
shaped like main.cu's legacy scenes (rigid-box stack, fluid block, ropes and
cloths), steps them and writes each last frame; with "dump" it also writes
each assembled scene, which the oracle then runs from the same input."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import brasspan as bp
import frames
import oracle_py as orc
import scenefile
from test_gpu_parity import POS_TOL_GATHER, per_substep_parity

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "brass-pan_amd", "bin", "legacy_scenes")


def test_legacy_scenes_run():
    print("[gpu case] legacy scenes, 60 frames", flush=True)
    with tempfile.TemporaryDirectory() as d:
        prefix = os.path.join(d, "f_")
        r = subprocess.run([EXE, "60", prefix], capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stderr
        print(r.stdout)
        rows = {ln.split()[0]: dict(zip(ln.split()[1::2], ln.split()[2::2])) for ln in r.stdout.splitlines()}
        fr = {k: frames.read_frame(prefix + k + ".bpfr") for k in ("rigid", "fluid", "ropes")}
    # counters the legacy constructor advances
    assert rows["rigid"]["particles"] == "75" and rows["rigid"]["rigid"] == "3"
    assert rows["fluid"]["particles"] == "4096"
    n = 10
    cloth_pairs = 2 * n * (n - 1) + 2 * (n - 1) ** 2 + 6 * (n - 2) ** 2
    assert rows["ropes"]["particles"] == str(20 + 2 * n * n)
    assert rows["ropes"]["pairs"] == str(19 + 18 + 2 * cloth_pairs)
    assert rows["ropes"]["faces"] == str(2 * 2 * (n - 1) ** 2) and rows["ropes"]["pins"] == "3"
    for k in rows:
        assert rows[k]["finite"] == "1"
    # after a second everything rests on (or hangs above) the floor plane y = 0
    for k in ("rigid", "fluid"):
        assert abs(float(rows[k]["ymin"]) - 0.05) < 0.02, rows[k]
    # frames: one group per part in insertion order, rigid matrices are rotations about the body's centre
    assert np.array_equal(np.unique(fr["rigid"]["groups"]), [0, 1, 2])
    assert np.array_equal(np.unique(fr["ropes"]["groups"]), [0, 1, 2])
    pos = fr["rigid"]["positions"].astype(np.float64)
    for b, (s, e) in enumerate([(0, 27), (27, 51), (51, 75)]):
        m = fr["rigid"]["matrices"][b].astype(np.float64)
        assert np.abs(m[:3, :3] @ m[:3, :3].T - np.eye(3)).max() < 1e-5
        assert np.abs(m[:3, 3] - pos[s:e].mean(0)).max() < 1e-5


@pytest.fixture(scope="module")
def legacy_dump():
    d = tempfile.mkdtemp()
    prefix = os.path.join(d, "f_")
    print("[gpu case] legacy scenes, one frame with scene dumps", flush=True)
    r = subprocess.run([EXE, "1", prefix, "dump"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    return prefix


def _legacy_params(rho):
    # ParticleSolver(OldSceneFormat) with the reference literals and the scene's rest density
    return bp.default_params(**({} if rho is None else {"fluid_rest_density": rho}))


@pytest.mark.parametrize("name", ["rigid", "fluid", "ropes"])
def test_legacy_scene_frame_matches_oracle(legacy_dump, name):
    """The C++ program's frame after update(2, 1/60) from the assembled
    legacy scene, against the oracle run on the scene file the program wrote
    (the same assembly, read back in Python)."""
    sc, rho = scenefile.read_scene_file(legacy_dump + name + ".bpsc")
    fr = frames.read_frame(legacy_dump + name + ".bpfr")
    o = orc.Oracle(bp.SceneArrays(sc), _legacy_params(rho), fp64=False, mode=orc.GATHER)
    o.step(2, 1.0 / 60.0)
    op, _ = o.state_original_order()
    assert np.array_equal(fr["groups"], sc.group_ids)
    err = float(np.abs(fr["positions"].astype(np.float64) - op).max())
    print(f"legacy {name}: n={sc.num_particles()} frame vs oracle after 2 substeps: {err:.3e}")
    assert err <= 4 * POS_TOL_GATHER


@pytest.mark.parametrize("name", ["rigid", "ropes"])
def test_legacy_scene_per_substep_parity(legacy_dump, name):
    """Per-substep GPU vs oracle on the assembled legacy scene (rigid stack:
    shape matching on a plane; ropes: rope and cloth distance constraints,
    pins and wind faces)."""
    sc, rho = scenefile.read_scene_file(legacy_dump + name + ".bpsc")
    per_substep_parity(sc, _legacy_params(rho), 3, f"legacy {name}")
