"""CPU: the legacy scene format (OldSceneFormat and the assembly of the
reference's legacy constructor, particlesolver.cuh:218-342, 395-710) as the
C++ shim restates it (brass-pan_amd/cpp/brasspan/oldscene.h), the defaults of
the C++ Scene (scene.h) and scene files.  The C++ checks are a host-only
program (tests/cpp/legacy_assembly.cpp) compiled here."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import scenefile
import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def assembly(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.exists("/opt/rocm/include/hip/hip_vector_types.h"):
        pytest.skip("g++ or the HIP headers are not available")
    d = tmp_path_factory.mktemp("legacy")
    exe = d / "legacy_assembly"
    # the shim's headers need only HIP's vector types (as the Makefile's examples)
    r = subprocess.run([gxx, "-O1", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-I", os.path.join(ROOT, "brass-pan_amd", "cpp"),
                        os.path.join(ROOT, "tests", "cpp", "legacy_assembly.cpp"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    path = d / "sample.bpsc"
    r = subprocess.run([str(exe), str(path)], capture_output=True, text=True, timeout=60)
    return r, path


def test_legacy_assembly(assembly):
    r, _ = assembly
    assert r.returncode == 0, r.stdout + r.stderr
    assert "legacy assembly ok" in r.stdout


def _bits(v):
    return int(np.float32(v).view(np.uint32))


def test_scene_cell_size_and_kernel_radius_follow_the_reference(assembly):
    """The reference derives the hash cell as particleRadius * 2.3f (float)
    when the solver is built (particlesolver.cuh:150), and initialises
    fluidKernelRadius = 2.3 * particleRadius in double at Scene construction
    (scene.h:501), so a radius set afterwards changes the cell but not h.
    The C++ Scene must hand the solver the same bits; an explicit cellSize is
    kept as set.  The Python Scene derives both from its radius argument with
    the same expressions."""
    r, _ = assembly
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("scene defaults")][0]
    got = [int(v, 16) for v in line.split()[2:]]
    f32 = np.float32
    cell = lambda rad: _bits(f32(rad) * f32(2.3))
    h_ref = lambda rad: _bits(2.3 * float(f32(rad)))
    assert got[0] == cell(0.05) and got[1] == h_ref(0.05)
    assert got[2] == cell(0.025) and got[3] == h_ref(0.05) and got[5] == cell(0.025)
    assert got[4] == _bits(0.2)
    sc = scenes.Scene(particle_radius=0.025)
    assert _bits(sc.cell_size[0]) == cell(0.025) and _bits(sc.fluid_kernel_radius) == h_ref(0.025)


def test_scene_file_from_cpp_assembly(assembly):
    """The legacy assembly written by the C++ writer (scenefile.h) reads back
    with the counts, offsets and rest density the assembly check asserts."""
    r, path = assembly
    assert r.returncode == 0
    sc, rho = scenefile.read_scene_file(str(path))
    assert sc.num_particles() == 159 and rho == pytest.approx(1200.0)
    assert sc.rigidbody_particle_id_ranges == [(0, 27), (27, 51)]
    assert np.all(sc.phases[59:86] == -1) and np.all(sc.phases[86:113] == -2)
    n = 6
    assert len(sc.distance_pairs) == 17 + 2 * n * (n - 1) + 2 * (n - 1) ** 2 + 6 * (n - 2) ** 2
    assert sc.faces[0] == (123, 124, 123 + n) and sc.immovable_constraints == [123, 123 + n - 1]
    assert _bits(sc.cell_size[0]) == _bits(np.float32(0.05) * np.float32(2.3)) and sc.grid_dims == (512,) * 3
    rest = sc.rigidbody_initial_positions
    assert rest.shape == (51, 3) and np.abs(rest[0:27].sum(0)).max() < 1e-4


def test_scene_file_round_trip(tmp_path):
    sc = scenes.mixed_phase(fluid_dims=(6, 6, 6), num_boxes=2, cloth_n=8, grid=(64, 64, 64))
    sc.add_cloth((1.0, 2.0, 1.0), (0.05, 0, 0), (0, 0, 0.05), 4, 4, 1.0, 0.2, 0.1, dihedral_bending=True)
    p = tmp_path / "s.bpsc"
    scenefile.write_scene_file(str(p), sc, 987.5)
    back, rho = scenefile.read_scene_file(str(p))
    assert rho == 987.5
    for a in ("positions", "masses", "phases", "group_ids", "rigidbody_initial_positions"):
        assert np.array_equal(getattr(back, a), getattr(sc, a)), a
    for a in ("distance_pairs", "bending_constraints", "faces", "rigidbody_particle_id_ranges",
              "immovable_constraints"):
        assert [tuple(np.atleast_1d(v)) for v in getattr(back, a)] == [tuple(np.atleast_1d(v)) for v in
                                                                       getattr(sc, a)], a
    assert np.allclose(np.asarray(back.distance_params), np.asarray(sc.distance_params, np.float32))
    assert len(back.bending_constraints) == 9
    assert back.cell_size == sc.cell_size and back.grid_dims == sc.grid_dims
    with open(p, "rb") as f:
        data = f.read()
    with open(tmp_path / "t.bpsc", "wb") as f:
        f.write(data[:-7])
    with pytest.raises(ValueError):
        scenefile.read_scene_file(str(tmp_path / "t.bpsc"))
