"""The C-ABI library: loads on a CPU-only host, exports every entry point
include/brasspan.h declares, fills the reference's literals and rejects bad
scenes before touching a device.  No compute calls (CPU only)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import brasspan as bp
import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    txt = open(os.path.join(ROOT, "include", "brasspan.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bp_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = bp.load_library()
    names = declared_entry_points()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(bp.SIGNATURES), "binding and header disagree"
    assert lib.bp_abi_version() == 1


def test_default_params_are_reference_literals():
    p = bp.default_params()
    assert p.iterations == 2 and p.stabilization_passes == 2          # particlesolver.cuh:862,822
    assert tuple(p.gravity) == pytest.approx((0.0, -9.8, 0.0))       # :56
    assert p.solid_density_scale == 2.0 and p.lambda_relaxation == 300.0  # :909-910
    assert p.lambda_denominator == 15000.0 and p.lambda_mode == 0    # fluid.cuh:76,153
    assert p.akinci == 1                                             # :785
    assert p.scorr_k == pytest.approx(1e-4) and p.scorr_n == 4 and p.scorr_dq == pytest.approx(0.03)
    assert p.delta_clamp == 5.0                                      # fluid.cuh:214
    assert p.vorticity_eps == pytest.approx(0.001) and p.tension_gamma == 0.5
    assert p.xsph_c == pytest.approx(0.0002)                         # :1055
    assert p.friction_static == 0.5 and p.friction_dynamic == 0.0    # cubrasstower.cuh:18-19
    assert p.mass_scaling_k == 2.0 and p.sleep_epsilon == 0.0        # cubrasstower.cuh:20-21
    assert p.drag_coeff == 10.0 and p.lift_coeff == 100.0            # wind.cuh:40,43
    assert p.bend_stiffness == pytest.approx(0.8)                    # bending.cuh:65
    assert p.shape_matching_iterations == 20                         # shapematching.cuh:72


def create_status(scene, params=None):
    lib = bp.load_library()
    arr = bp.SceneArrays(scene)
    h = C.c_void_p()
    st = lib.bp_solver_create(C.byref(arr.desc), C.byref(params or bp.default_params()), None, C.byref(h))
    return st, lib.bp_last_error_string().decode()


def small_scene(**kw):
    sc = scenes.Scene(particle_radius=0.025, fluid_kernel_radius=0.1, cell_size=0.1, **kw)
    sc.add_fluid_block((4, 4, 4), (0.5, 0.5, 0.5), 0.05, 1.0)
    return sc


def test_create_rejects_invalid_grid():
    st, msg = create_status(small_scene(grid_dims=(64, 48, 64)))
    assert st == 1 and "power" in msg
    st, _ = create_status(small_scene(grid_dims=(2, 64, 64)))
    assert st == 1


def test_create_rejects_large_rigid_body():
    sc = small_scene(grid_dims=(64, 64, 64))
    sc.add_rigid_box((5, 5, 5), (1.0, 1.0, 1.0), 0.05, 1.0)  # 125 > 64 particles
    st, msg = create_status(sc)
    assert st == 2 and "64" in msg


def test_create_rejects_bad_ids_and_params():
    sc = small_scene(grid_dims=(64, 64, 64))
    sc.distance_pairs.append((0, 10_000))
    sc.distance_params.append((0.1, 1.0))
    st, _ = create_status(sc)
    assert st == 1
    st, msg = create_status(small_scene(grid_dims=(64, 64, 64)), bp.default_params(fluid_rest_density=0.0))
    assert st == 1 and "rest_density" in msg


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="device present")
def test_create_without_device_fails_loudly():
    st, msg = create_status(small_scene(grid_dims=(64, 64, 64)))
    assert st == 5, msg  # BP_ERR_NO_DEVICE: no CPU fallback exists
    with pytest.raises(bp.BrassPanError):
        bp.Solver(small_scene(grid_dims=(64, 64, 64)))


EXAMPLES = [os.path.join(ROOT, "brass-pan_amd", "bin", name) for name in ("dam_break", "batch_and_grad", "legacy_scenes")]


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="device present")
@pytest.mark.parametrize("exe", EXAMPLES)
def test_cpp_api_example_builds_and_fails_loudly_without_device(exe):
    import subprocess
    assert os.path.exists(exe), f"{exe} not built"
    r = subprocess.run([exe, "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "no HIP device" in r.stderr


def _create_desc_status(arr, params=None):
    lib = bp.load_library()
    h = C.c_void_p()
    st = lib.bp_solver_create(C.byref(arr.desc), C.byref(params or bp.default_params()), None, C.byref(h))
    return st, lib.bp_last_error_string().decode()


def test_create_rejects_null_and_overlapping_rigid_ranges():
    """One shape-matching workgroup per body writes that body's particles, so
    ranges must exist and be disjoint; both are checked before any device
    call (BP_ERR_INVALID_ARG)."""
    sc = small_scene(grid_dims=(64, 64, 64))
    sc.add_rigid_box((2, 2, 2), (1.0, 1.0, 1.0), 0.05, 1.0)
    sc.add_rigid_box((2, 2, 2), (2.0, 1.0, 1.0), 0.05, 1.0)
    arr = bp.SceneArrays(sc)
    arr.desc.rigid_ranges = None
    st, msg = _create_desc_status(arr)
    assert st == 1 and "rigid_ranges" in msg
    arr = bp.SceneArrays(sc)
    arr.rranges[1, 0] = arr.rranges[0, 1] - 1  # second body starts inside the first
    st, msg = _create_desc_status(arr)
    assert st == 1 and "overlap" in msg
    arr = bp.SceneArrays(sc)
    arr.rranges[:] = arr.rranges[::-1].copy()  # disjoint but listed out of order: accepted
    st, msg = _create_desc_status(arr)
    assert st != 1, msg


def test_create_rejects_negative_masses():
    """A partner's mass travels in the lambda pass's record with the partner's
    fluid flag in its sign, so every mass must be >= 0 (and not NaN); the
    check runs before any device call (BP_ERR_INVALID_ARG)."""
    for bad in (-1.0, float("nan")):
        arr = bp.SceneArrays(small_scene(grid_dims=(64, 64, 64)))
        arr.masses[3] = bad
        st, msg = _create_desc_status(arr)
        assert st == 1 and "masses" in msg, (bad, st, msg)
    arr = bp.SceneArrays(small_scene(grid_dims=(64, 64, 64)))
    arr.masses[3] = 0.0  # a massless particle is accepted
    st, msg = _create_desc_status(arr)
    assert st != 1, msg
