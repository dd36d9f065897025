"""HIP solver against the committed golden vectors (tests/golden/*.npz), through
the C ABI.  No oracle runs here: the expected values are data.

Tolerances (positions in scene units, h = 0.1):
  - integer work (cell keys, sort permutation, cell ranges): bit-exact;
  - one substep from the initial state: 2e-5 against the GPU-order oracle,
    2e-3 against the reference-order oracle; density and lambda 1e-5
    relative to the largest value;
  - the 10-substep dam-break horizon: 2e-5 against FP32 (GPU order) and
    against FP64 at every substep;
  - the mixed-phase scene (rigid boxes, fluid, pinned cloth in wind): 1e-4
    against FP32 and FP64 at each of 4 substeps, rigid rotations 1e-4.
"""
import os

import numpy as np
import pytest

import brasspan as bp

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

import importlib.util  # noqa: E402

_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
mg = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(mg)


def load(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"))


def case(msg):
    print(f"[golden] {msg}", flush=True)


@pytest.mark.parametrize("name", sorted(mg.GRID_CASES))
def test_gpu_grid_matches_golden(name):
    f = load("grid")
    _, _, cs, dims, _ = mg.GRID_CASES[name]
    case(f"grid {name}")
    out = bp.neighbor_grid(f[f"{name}_xs"], cs, dims)
    for a, key in zip(out, ("keys", "perm", "start", "end")):
        assert np.array_equal(a, f[f"{name}_{key}"]), key


def test_gpu_dam_break_one_substep_matches_golden():
    f = load("dam_break_512")
    sc, prm = mg.dam_break_512()
    assert np.array_equal(bp.SceneArrays(sc).positions, f["in_positions"])
    case("dam break 512, one substep")
    s = bp.Solver(sc, prm)
    s.step(1, mg.DT)
    assert np.array_equal(s.debug(2), f["step1_keys"])
    assert np.array_equal(s.debug(3), f["step1_perm"])
    rho, lam = s.debug(0), s.debug(1)
    assert np.abs(rho - f["step1_density"]).max() <= 1e-5 * np.abs(f["step1_density"]).max()
    assert np.abs(lam - f["step1_lambda"]).max() <= 1e-5 * max(np.abs(f["step1_lambda"]).max(), 1e-12)
    p, v = s.state_original_order()
    eg = np.abs(p - f["step1_gather_positions"]).max()
    es = np.abs(p - f["step1_scatter_positions"]).max()
    print(f"dam break 512 substep 1: vs gather {eg:.3e}, vs scatter {es:.3e}")
    assert eg <= 2e-5 and es <= 2e-3
    s.close()


def test_gpu_dam_break_horizon_matches_golden():
    f = load("dam_break_512")
    sc, prm = mg.dam_break_512()
    case("dam break 512, 10-substep horizon")
    s = bp.Solver(sc, prm)
    for k in range(mg.HORIZON):
        s.step(1, mg.DT)
        p, _ = s.state_original_order()
        e32 = np.abs(p - f["horizon_fp32_gather"][k]).max()
        e64 = np.abs(p - f["horizon_fp64"][k]).max()
        r64 = float(np.sqrt(((p - f["horizon_fp64"][k]) ** 2).mean()))
        print(f"substep {k + 1}: vs FP32 max {e32:.3e} | vs FP64 max {e64:.3e} rms {r64:.3e}")
        assert e32 <= 2e-5 and e64 <= 2e-5, k
    s.close()


def test_gpu_mixed_phase_matches_golden():
    f = load("mixed_small")
    sc, prm = mg.mixed_small()
    assert np.array_equal(bp.SceneArrays(sc).positions, f["in_positions"])
    case("mixed phase 704, 4 substeps")
    s = bp.Solver(sc, prm)
    for k in range(mg.MIXED_SUBSTEPS):
        s.step(1, mg.DT)
        p, _ = s.state_original_order()
        e32 = np.abs(p - f["traj_fp32_gather"][k]).max()
        e64 = np.abs(p - f["traj_fp64"][k]).max()
        q, _ = s.rigid_state()
        eq = np.abs(q - f["quats_fp32_gather"][k]).max()
        print(f"substep {k + 1}: vs FP32 {e32:.3e}, vs FP64 {e64:.3e}, rigid rotations {eq:.3e}")
        assert e32 <= 1e-4 and e64 <= 1e-4 and eq <= 1e-4, k
    s.close()
