"""Writes the golden vectors under tests/golden/ from the host oracle.

The reference ships no tests and no golden vectors, so these fixtures come
from this build's own oracle (oracle/bp_oracle.cpp, pinned independently by
tests/test_oracle_pins.py).  They freeze, as data:
  - the inputs the seeded scene generator produces, so a change to the
    generator shows up as a changed input;
  - the neighbour grid (cell keys, stable sort permutation, cell ranges;
    grid.cuh:30-99, particlesolver.cuh:721-738) on random and lattice inputs;
  - one substep of the 512-particle dam break from its initial state (keys,
    permutation, density, lambda, positions, velocities), in the GPU's
    summation order (gather) and in the reference's (scatter);
  - the positions of a 10-substep rollout of the same scene in FP32 (gather
    and scatter) and FP64 (scatter);
  - a small mixed-phase scene (rigid boxes, fluid, a pinned cloth in wind)
    over 4 substeps in FP32 (gather) and FP64, with the rigid rotations.

tests/test_golden.py checks that the oracle still reproduces them (CPU) and
tests/test_gpu_golden.py checks the HIP solver against them (GPU).

Run from the repository root after building the oracle:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("brass-pan_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import brasspan as bp  # noqa: E402  (scene descriptors only; no device use)
import oracle_py as orc  # noqa: E402
import scenes  # noqa: E402

DT = 1.0 / 120.0
HORIZON = 10
MIXED_SUBSTEPS = 4

# the scenes the fixtures are made from (the tests rebuild them the same way)
GRID_CASES = {
    # name: (seed, n, cell size per axis, grid dims, coordinate range)
    "random_dup_16": (11, 2000, (0.1, 0.1, 0.1), (16, 16, 16), (-0.5, 2.1)),
    "aniso_64x32x8": (12, 3000, (0.05, 0.1, 0.3), (64, 32, 8), (-1.0, 4.0)),
}


def dam_break_512():
    """8x8x8 jittered lattice; XSPH c = 0.0002, the reference's own literal
    (particlesolver.cuh:1055).  At c = 0.01 this block splashes across the box
    within 4 substeps and a 10-substep horizon only measures chaos."""
    return scenes.dam_break(dims=(8, 8, 8), grid=(64, 64, 64)), scenes.dam_break_params(xsph_c=0.0002)


def mixed_small():
    sc = scenes.mixed_phase(fluid_dims=(8, 8, 8), num_boxes=2, cloth_n=8, grid=(128, 128, 128))
    return sc, scenes.mixed_phase_params()


def grid_input(seed, n, lo, hi):
    rng = np.random.default_rng(seed)
    xs = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    xs[: n // 8] = xs[n // 8: 2 * (n // 8)]  # repeated positions: stability matters
    return xs


def scene_inputs(sc):
    a = bp.SceneArrays(sc)
    return {"in_positions": a.positions, "in_masses": a.masses, "in_phases": a.phases}


def make_grid():
    out = {}
    for name, (seed, n, cs, dims, (lo, hi)) in GRID_CASES.items():
        xs = grid_input(seed, n, lo, hi)
        keys, perm, start, end = orc.neighbor_grid(xs, cs, dims)
        out.update({f"{name}_xs": xs, f"{name}_keys": keys, f"{name}_perm": perm,
                    f"{name}_start": start, f"{name}_end": end})
    return out


def make_dam_break():
    sc, prm = dam_break_512()
    arr = bp.SceneArrays(sc)
    out = scene_inputs(sc)
    for mode, tag in ((orc.GATHER, "gather"), (orc.SCATTER, "scatter")):
        o = orc.Oracle(arr, prm, fp64=False, mode=mode)
        o.step(1, DT)
        p, v = o.state_original_order()
        out[f"step1_{tag}_positions"] = p.astype(np.float32)
        out[f"step1_{tag}_velocities"] = v.astype(np.float32)
        if tag == "gather":
            out["step1_keys"] = o.debug(2)
            out["step1_perm"] = o.debug(3)
            out["step1_density"] = o.debug(0).astype(np.float32)
            out["step1_lambda"] = o.debug(1).astype(np.float32)
    for fp64, mode, tag in ((False, orc.GATHER, "fp32_gather"), (False, orc.SCATTER, "fp32_scatter"),
                            (True, orc.SCATTER, "fp64")):
        o = orc.Oracle(arr, prm, fp64=fp64, mode=mode)
        traj = []
        for _ in range(HORIZON):
            o.step(1, DT)
            traj.append(o.state_original_order()[0])
        out[f"horizon_{tag}"] = np.asarray(traj, np.float64 if fp64 else np.float32)
    return out


def make_mixed():
    sc, prm = mixed_small()
    arr = bp.SceneArrays(sc)
    out = scene_inputs(sc)
    for fp64, tag in ((False, "fp32_gather"), (True, "fp64")):
        o = orc.Oracle(arr, prm, fp64=fp64, mode=orc.GATHER if not fp64 else orc.SCATTER)
        traj, quats = [], []
        for _ in range(MIXED_SUBSTEPS):
            o.step(1, DT)
            traj.append(o.state_original_order()[0])
            quats.append(o.rigid_state()[0])
        out[f"traj_{tag}"] = np.asarray(traj, np.float64 if fp64 else np.float32)
        out[f"quats_{tag}"] = np.asarray(quats, np.float64)
    return out


def main():
    for name, fn in (("grid", make_grid), ("dam_break_512", make_dam_break), ("mixed_small", make_mixed)):
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **fn())
        print(f"wrote {os.path.relpath(path, ROOT)} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
