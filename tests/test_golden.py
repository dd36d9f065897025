"""The committed golden vectors (tests/golden/*.npz, written by
tests/golden/make_golden.py from the host oracle) against the scene
generator, an independent numpy restatement of the grid build, and the
oracle itself.  CPU only.

The reference ships no golden vectors, so the fixtures are this build's own.
What keeps them honest:
  - the grid vectors are recomputed here from the hash formula in numpy
    (grid.cuh:9-47: (x + 50000) / cellSize in FP32, truncation, & (D - 1)
    wrap, x-major address) and numpy's stable argsort, independently of the
    oracle's C++;
  - the physics vectors are checked here against what they must satisfy: the
    first substep's keys and permutation recomputed by the numpy hash from
    the stored inputs, positive densities, the FP32-vs-FP64 gap of the
    rollout, and agreement of the GPU and reference summation orders;
  - the oracle must still reproduce all of them, so a change to the oracle or
    to the scene generator fails here first.
"""
import os

import numpy as np
import pytest

import oracle_py as orc
import brasspan as bp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

import importlib.util  # noqa: E402

_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
mg = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(mg)


def load(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"))


def numpy_grid(xs, cs, dims):
    """Cell keys, stable permutation and [start, end) ranges by the hash formula."""
    cs = np.asarray(cs, np.float32)
    g = ((xs + np.float32(50000.0)) / cs).astype(np.int64)  # FP32 add and divide, truncation
    d = np.asarray(dims, np.int64)
    w = g & (d - 1)
    keys = (w[:, 0] * d[1] * d[2] + w[:, 1] * d[2] + w[:, 2]).astype(np.int32)
    perm = np.argsort(keys, kind="stable").astype(np.int32)
    sk = keys[perm]
    start = np.full(int(np.prod(d)), -1, np.int32)
    end = np.full(int(np.prod(d)), -1, np.int32)
    first = np.r_[True, sk[1:] != sk[:-1]]
    last = np.r_[sk[1:] != sk[:-1], True]
    start[sk[first]] = np.nonzero(first)[0]
    end[sk[last]] = np.nonzero(last)[0] + 1
    return sk, perm, start, end


@pytest.mark.parametrize("name", sorted(mg.GRID_CASES))
def test_grid_vectors_match_numpy_restatement(name):
    f = load("grid")
    seed, n, cs, dims, (lo, hi) = mg.GRID_CASES[name]
    xs = mg.grid_input(seed, n, lo, hi)
    assert np.array_equal(xs, f[f"{name}_xs"])
    keys, perm, start, end = numpy_grid(xs, cs, dims)
    assert np.array_equal(keys, f[f"{name}_keys"])
    assert np.array_equal(perm, f[f"{name}_perm"])
    assert np.array_equal(start, f[f"{name}_start"])
    assert np.array_equal(end, f[f"{name}_end"])


@pytest.mark.parametrize("name", sorted(mg.GRID_CASES))
def test_oracle_reproduces_grid_vectors(have_oracle, name):
    f = load("grid")
    seed, n, cs, dims, _ = mg.GRID_CASES[name]
    out = orc.neighbor_grid(f[f"{name}_xs"], cs, dims)
    for a, key in zip(out, ("keys", "perm", "start", "end")):
        assert np.array_equal(a, f[f"{name}_{key}"]), key


def test_scene_generator_reproduces_inputs():
    for name, make in (("dam_break_512", mg.dam_break_512), ("mixed_small", mg.mixed_small)):
        f = load(name)
        a = bp.SceneArrays(make()[0])
        assert np.array_equal(a.positions, f["in_positions"]), name
        assert np.array_equal(a.masses, f["in_masses"]), name
        assert np.array_equal(a.phases, f["in_phases"]), name


def test_dam_break_vectors_invariants():
    f = load("dam_break_512")
    sc, prm = mg.dam_break_512()
    # the keys are those of the predicted positions x + v dt + g dt^2, which
    # start at rest: recompute them from the inputs
    x = f["in_positions"]
    dt = np.float32(mg.DT)
    xp = x.copy()
    xp[:, 1] = xp[:, 1] + (np.float32(-9.8) * dt) * dt
    keys, perm, _, _ = numpy_grid(xp, sc.cell_size, sc.grid_dims)
    assert np.array_equal(keys, f["step1_keys"]) and np.array_equal(perm, f["step1_perm"])
    # every fluid particle has a positive density; lambda = -max(C, 0)/15000
    # is non-positive only if the Akinci clamp is on, so check its bound
    assert np.all(f["step1_density"] > 0)
    assert np.all(np.abs(f["step1_lambda"]) < 1.0)
    # the FP32 rollout stays within 2e-6 of FP64 over the 10 substeps
    gap = np.abs(f["horizon_fp32_gather"] - f["horizon_fp64"]).max(axis=(1, 2))
    assert np.all(gap < 2e-6), gap
    # reference order and GPU order agree to rounding over the horizon
    assert np.abs(f["horizon_fp32_gather"] - f["horizon_fp32_scatter"]).max() < 1e-6


def test_oracle_reproduces_dam_break_vectors(have_oracle):
    f = load("dam_break_512")
    sc, prm = mg.dam_break_512()
    arr = bp.SceneArrays(sc)
    o = orc.Oracle(arr, prm, fp64=False, mode=orc.GATHER)
    o.step(1, mg.DT)
    assert np.array_equal(o.debug(2), f["step1_keys"])
    assert np.array_equal(o.debug(3), f["step1_perm"])
    np.testing.assert_allclose(o.debug(0), f["step1_density"], rtol=1e-6)
    np.testing.assert_allclose(o.debug(1), f["step1_lambda"], rtol=1e-6, atol=1e-12)
    p, v = o.state_original_order()
    np.testing.assert_allclose(p, f["step1_gather_positions"], rtol=0, atol=1e-6)
    o64 = orc.Oracle(arr, prm, fp64=True, mode=orc.SCATTER)
    for k in range(mg.HORIZON):
        o64.step(1, mg.DT)
    np.testing.assert_allclose(o64.state_original_order()[0], f["horizon_fp64"][-1], rtol=0, atol=1e-9)


def test_oracle_reproduces_mixed_vectors(have_oracle):
    f = load("mixed_small")
    sc, prm = mg.mixed_small()
    o = orc.Oracle(bp.SceneArrays(sc), prm, fp64=False, mode=orc.GATHER)
    for k in range(mg.MIXED_SUBSTEPS):
        o.step(1, mg.DT)
        np.testing.assert_allclose(o.state_original_order()[0], f["traj_fp32_gather"][k], rtol=0, atol=1e-6)
    q, _ = o.rigid_state()
    np.testing.assert_allclose(q, f["quats_fp32_gather"][-1], rtol=0, atol=1e-6)
    assert np.abs(f["traj_fp32_gather"] - f["traj_fp64"]).max() < 5e-5


def test_oracle_thread_count_does_not_change_results(have_oracle):
    """The oracle's per-particle loops run over OpenMP threads (each iteration
    writes only its own slot; pair deliveries are atomic integer adds), so one
    thread and several give the same bits, in both summation orders."""
    sc, prm = mg.mixed_small()
    arr = bp.SceneArrays(sc)
    before = orc.set_threads()
    try:
        for mode in (orc.GATHER, orc.SCATTER):
            out = []
            for threads in (1, 4):
                orc.set_threads(threads)
                o = orc.Oracle(arr, prm, fp64=False, mode=mode)
                o.step(3, 3 * mg.DT)
                out.append(o.get_state())
            for a, b in zip(*out):
                assert np.array_equal(a, b), mode
    finally:
        orc.set_threads(before)
