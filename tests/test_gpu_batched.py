"""GPU: the batched solver (bp_batch_create) against the plain solver run on
each scene by itself.

Scenes in a batch occupy disjoint cell-id ranges, so the stable sort keeps
every scene's particles in the order the plain solver gives them and every
pass sees the same neighbours in the same order: the batched state must equal
the looped plain state bit for bit.  Graph replay (bp_solver_set_graph) and
block sizes chosen by bp_solver_autotune launch the same kernels, so they
must not change a single bit either."""
import numpy as np
import pytest

import brasspan as bp
import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def case(msg):
    print(f"[gpu case] {msg}", flush=True)


def small_batch(num=6, seed=4):
    return scenes.batched_scenes(num=num, seed=seed, lo=2048, hi=8192)


def test_batch_scene_ranges():
    scs, prms = small_batch()
    bs = bp.BatchSolver(scs, prms)
    first = 0
    for k, sc in enumerate(scs):
        assert bs.ranges[k] == (first, sc.num_particles())
        first += sc.num_particles()
    assert bs.n == first
    f, c = bp.C.c_int(), bp.C.c_int()
    assert bs.lib.bp_batch_scene_range(bs.h, len(scs), bp.C.byref(f), bp.C.byref(c)) == 1  # BP_ERR_INVALID_ARG


@pytest.mark.parametrize("steps", [1, 3])
def test_batched_equals_plain_loop(steps):
    scs, prms = small_batch()
    case(f"batch of {len(scs)} scenes, n={sum(s.num_particles() for s in scs)}, {steps} frames x 2 substeps")
    bs = bp.BatchSolver(scs, prms)
    for _ in range(steps):
        bs.step(2, DT)
    bpos, bvel = bs.state_original_order()
    bq, bcm = bs.rigid_state()
    r0 = 0
    for k, (sc, prm) in enumerate(zip(scs, prms)):
        case(f"plain scene {k} n={sc.num_particles()}")
        s = bp.Solver(sc, prm)
        for _ in range(steps):
            s.step(2, DT)
        p, v = s.state_original_order()
        q, cm = s.rigid_state()
        f, c = bs.ranges[k]
        assert np.all(np.isfinite(p))
        assert np.array_equal(bpos[f:f + c], p), f"scene {k}: positions differ, max {np.abs(bpos[f:f + c] - p).max()}"
        assert np.array_equal(bvel[f:f + c], v), f"scene {k}: velocities differ"
        nr = s.num_rigid
        assert np.array_equal(bq[r0:r0 + nr], q) and np.array_equal(bcm[r0:r0 + nr], cm), f"scene {k}: rigid state"
        r0 += nr
        s.close()
    assert r0 == bs.num_rigid


def test_batch_of_tiny_scenes_equals_plain_loop():
    """Scenes of fewer than 64 particles put three or more scenes in one
    64-slot wave (12 + 27 + 2048 particles here), so a pass's per-wave loop
    over the distinct scenes of its lanes runs more than twice.  The batch
    must still give the plain solver's results bit for bit."""
    dims = [(2, 2, 3), (3, 3, 3), (16, 16, 8), (2, 3, 4), (4, 4, 4), (8, 8, 4)]
    scs = [scenes.dam_break(dims=d, seed=5 + k, grid=(32, 32, 32), box=3.2) for k, d in enumerate(dims)]
    prms = [scenes.dam_break_params() for _ in scs]
    case(f"tiny-scene batch n={sum(s.num_particles() for s in scs)}")
    bs = bp.BatchSolver(scs, prms)
    for _ in range(3):
        bs.step(2, DT)
    bpos, bvel = bs.state_original_order()
    for k, (sc, prm) in enumerate(zip(scs, prms)):
        s = bp.Solver(sc, prm)
        for _ in range(3):
            s.step(2, DT)
        p, v = s.state_original_order()
        f, c = bs.ranges[k]
        assert np.array_equal(bpos[f:f + c], p) and np.array_equal(bvel[f:f + c], v), f"scene {k} differs"
        s.close()


def test_graph_replay_equals_eager():
    scs, prms = small_batch(num=4, seed=11)
    case("batch graph vs eager")
    eager = bp.BatchSolver(scs, prms)
    graph = bp.BatchSolver(scs, prms)
    graph.set_graph(True)
    for frame in range(4):
        eager.step(2, DT)
        graph.step(2, DT)
        # host reads between replays, as a caller would do
        pe, ve = eager.state_original_order()
        pg, vg = graph.state_original_order()
        assert np.array_equal(pe, pg) and np.array_equal(ve, vg), f"frame {frame}: graph replay differs"
    # an odd substep count and a new dt force a recapture
    eager.step(3, DT / 2)
    graph.step(3, DT / 2)
    assert np.array_equal(eager.state_original_order()[0], graph.state_original_order()[0])


def test_graph_replay_plain_solver():
    sc = scenes.dam_break(dims=(16, 16, 8))
    prm = scenes.dam_break_params()
    case("plain graph vs eager")
    a, b = bp.Solver(sc, prm), bp.Solver(sc, prm)
    b.set_graph(True)
    for _ in range(5):
        a.step(2, DT)
        b.step(2, DT)
    assert np.array_equal(a.state_original_order()[0], b.state_original_order()[0])
    b.set_graph(False)
    a.step(2, DT)
    b.step(2, DT)
    assert np.array_equal(a.state_original_order()[0], b.state_original_order()[0])


def test_autotune_keeps_results():
    scs, prms = small_batch(num=4, seed=12)
    case("autotune")
    ref = bp.BatchSolver(scs, prms)
    tuned = bp.BatchSolver(scs, prms)
    report = tuned.autotune(DT / 2, 2)
    print("autotune report", report)
    passes = {"fluid_lambda", "fluid_position", "omega_normal", "vorticity_tension", "xsph", "particle_particle"}
    assert set(report) == passes | {"iteration_fusion", "lambda_walk", "position_walk", "omega_normal_walk",
                                    "vorticity_tension_walk", "lambda_stage_records", "omega_normal_stage_records",
                                    "position_partners", "lambda_stage_screen", "omega_normal_stage_screen"}
    assert report["position_partners"] in (2, 4)
    assert report["lambda_stage_screen"] in (4, 8) and report["omega_normal_stage_screen"] in (4, 8)
    assert report["lambda_walk"] in ("staged", "global") and report["omega_normal_walk"] in ("staged", "global")
    for walk, records in (("lambda_walk", "lambda_stage_records"),
                          ("omega_normal_walk", "omega_normal_stage_records")):
        # a staged walk names its LDS region size, the per-lane walk none
        assert report[records] in ((128, 256, 512) if report[walk] == "staged" else (0,))
    assert all(report[k] in (64, 128, 256, 512) for k in passes)
    assert report["iteration_fusion"] in (0, 1)
    assert report["position_walk"] in ("flat", "chunk") and report["vorticity_tension_walk"] in ("flat", "chunk")
    # the tuner advances its own copy of the state and restores it
    assert np.array_equal(ref.state_original_order()[0], tuned.state_original_order()[0])
    # and its choices do not change results
    ref.step(2, DT)
    tuned.step(2, DT)
    assert np.array_equal(ref.state_original_order()[0], tuned.state_original_order()[0])
    ref.step(2, DT)
    tuned.step(2, DT)
    assert np.array_equal(ref.state_original_order()[0], tuned.state_original_order()[0])


def test_full_batch_512_scenes_invariants():
    """The batched workload at full size: 512 scenes, 2k..64k particles each.
    Every particle stays finite and inside its own scene's box."""
    scs, prms = scenes.batched_scenes(512)
    n = sum(s.num_particles() for s in scs)
    case(f"512-scene batch n={n}")
    bs = bp.BatchSolver(scs, prms)
    bs.set_graph(True)
    for _ in range(5):
        bs.step(2, DT)
    pos, vel = bs.state_original_order()
    assert np.all(np.isfinite(pos)) and np.all(np.isfinite(vel))
    for k, sc in enumerate(scs):
        f, c = bs.ranges[k]
        lo, hi = sc.box
        p = pos[f:f + c]
        slack = 0.5 * sc.fluid_kernel_radius
        assert np.all(p >= lo - slack) and np.all(p <= hi + slack), f"scene {k} left its box"
