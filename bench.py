"""Benchmark: particle-substeps/s of the gfx950 particle solver substep.

A step is one substep (grid build + sort + 4 solver iterations + post passes)
of the plain solver over a synthetic scene resident in HBM.  Prints one JSON
line.  The default workload is the 282,624-particle mixed-phase scene (fluid +
64 shape-matched boxes + 128x128 cloth under wind); the same line also carries
the 4,194,304-particle large block, the HBM-roofline case, and the 512-scene
batch (2k..64k particles per scene) against the plain solver looped over the
same scenes, and the adjoint/forward time ratio of the 64-scene
differentiable rollout (32 substeps, checkpoints every 8).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "brass-pan_amd"))

import numpy as np  # noqa: E402

import brasspan as bp  # noqa: E402
import scenes  # noqa: E402

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E

# Algorithmic bytes per particle of each kernel family (each array read or
# written once per pass; FP32 vec3 = 12 B, scalars 4 B; cell-range table reads
# 8 B per occupied cell, `cpn` = occupied cells / particles).
ALG_BYTES = {
    "predict_hash": lambda cpn: 12 + 12 + 4 + 8 * cpn,           # R x, v; W key; clear old ranges
    "radix_sort": lambda cpn: 16,                                # R+W key, value (single-pass bound)
    "cell_ranges": lambda cpn: 4 + 8 * cpn,                      # R key; W start/end
    "reorder": lambda cpn: 104,                                  # R perm + 6 fields; W 6 fields + inverse map
    "fluid_lambda": lambda cpn: 28 + 8 * cpn,                    # R x*, m, phase; W lambda, rho
    "fluid_position": lambda cpn: 36 + 8 * cpn,                  # R x*, lambda, m, phase; W x*
    "particle_particle": lambda cpn: 44 + 8 * cpn,               # R x*, x, w, phase; W x*
    "omega_normal": lambda cpn: 56 + 8 * cpn,                    # R x, v, rho, phase; W omega, n
    "vorticity_tension": lambda cpn: 68 + 8 * cpn,               # R x, omega, v, n, rho, phase; W v
    "xsph": lambda cpn: 40 + 8 * cpn,                            # R x, v, phase; W v
    "finish": lambda cpn: 48,                                    # R x*, x; W v, x
    "plane_collide": lambda cpn: 36,                             # R x*, x; W x*
}


def build_workload(name):
    builder, pbuilder, dt, substeps = scenes.WORKLOADS[name]
    sc = builder()
    return sc, pbuilder(), dt / substeps


def time_plain(sc, prm, subdt, steps, warmup):
    """Workgroup sizes from bp_solver_autotune (which restores the state it
    started from), then `warmup` untimed and `steps` timed substeps."""
    s = bp.Solver(sc, prm)
    s.block_sizes = s.autotune(subdt, 2)
    for _ in range(warmup):
        s.step(1, subdt)
    s.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(1, subdt)
    s.sync()
    wall = time.perf_counter() - t0
    return s, wall


def profiled_traffic(family, workload):
    """HBM bytes per launch of `family` on `workload` from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summaries (profiles/), or None."""
    import glob
    import re

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.txt"))):
        for line in open(path):
            m = re.match(rf"traffic_bytes_per_launch {family} {workload}: (\d+)", line)
            if m:
                return int(m.group(1))
    return None


def kernel_roofline(s, subdt, substeps=5, workload=None):
    """Per-family HIP-event timing on the solver's stream over `substeps`."""
    s.enable_kernel_timing(True)
    s.timed_substeps(substeps, subdt)
    times = s.kernel_times()
    s.enable_kernel_timing(False)
    keys = s.debug(2)
    cpn = float(len(np.unique(keys))) / float(len(keys))
    fam = max(times, key=lambda k: times[k][0])
    total_ms, launches = times[fam]
    avg_ms = total_ms / launches
    per = ALG_BYTES.get(fam, lambda c: float("nan"))(cpn)
    alg_bytes = per * s.n
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    return {
        "kernel": fam, "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": profiled_traffic(fam, workload),
        "alg_bytes_per_launch": int(alg_bytes), "avg_launch_ms": round(avg_ms, 4),
        "alg_bytes_per_particle": round(per, 2), "occupied_cells_per_particle": round(cpn, 4),
        "launches": launches,
    }, {k: [round(v[0], 3), v[1]] for k, v in sorted(times.items(), key=lambda kv: -kv[1][0])}


def _oracle_rate(o, n, subdt, budget_s, max_substeps=50):
    o.step(1, subdt)  # warm-up substep (untimed)
    done = 0
    t0 = time.perf_counter()
    while True:
        o.step(1, subdt)
        done += 1
        el = time.perf_counter() - t0
        if el >= budget_s or done >= max_substeps:
            break
    return n * done / el, done, el


def host_threads():
    """Threads this process may use: the OpenMP setting, else the CPUs in its affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(sc, prm, subdt, budget_s=10.0):
    """The FP32 oracle on a bounded sample of the same scene, in two forms:
    gather form (the GPU's pair order) over all host threads with OpenMP, and
    the reference's own order (scatter form) on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc

    n = sc.num_particles()
    threads = orc.set_threads(host_threads())
    o = orc.Oracle(bp.SceneArrays(sc), prm, fp64=False, mode=orc.GATHER)
    rate_mt, done_mt, el_mt = _oracle_rate(o, n, subdt, budget_s)
    del o
    orc.set_threads(1)
    o = orc.Oracle(bp.SceneArrays(sc), prm, fp64=False, mode=orc.SCATTER)
    rate_1, done_1, el_1 = _oracle_rate(o, n, subdt, budget_s)
    return ({"value": round(rate_mt, 1), "unit": "particle-substeps/s", "cores": threads, "kind": "port",
             "sample": f"{done_mt} substeps after 1 warm-up substep, {el_mt:.1f} s, FP32 oracle in gather form "
                       f"(the GPU's pair order) over {threads} OpenMP threads"},
            {"value": round(rate_1, 1), "unit": "particle-substeps/s", "cores": 1, "kind": "port",
             "sample": f"{done_1} substeps after 1 warm-up substep, {el_1:.1f} s, single-threaded FP32 oracle "
                       f"in reference order"})


def hip_stream():
    """A new HIP stream at the default priority (a raw handle for the C ABI)."""
    import ctypes as C

    hip = bp.hip_runtime()
    h = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(h)) == 0
    return h.value


def time_batched(frames=3, warmup=1, num=512, graphs=True):
    """The 512-scene batched workload (2 substeps per frame of 1/60 s) in one
    bp_batch_create handle with graphs and tuned block sizes, against the
    plain solver looped over the same scenes on one stream, eager and with
    graphs.  The batch must equal the eager loop bit for bit."""
    scs, prms = scenes.batched_scenes(num)
    dt, sub = 1.0 / 60.0, 2
    n = sum(sc.num_particles() for sc in scs)
    bs = bp.BatchSolver(scs, prms)
    tune = bs.autotune(dt / sub, 2)
    bs.set_graph(graphs)
    for _ in range(warmup):
        bs.step(sub, dt)
    bs.sync()
    t0 = time.perf_counter()
    for _ in range(frames):
        bs.step(sub, dt)
    bs.sync()
    wall_b = time.perf_counter() - t0
    bpos, _ = bs.state_original_order()
    ranges = bs.ranges
    bs.close()

    stream = hip_stream()
    loop = [bp.Solver(sc, prm, stream=stream) for sc, prm in zip(scs, prms)]

    def run_loop(count):
        for _ in range(count):
            for s in loop:
                s.step(sub, dt)
        loop[-1].sync()

    run_loop(warmup)
    t0 = time.perf_counter()
    run_loop(frames)
    wall_e = time.perf_counter() - t0
    same = all(np.array_equal(bpos[f:f + c], s.state_original_order()[0]) for (f, c), s in zip(ranges, loop))
    for s in loop:
        s.set_graph(graphs)
    run_loop(warmup)
    t0 = time.perf_counter()
    run_loop(frames)
    wall_g = time.perf_counter() - t0
    for s in loop:
        s.close()
    ps = n * sub * frames
    best_loop = min(wall_e, wall_g)
    return {
        "value": round(ps / wall_b, 1), "unit": "particle-substeps/s", "ms_per_substep": round(wall_b / (frames * sub) * 1e3, 3),
        "plain_loop_eager": round(ps / wall_e, 1), "plain_loop_graph": round(ps / wall_g, 1),
        "speedup_vs_plain_loop": round(best_loop / wall_b, 2),
        "bitexact_vs_plain_loop": bool(same),
        "config": {"workload": f"batched_{num}_scenes", "scenes": num, "particles": n,
                   "particles_min": min(c for _, c in ranges), "particles_max": max(c for _, c in ranges),
                   "iterations": prms[0].iterations, "substeps_per_frame": sub, "frames": frames,
                   "warmup_frames": warmup, "block_sizes": tune},
    }


def time_adjoint(num=64, graphs=True):
    """The differentiable workload: `num` scenes of 16,384 fluid particles and
    a 64 x 64 cloth (scenes.adjoint_scenes, seed 5), 32 substeps of 1/120 s,
    checkpoints every 8.  Loss = mean |x_T - x_target|^2 with x_target from a
    seeded forward run (wind and XSPH c changed, v0 perturbed).  Times the
    checkpointed forward and bp_backward (replay + adjoint) on the GPU."""
    scs, prms = scenes.adjoint_scenes(num)
    K, dt, every = scenes.ADJOINT_SUBSTEPS, scenes.ADJOINT_SUBDT, scenes.ADJOINT_EVERY
    n = sum(sc.num_particles() for sc in scs)
    rng = np.random.default_rng(5)
    # target: seeded forward run with other parameters and initial velocities
    tprms = []
    for p in prms:
        q = bp.bp_params.from_buffer_copy(p)
        q.wind[0] += 1.0
        q.wind[2] -= 0.5
        q.xsph_c *= 1.5
        tprms.append(q)
    ts = bp.BatchSolver(scs, tprms)
    pos, vel, mp = ts.get_state()
    ts.set_state(pos, rng.normal(0.0, 0.05, vel.shape).astype(np.float32), mp)
    for _ in range(K):
        ts.step(1, dt)
    target = ts.state_original_order()[0].astype(np.float64)
    ts.close()

    s = bp.BatchSolver(scs, prms)
    tune = s.autotune(dt, 2)
    pos0, vel0, mp0 = s.get_state()

    def run_once():
        s.set_state(pos0, vel0, mp0)
        s.sync()
        t0 = time.perf_counter()
        s.forward_checkpointed(K, dt, every)
        s.sync()
        t_f = time.perf_counter() - t0
        xT = s.state_original_order()[0].astype(np.float64)
        gx = (2.0 * (xT - target) / n).astype(np.float32)
        t0 = time.perf_counter()
        gx0, gv0, gth = s.backward(gx)
        t_b = time.perf_counter() - t0
        return t_f, t_b, float(np.sum((xT - target) ** 2) / n), gth, gv0

    run_once()  # warm-up (allocates checkpoints and tape)
    t_f, t_b, loss, gth, gv0 = run_once()
    # the plain solver over the same substeps (graphs on), for reference
    s.set_state(pos0, vel0, mp0)
    s.set_graph(graphs)
    s.step(2, 2 * dt)
    s.set_state(pos0, vel0, mp0)
    s.sync()
    t0 = time.perf_counter()
    s.step(K, K * dt)
    s.sync()
    t_plain = time.perf_counter() - t0
    s.close()
    return {
        "adjoint_over_forward": round(t_b / t_f, 3), "forward_ms": round(t_f * 1e3, 2),
        "backward_ms": round(t_b * 1e3, 2), "plain_forward_ms": round(t_plain * 1e3, 2),
        "adjoint_over_plain_forward": round(t_b / t_plain, 3),
        "particle_substeps_per_s_backward": round(n * K / t_b, 1), "loss": loss,
        "grad_theta_scene0": [float(x) for x in gth[0]], "grad_v0_norm": float(np.linalg.norm(gv0)),
        "config": {"workload": f"adjoint_{num}_scenes", "scenes": num, "particles": n, "substeps": K,
                   "substep_dt": dt, "checkpoint_every": every, "iterations": prms[0].iterations,
                   "block_sizes": tune},
    }


def copy_bandwidth(mib=1024, reps=5):
    """Device-to-device copy of `mib` MiB timed with HIP events (best of
    `reps`): a STREAM-style check of the HBM peak, counting the bytes read
    plus the bytes written."""
    import ctypes as C

    hip = bp.hip_runtime()
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    e0, e1 = C.c_void_p(), C.c_void_p()
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    best = float("inf")
    for _ in range(reps + 1):  # the first copy is a warm-up
        hip.hipEventRecord(e0, None)
        hip.hipMemcpyAsync(b, a, C.c_size_t(n), 3, None)  # hipMemcpyDeviceToDevice
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        best = min(best, ms.value) if _ else best
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return {"bytes": 2 * n, "ms": round(best, 4), "GBs": round(2 * n / (best * 1e-3) / 1e9, 1)}


def copy_calibration(mib):
    """A device-to-device copy of `mib` MiB through the HIP runtime: the known
    traffic (mib read + mib written) that scales rocprof FETCH_SIZE and
    WRITE_SIZE readings of the solver kernels."""
    import ctypes as C

    hip = bp.hip_runtime()
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(3):
        hip.hipMemcpy(b, a, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="mixed_phase_283k", choices=sorted(scenes.WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the 4M large-block measurement")
    ap.add_argument("--no-batched", action="store_true", help="skip the 512-scene batched measurement")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward measurement")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every kernel (no hipGraph replay; for kernel-trace profiling)")
    ap.add_argument("--leg", choices=("all", "batched", "adjoint"), default="all",
                    help="batched / adjoint: run only that measurement and print its object (for profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device (rocprof counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU slab sharding is a design, not part of this build)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    if args.leg == "batched":
        print(json.dumps({"batched_512": time_batched(graphs=not args.no_graph)}), flush=True)
        return
    if args.leg == "adjoint":
        print(json.dumps({"adjoint_64": time_adjoint(graphs=not args.no_graph)}), flush=True)
        return
    sc, prm, subdt = build_workload(args.workload)
    s, wall = time_plain(sc, prm, subdt, args.steps, args.warmup)
    n = s.n
    value = n * args.steps / wall
    roof, fam_times = kernel_roofline(s, subdt, workload=args.workload)
    s.close()
    out = {
        "metric": "particle-substeps/s (4 iters) plain vs 512-scene batched; adjoint/forward time ratio", "value": round(value, 1), "unit": "particle-substeps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded splitmix64 scene, no checkpoint)",
        "config": {"workload": args.workload, "particles": n, "iterations": prm.iterations,
                   "substep_dt": subdt, "distance_constraints": len(sc.distance_pairs),
                   "wind_faces": len(sc.faces), "rigid_bodies": len(sc.rigidbody_particle_id_ranges),
                   "block_sizes": s.block_sizes, "parallelism": "single GPU"},
        "roofline": roof,
        "kernel_ms": fam_times,
        "hbm_copy_check": copy_bandwidth(),
    }
    if not args.no_large and args.workload != "large_block_4m":
        # The large block's Akinci tension (gamma 0.5, no mass weighting,
        # fluid.cuh:436-489) drives it unstable after ~20 substeps; time
        # substeps 3-12 from the lattice.
        lsc, lprm, lsub = build_workload("large_block_4m")
        lsteps, lwarm = 10, 2
        ls, lwall = time_plain(lsc, lprm, lsub, lsteps, lwarm)
        lroof, ltimes = kernel_roofline(ls, lsub, substeps=2, workload="large_block_4m")
        out["large_block_4m"] = {"value": round(ls.n * lsteps / lwall, 1),
                                 "ms_per_step": round(lwall / lsteps * 1e3, 4), "steps": lsteps,
                                 "warmup": lwarm, "block_sizes": ls.block_sizes, "roofline": lroof,
                                 "kernel_ms": ltimes}
        ls.close()
    if not args.no_batched:
        out["batched_512"] = time_batched(graphs=not args.no_graph)
    if not args.no_adjoint:
        out["adjoint_64"] = time_adjoint(graphs=not args.no_graph)
    if not args.no_cpu_baseline:
        out["cpu_baseline"], out["cpu_baseline_reference_order"] = cpu_baseline(sc, prm, subdt)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
