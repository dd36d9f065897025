"""Benchmark: particle-substeps/s of the gfx950 particle solver substep.

A step is one substep (grid build + sort + 4 solver iterations + post passes)
over synthetic scenes resident in HBM.  Prints one JSON line.  The headline
workload is the metric's own configuration: 512 independent scenes of 2k..64k
particles (fluid block, 1-8 shape-matched boxes and a pinned cloth in wind
each; 9.5M particles in all) advanced by the batched solver, one launch per
pass for all scenes, replayed from a hipGraph.  The same line carries the
plain solver looped over the same scenes (the comparison the metric names),
the roofline of the batch's dominant kernel, a CPU baseline on a bounded
sample of the same scenes, the per-substep error against the FP32 and FP64
oracles, and sub-objects for the 16,384-particle dam break, the
282,624-particle mixed scene, the 4,194,304-particle block and the adjoint /
forward ratio of the 64-scene differentiable rollout.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
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


TUNE_FILE = None  # --tune-file: record the first tuning of each workload, replay it in later runs


def tune(s, subdt, key):
    """bp_solver_autotune (which restores the state it started from), or, with
    --tune-file, the choices a previous run recorded there for `key`, so that
    the runs of one profiling session time the same kernel forms."""
    saved = {}
    if TUNE_FILE and os.path.exists(TUNE_FILE):
        with open(TUNE_FILE) as f:
            saved = json.load(f)
    if key in saved:
        s.apply_tuning(saved[key])
        return saved[key]
    report = s.autotune(subdt, 2)
    if TUNE_FILE:
        saved[key] = report
        with open(TUNE_FILE, "w") as f:
            json.dump(saved, f)
    return report


def time_plain(sc, prm, subdt, steps, warmup, key="plain"):
    """`warmup` untimed and `steps` timed substeps, with workgroup sizes and
    kernel forms from the autotuner (tune()) run on the state the timed
    substeps start from (it restores that state)."""
    s = bp.Solver(sc, prm)
    for _ in range(warmup):
        s.step(1, subdt)
    s.block_sizes = tune(s, subdt, key)
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


def host_threads():
    """Threads this process may use: the OpenMP setting, else the CPUs in its affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def hip_stream():
    """A new HIP stream at the default priority (a raw handle for the C ABI)."""
    import ctypes as C

    hip = bp.hip_runtime()
    h = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(h)) == 0
    return h.value


BATCH_DT, BATCH_SUB = 1.0 / 60.0, 2  # frames of update(2, 1/60): substeps of 1/120 s


def run_substeps(s, count, subdt):
    """`count` substeps of `subdt`, two per call so that graph replay engages
    (a graph holds two substeps; 2 * subdt / 2 is exact in FP32)."""
    while count >= 2:
        s.step(2, 2 * subdt)
        count -= 2
    if count:
        s.step(1, subdt)


def batched_headline(scs, prms, steps, warmup, graphs=True):
    """The 512-scene batch in one bp_batch_create handle: `warmup` untimed
    substeps, then `steps` timed substeps replayed from the handle's hipGraph.
    Workgroup sizes, walk forms and iteration fusion come from
    bp_solver_autotune, run on the state the timed substeps start from (the
    passes' costs change as the scenes evolve) and restoring it; the last two
    warm-up substeps capture the graph."""
    subdt = BATCH_DT / BATCH_SUB
    bs = bp.BatchSolver(scs, prms)
    pre = max(warmup - 2, 0)
    run_substeps(bs, pre, subdt)
    choice = tune(bs, subdt, "batched_512_scenes")
    bs.set_graph(graphs)
    run_substeps(bs, warmup - pre, subdt)
    bs.sync()
    t0 = time.perf_counter()
    run_substeps(bs, steps, subdt)
    bs.sync()
    return bs, choice, time.perf_counter() - t0


def plain_loop(scs, prms, frames=3, graphs=True):
    """The plain solver looped over the same scenes on one stream (the
    per-scene loop the batched mode is measured against), eager and with
    per-solver graphs, over `frames` frames of update(2, 1/60) after one
    warm-up frame.  Its eager state must equal a batch stepped as far, bit
    for bit."""
    sub, dt = BATCH_SUB, BATCH_DT
    n = sum(sc.num_particles() for sc in scs)
    stream = hip_stream()
    loop = [bp.Solver(sc, prm, stream=stream) for sc, prm in zip(scs, prms)]

    def run_loop(count):
        for _ in range(count):
            for s in loop:
                s.step(sub, dt)
        loop[-1].sync()

    run_loop(1)
    t0 = time.perf_counter()
    run_loop(frames)
    wall_e = time.perf_counter() - t0
    ref = bp.BatchSolver(scs, prms)
    ref.set_graph(graphs)
    for _ in range(1 + frames):
        ref.step(sub, dt)
    bpos, _ = ref.state_original_order()
    same = all(np.array_equal(bpos[f:f + c], s.state_original_order()[0]) for (f, c), s in zip(ref.ranges, loop))
    ref.close()
    for s in loop:
        s.set_graph(graphs)
    run_loop(1)
    t0 = time.perf_counter()
    run_loop(frames)
    wall_g = time.perf_counter() - t0
    for s in loop:
        s.close()
    ps = n * sub * frames
    return {"eager": round(ps / wall_e, 1), "graph": round(ps / wall_g, 1), "unit": "particle-substeps/s",
            "substeps": sub * frames, "bitexact_vs_batch": bool(same)}


def time_batched(frames=3, warmup=1, num=512, graphs=True):
    """The batched leg alone (for profiling): the batch over `frames` frames
    and the plain loop over the same scenes."""
    scs, prms = scenes.batched_scenes(num)
    bs, choice, wall = batched_headline(scs, prms, BATCH_SUB * frames, BATCH_SUB * warmup, graphs)
    n = bs.n
    bs.close()
    ps = n * BATCH_SUB * frames
    loop = plain_loop(scs, prms, frames, graphs)
    return {"value": round(ps / wall, 1), "unit": "particle-substeps/s",
            "ms_per_substep": round(wall / (frames * BATCH_SUB) * 1e3, 3), "plain_loop": loop,
            "speedup_vs_plain_loop": round(ps / wall / max(loop["eager"], loop["graph"]), 2),
            "config": {"workload": f"batched_{num}_scenes", "scenes": num, "particles": n, "block_sizes": choice}}


def host_info():
    """The machine the line was measured on: CPU model and logical CPUs of the
    host, the GPU's name and architecture as the HIP runtime reports them."""
    cpu = None
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), None)
    except OSError:
        pass
    gpu = None
    try:
        gpu = bp.device_name()
    except Exception:  # noqa: BLE001 - a missing name does not void the line
        pass
    return {"cpu": cpu, "logical_cpus": os.cpu_count(), "gpu": gpu}


def cpu_baseline_batch(scs, prms, subdt, budget_s=15.0, ref_budget_s=5.0):
    """The FP32 oracle on a bounded sample of the batched workload: scenes in
    batch order, each one warm-up substep and two timed substeps, until the
    timed work reaches the budget.  Gather form (the GPU's pair order) over
    all host threads with OpenMP, and the reference's own order (scatter
    form) on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc

    def sample(mode, budget):
        done = el = 0.0
        used = 0
        for sc, prm in zip(scs, prms):
            o = orc.Oracle(bp.SceneArrays(sc), prm, fp64=False, mode=mode)
            o.step(1, subdt)
            t0 = time.perf_counter()
            o.step(1, subdt)
            o.step(1, subdt)
            el += time.perf_counter() - t0
            done += 2 * sc.num_particles()
            used += 1
            del o
            if el >= budget:
                break
        return done / el, used, int(done), el

    threads = orc.set_threads(host_threads())
    r_mt, u_mt, d_mt, e_mt = sample(orc.GATHER, budget_s)
    orc.set_threads(1)
    r_1, u_1, d_1, e_1 = sample(orc.SCATTER, ref_budget_s)
    orc.set_threads(host_threads())
    return ({"value": round(r_mt, 1), "unit": "particle-substeps/s", "cores": threads, "kind": "port",
             "sample": f"the first {u_mt} of the 512 scenes, 2 timed substeps each after 1 warm-up substep "
                       f"({d_mt} particle-substeps, {e_mt:.1f} s), FP32 oracle in gather form (the GPU's pair "
                       f"order) over {threads} OpenMP threads"},
            {"value": round(r_1, 1), "unit": "particle-substeps/s", "cores": 1, "kind": "port",
             "sample": f"the first {u_1} scenes, 2 timed substeps each ({d_1} particle-substeps, {e_1:.1f} s), "
                       f"single-threaded FP32 oracle in reference order"})


def substep_errors(sc, prm, subdt, substeps=3, h=0.1):
    """Per-substep error of the GPU: from the same state as the FP32 oracle in
    reference order and its FP64 twin, one substep each, `substeps` times;
    max and RMS of |x_gpu - x_oracle| over particles, in units of h."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc

    arr = bp.SceneArrays(sc)
    o32 = orc.Oracle(arr, prm, fp64=False, mode=orc.SCATTER)
    o64 = orc.Oracle(arr, prm, fp64=True, mode=orc.SCATTER)
    g = bp.Solver(sc, prm)
    rows = []
    for _ in range(substeps):
        p, v, m = o32.get_state()
        o64.set_state(p, v, m)
        g.set_state(p.astype(np.float32), v.astype(np.float32), m)
        if o32.num_rigid:
            q, _ = o32.rigid_state()
            o64.set_rigid_state(q)
            g.set_rigid_state(q.astype(np.float32))
        g.step(1, subdt)
        o32.step(1, subdt)
        o64.step(1, subdt)
        x = g.state_original_order()[0].astype(np.float64)
        e32 = np.linalg.norm(x - o32.state_original_order()[0], axis=1) / h
        e64 = np.linalg.norm(x - o64.state_original_order()[0], axis=1) / h
        rows.append((e32.max(), np.sqrt((e32 ** 2).mean()), e64.max(), np.sqrt((e64 ** 2).mean())))
    g.close()
    r = np.array(rows)
    return {"max_vs_fp32": float(r[:, 0].max()), "rms_vs_fp32": float(r[:, 1].max()),
            "max_vs_fp64": float(r[:, 2].max()), "rms_vs_fp64": float(r[:, 3].max()), "unit": "h",
            "substeps": substeps, "particles": sc.num_particles()}


def horizon_errors(sc, prm, subdt, substeps=10, h=0.1):
    """Free-running error over `substeps` substeps from the scene's initial
    state: max |x_gpu - x_oracle| / h after each substep, against the FP32
    oracle (reference order) and the FP64 oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc

    arr = bp.SceneArrays(sc)
    o32 = orc.Oracle(arr, prm, fp64=False, mode=orc.SCATTER)
    o64 = orc.Oracle(arr, prm, fp64=True, mode=orc.SCATTER)
    g = bp.Solver(sc, prm)
    e32, e64 = [], []
    for _ in range(substeps):
        g.step(1, subdt)
        o32.step(1, subdt)
        o64.step(1, subdt)
        x = g.state_original_order()[0].astype(np.float64)
        e32.append(float(np.abs(x - o32.state_original_order()[0]).max() / h))
        e64.append(float(np.abs(x - o64.state_original_order()[0]).max() / h))
    g.close()
    return {"max_vs_fp32": [float("%.3g" % e) for e in e32], "max_vs_fp64": [float("%.3g" % e) for e in e64],
            "unit": "h", "particles": sc.num_particles()}


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
    choice = tune(s, dt, "adjoint_64_scenes")
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
                   "block_sizes": choice},
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


def plain_leg(name, steps, warmup, roof_substeps=5):
    """A plain-solver workload: timing, roofline of its dominant kernel."""
    sc, prm, subdt = build_workload(name)
    s, wall = time_plain(sc, prm, subdt, steps, warmup, key=name)
    roof, times = kernel_roofline(s, subdt, substeps=roof_substeps, workload=name)
    out = {"value": round(s.n * steps / wall, 1), "unit": "particle-substeps/s",
           "ms_per_step": round(wall / steps * 1e3, 4), "steps": steps, "warmup": warmup, "particles": s.n,
           "block_sizes": s.block_sizes, "roofline": roof, "kernel_ms": times}
    return out, s, sc, prm, subdt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40, help="timed substeps of the 512-scene batch")
    ap.add_argument("--warmup", type=int, default=10, help="untimed substeps before them")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the per-substep errors against the oracle")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-solver sub-objects (c1, c2, c3)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward measurement")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every kernel (no hipGraph replay; for kernel-trace profiling)")
    ap.add_argument("--leg", choices=("all", "headline", "batched", "adjoint", "plain"), default="all",
                    help="headline (the batch's timing and roofline) / batched (with the plain loop) / adjoint / "
                         "plain (--workload): run only that measurement and print its object (for profiling)")
    ap.add_argument("--workload", default="mixed_phase_283k", choices=sorted(scenes.WORKLOADS),
                    help="the plain workload of --leg plain")
    ap.add_argument("--tune-file", default=None, metavar="JSON",
                    help="record each workload's autotuner choices in JSON on the first run and reuse them on "
                         "later runs (so the kernel trace and the counter runs of a profile time the same forms)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device (rocprof counter calibration)")
    args = ap.parse_args()
    global TUNE_FILE
    TUNE_FILE = args.tune_file
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
    if args.leg == "plain":
        steps = 10 if args.workload == "large_block_4m" else 50
        out, s, *_ = plain_leg(args.workload, steps, 2 if args.workload == "large_block_4m" else 10)
        s.close()
        print(json.dumps({args.workload: out}), flush=True)
        return

    # headline: the 512-scene batch
    scs, prms = scenes.batched_scenes(512)
    subdt = BATCH_DT / BATCH_SUB
    graphs = not args.no_graph
    bs, choice, wall = batched_headline(scs, prms, args.steps, args.warmup, graphs)
    n = bs.n
    value = n * args.steps / wall
    roof, fam_times = kernel_roofline(bs, subdt, substeps=2, workload="batched_512_scenes")
    bs.close()
    if args.leg == "headline":
        print(json.dumps({"batched_512_scenes": {"value": round(value, 1), "ms_per_step": round(wall / args.steps * 1e3, 4),
                                                 "roofline": roof, "kernel_ms": fam_times, "block_sizes": choice}}),
              flush=True)
        return
    loop = plain_loop(scs, prms, frames=3, graphs=graphs)
    best_loop = max(loop["eager"], loop["graph"])
    ranges_n = [sc.num_particles() for sc in scs]
    out = {
        "metric": "particle-substeps/s (4 iters) plain vs 512-scene batched; adjoint/forward time ratio",
        "value": round(value, 1), "unit": "particle-substeps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded splitmix64 scenes, no checkpoint)",
        "config": {"workload": "batched_512_scenes", "scenes": len(scs), "particles": n,
                   "particles_min": min(ranges_n), "particles_max": max(ranges_n),
                   "iterations": prms[0].iterations, "substep_dt": subdt,
                   "distance_constraints": sum(len(sc.distance_pairs) for sc in scs),
                   "wind_faces": sum(len(sc.faces) for sc in scs),
                   "rigid_bodies": sum(len(sc.rigidbody_particle_id_ranges) for sc in scs),
                   "graph_replay": graphs, "block_sizes": choice, "parallelism": "single GPU"},
        "roofline": roof,
        "kernel_ms": fam_times,
        "plain_loop": loop,
        "speedup_vs_plain_loop": round(value / best_loop, 2),
        "host": host_info(),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"], out["cpu_baseline_reference_order"] = cpu_baseline_batch(scs, prms, subdt)
    if not args.no_accuracy:
        acc = {"batched_512_scenes": {f"scene_{k}": substep_errors(scs[k], prms[k], subdt) for k in (0, 1)}}
        dsc, dprm, dsub = build_workload("dam_break_16k")
        acc["dam_break_16k_horizon"] = horizon_errors(dsc, dprm, dsub)
        out["accuracy"] = acc
    del scs, prms
    if not args.no_plain:
        out["dam_break_16k"], s, *_ = plain_leg("dam_break_16k", 50, 10)
        s.close()
        out["mixed_phase_283k"], s, msc, mprm, msub = plain_leg("mixed_phase_283k", 50, 10)
        s.close()
        if not args.no_accuracy:
            out["mixed_phase_283k"]["accuracy"] = substep_errors(msc, mprm, msub, substeps=2)
        del msc
        # The large block's Akinci tension (gamma 0.5, no mass weighting,
        # fluid.cuh:436-489) drives it unstable after ~20 substeps; time
        # substeps 3-12 from the lattice.
        out["large_block_4m"], s, *_ = plain_leg("large_block_4m", 10, 2, roof_substeps=2)
        s.close()
    if not args.no_adjoint:
        out["adjoint_64"] = time_adjoint(graphs=graphs)
    out["hbm_copy_check"] = copy_bandwidth()
    # the same check with this library's streaming copy kernel instead of the runtime's copy
    out["hbm_copy_check_kernel"] = bp.stream_copy_bandwidth(1 << 30, 5)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
