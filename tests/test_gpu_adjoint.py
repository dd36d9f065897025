"""GPU: the hand-written adjoint (bp_forward_checkpointed / bp_backward)
against the FP64 host derivative of the same rollout.

The host derivative is the oracle's reference-order FP64 rollout evaluated in
forward-mode dual numbers (oracle/bp_oracle.cpp, or_rollout_tangent), which
tests/test_oracle_adjoint.py pins against central finite differences.  The
GPU rollout is FP32; on a small scene over 8-16 substeps it takes the same
branches and neighbour sets as the FP64 one, so its gradient must match the
FP64 derivative to ADJ_TOL relative error: every parameter component
(wind xyz, drag, lift, XSPH c) and the initial velocities along random
directions."""
import numpy as np
import pytest

import brasspan as bp
import oracle_py as orc
import scenes

pytestmark = pytest.mark.gpu

ADJ_TOL = 1e-3   # relative error of each gradient component vs the FP64 derivative
DT = scenes.ADJOINT_SUBDT


def case(msg):
    print(f"[gpu case] {msg}", flush=True)


def small(num=1, seed=5, fluid=(8, 8, 8), cloth=12):
    return scenes.adjoint_scenes(num, seed=seed, fluid_dims=fluid, cloth_n=cloth, box=3.2, grid=(32, 32, 32))


def seeded_target(sc, seed):
    rng = np.random.default_rng(seed)
    x0 = np.asarray(sc.positions, np.float64).reshape(-1, 3)
    return x0 + rng.normal(0.0, 0.05, x0.shape)


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-30)


def gpu_gradient(solver, target, norm, substeps, every):
    solver.forward_checkpointed(substeps, DT, every)
    xT, _ = solver.state_original_order()
    gx = 2.0 * (xT.astype(np.float64) - target) / norm
    L = float(np.sum((xT.astype(np.float64) - target) ** 2) / norm)
    gx0, gv0, gth = solver.backward(gx.astype(np.float32))
    return L, gx0, gv0, gth


@pytest.mark.parametrize("substeps,every", [(8, 4), (16, 8)])
def test_adjoint_matches_fp64_derivative(substeps, every):
    scs, prms = small()
    sc, prm = scs[0], prms[0]
    n = sc.num_particles()
    target = seeded_target(sc, 1)
    case(f"adjoint n={n} substeps={substeps} every={every}")
    s = bp.Solver(sc, prm)
    L, gx0, gv0, gth = gpu_gradient(s, target, n, substeps, every)
    rng = np.random.default_rng(7)
    dirs = [rng.normal(0, 1, (n, 3)) for _ in range(2)]
    Lh, dL = orc.rollout_tangent(bp.SceneArrays(sc), prm, substeps, DT, target, n, dirs)
    print(f"loss gpu {L:.9e} fp64 {Lh:.9e}")
    names = ["wind_x", "wind_y", "wind_z", "drag", "lift", "xsph_c"]
    worst = 0.0
    for k in range(6):
        e = rel(float(gth[0, k]), dL[k])
        worst = max(worst, e)
        print(f"  d/d{names[k]}: gpu {gth[0, k]:.6e} fp64 {dL[k]:.6e} rel {e:.2e}")
    for j, d in enumerate(dirs):
        g = float(np.sum(gv0.astype(np.float64) * d))
        e = rel(g, dL[6 + j])
        worst = max(worst, e)
        print(f"  dL/dv0 . u{j}: gpu {g:.6e} fp64 {dL[6 + j]:.6e} rel {e:.2e}")
    assert rel(L, Lh) < 1e-4
    assert worst <= ADJ_TOL, f"worst relative gradient error {worst:.3e}"


def test_adjoint_batched_per_scene():
    scs, prms = small(num=3, seed=9, fluid=(6, 6, 6), cloth=10)
    ntot = sum(sc.num_particles() for sc in scs)
    targets = [seeded_target(sc, 2 + k) for k, sc in enumerate(scs)]
    case(f"batched adjoint scenes=3 n={ntot}")
    bs = bp.BatchSolver(scs, prms)
    target = np.concatenate(targets)
    L, gx0, gv0, gth = gpu_gradient(bs, target, ntot, 8, 4)
    for k, (sc, prm) in enumerate(zip(scs, prms)):
        Lh, dL = orc.rollout_tangent(bp.SceneArrays(sc), prm, 8, DT, targets[k], ntot, [])
        for c in range(6):
            e = rel(float(gth[k, c]), dL[c])
            print(f"  scene {k} theta {c}: gpu {gth[k, c]:.6e} fp64 {dL[c]:.6e} rel {e:.2e}")
            assert e <= ADJ_TOL


def test_adjoint_rollout_is_the_plain_rollout_and_repeatable():
    scs, prms = small()
    sc, prm = scs[0], prms[0]
    n = sc.num_particles()
    target = seeded_target(sc, 3)
    case("rollout equals plain steps; backward repeatable; state restored")
    a, b = bp.Solver(sc, prm), bp.Solver(sc, prm)
    a.forward_checkpointed(8, DT, 4)
    for _ in range(8):
        b.step(1, DT)
    pa, va = a.state_original_order()
    pb, vb = b.state_original_order()
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    gx = (2.0 * (pa - target) / n).astype(np.float32)
    r1 = a.backward(gx)
    r2 = a.backward(gx)
    for u, v in zip(r1, r2):
        assert np.array_equal(u, v)
    pa2, va2 = a.state_original_order()
    assert np.array_equal(pa, pa2) and np.array_equal(va, va2)
    # the adjoint is linear in the seed
    r3 = a.backward(2.0 * gx)
    np.testing.assert_allclose(r3[2], 2.0 * r1[2], rtol=1e-5, atol=1e-12)
    # any state change discards the rollout
    a.step(1, DT)
    with pytest.raises(bp.BrassPanError):
        a.backward(gx)


def test_adjoint_refuses_unsupported_scenes():
    sc = scenes.Scene(particle_radius=0.025, fluid_kernel_radius=0.1, cell_size=0.1, grid_dims=(32, 32, 32))
    sc.add_rigid_box((4, 4, 4), (1.0, 1.0, 1.0), 0.05, 1.0)
    sc.add_box_planes((0, 0, 0), (3.2, 3.2, 3.2))
    s = bp.Solver(sc, bp.default_params(akinci=0, vorticity_eps=0.0))
    case("adjoint refuses rigid bodies")
    with pytest.raises(bp.BrassPanError) as e:
        s.forward_checkpointed(4, DT, 2)
    assert e.value.status == 6  # BP_ERR_UNSUPPORTED


def test_torch_harness_gradients():
    import torch

    from autograd import Rollout

    scs, prms = small(num=2, seed=21, fluid=(6, 6, 6), cloth=10)
    ntot = sum(sc.num_particles() for sc in scs)
    targets = [seeded_target(sc, 30 + k) for k, sc in enumerate(scs)]
    th0 = np.array([[p.wind[0], p.wind[1], p.wind[2], p.drag_coeff, p.lift_coeff, p.xsph_c] for p in prms],
                   np.float32)
    theta = torch.tensor(th0, requires_grad=True)
    v0 = torch.zeros(ntot, 3, requires_grad=True)
    case(f"torch harness scenes=2 n={ntot}")
    xT = Rollout.apply(theta, v0, scs, prms, 8, DT, 4)
    loss = ((xT.double() - torch.from_numpy(np.concatenate(targets))) ** 2).sum() / ntot
    loss.backward()
    rng = np.random.default_rng(11)
    off = 0
    for k, (sc, prm) in enumerate(zip(scs, prms)):
        n = sc.num_particles()
        u = rng.normal(0, 1, (n, 3))
        _, dL = orc.rollout_tangent(bp.SceneArrays(sc), prm, 8, DT, targets[k], ntot, [u])
        for c in range(6):
            assert rel(float(theta.grad[k, c]), dL[c]) <= ADJ_TOL
        g = float(np.sum(v0.grad[off:off + n].numpy().astype(np.float64) * u))
        assert rel(g, dL[6]) <= ADJ_TOL
        off += n


def test_full_size_adjoint_against_gpu_differences():
    """The differentiable workload at full size (64 scenes of 16,384 fluid
    particles + a 64 x 64 cloth, 32 substeps, checkpoints every 8): scene 0's
    wind and drag gradients against central differences of the GPU's own FP32
    rollout.  The cloth does not touch the fluid, so these parameters only
    move the cloth, whose rollout is smooth; FP32 rounding of the differences
    bounds the agreement (FD_TOL)."""
    FD_TOL = 2e-2
    scs, prms = scenes.adjoint_scenes(64)
    n = sum(sc.num_particles() for sc in scs)
    K, every = scenes.ADJOINT_SUBSTEPS, scenes.ADJOINT_EVERY
    rng = np.random.default_rng(4)
    target = np.concatenate([np.asarray(sc.positions, np.float64).reshape(-1, 3) for sc in scs])
    target += rng.normal(0.0, 0.05, target.shape)
    case(f"full-size adjoint n={n}")
    s = bp.BatchSolver(scs, prms)
    L, gx0, gv0, gth = gpu_gradient(s, target, n, K, every)
    s.close()
    assert np.all(np.isfinite(gv0)) and np.all(np.isfinite(gth))

    def loss_with(k, eps):
        p2 = list(prms)
        q = bp.bp_params.from_buffer_copy(prms[0])
        if k < 3:
            q.wind[k] += eps
        else:
            q.drag_coeff += eps
        p2[0] = q
        t = bp.BatchSolver(scs, p2)
        for _ in range(K):
            t.step(1, DT)
        x = t.state_original_order()[0].astype(np.float64)
        t.close()
        return float(np.sum((x - target) ** 2) / n)

    for k, eps in [(0, 0.05), (2, 0.05), (3, 0.5)]:
        fd = (loss_with(k, eps) - loss_with(k, -eps)) / (2 * eps)
        e = rel(float(gth[0, k]), fd)
        print(f"  scene 0 theta {k}: adjoint {gth[0, k]:.6e} fd {fd:.6e} rel {e:.2e}")
        assert e <= FD_TOL


def _scene_at(sc, pos, vel):
    """A copy of scene `sc` that starts from the state (pos, vel), original order."""
    import copy
    out = copy.deepcopy(sc)
    out._pos = [np.ascontiguousarray(pos, np.float32)]
    out._vel = [np.ascontiguousarray(vel, np.float32)]
    out._has_vel = True
    return out


def _compare(label, L, gv0, gth, Lh, dL, dirs, which=range(6)):
    names = ["wind_x", "wind_y", "wind_z", "drag", "lift", "xsph_c"]
    worst = 0.0
    print(f"{label}: loss gpu {L:.9e} fp64 {Lh:.9e}")
    for k in range(6):
        e = rel(float(gth[0, k]), dL[k])
        if k in which:
            worst = max(worst, e)
        print(f"  d/d{names[k]}: gpu {gth[0, k]:.6e} fp64 {dL[k]:.6e} rel {e:.2e}{'' if k in which else ' (reported)'}")
    for j, d in enumerate(dirs):
        g = float(np.sum(gv0.astype(np.float64) * d))
        e = rel(g, dL[6 + j])
        worst = max(worst, e)
        print(f"  dL/dv0 . u{j}: gpu {g:.6e} fp64 {dL[6 + j]:.6e} rel {e:.2e}")
    return worst


def test_full_size_scene_adjoint_matches_fp64_derivative():
    """One scene of the differentiable workload at its full size (16,384
    fluid particles + a 64 x 64 cloth in wind) against the FP64 host
    derivative.

    Over the whole rollout (32 substeps, checkpoints every 8) the FP32 and
    FP64 trajectories of the fluid drift apart (the half-sum density re-sorts
    when a particle changes cell, so the fluid's rollout is only piecewise
    smooth and the two take different branches), so the derivatives along the
    fluid's initial velocities differ by several percent while the loss agrees
    to 1e-4.  The cloth's parameters (wind, drag, lift) move only the cloth,
    which never meets the fluid: their gradients over the whole rollout are
    checked at ADJ_TOL.  Every component -- the six parameters and the initial
    velocities along two random directions -- is then checked at ADJ_TOL on
    each of the four 8-substep segments, GPU and host starting from the same
    FP32 state (the GPU rollout's state at the segment's first substep).  By
    the chain rule the segments' derivatives compose the rollout's.

    Also reports the largest term the backward delivered between slots in
    fixed point (bp_debug_fetch 6), which must stay inside the format's range
    (|term| <= 2^17 with the seeds scaled to a unit maximum)."""
    scs, prms = scenes.adjoint_scenes(1)
    sc, prm = scs[0], prms[0]
    n = sc.num_particles()
    assert n == 16_384 + 64 * 64
    K, every = scenes.ADJOINT_SUBSTEPS, scenes.ADJOINT_EVERY
    target = seeded_target(sc, 4)
    rng = np.random.default_rng(8)
    dirs = [rng.normal(0, 1, (n, 3)) for _ in range(2)]
    case(f"full-size scene adjoint n={n} substeps={K} every={every}")
    s = bp.Solver(sc, prm)
    L, gx0, gv0, gth = gpu_gradient(s, target, n, K, every)
    mx = float(s.debug(6)[0])
    s.close()
    print(f"largest delivered term {mx:.3e} (seed units)")
    assert mx <= 2.0 ** 17
    case("FP64 host derivative of the full-size rollout")
    Lh, dL = orc.rollout_tangent(bp.SceneArrays(sc), prm, K, DT, target, n, dirs)
    _compare("32 substeps", L, gv0, gth, Lh, dL, [], which=range(5))
    assert rel(L, Lh) < 1e-4
    worst = max(rel(float(gth[0, k]), dL[k]) for k in range(5))
    assert worst <= ADJ_TOL, f"cloth parameters over the rollout: worst relative error {worst:.3e}"

    # the segment starts: states entering substeps 0, 8, 16, 24 of the GPU rollout
    g = bp.Solver(sc, prm)
    starts = []
    for k in range(K):
        if k % every == 0:
            starts.append(g.state_original_order())
        g.step(1, DT)
    g.close()
    worst = 0.0
    for seg, (pos, vel) in enumerate(starts):
        sseg = _scene_at(sc, pos, vel)
        case(f"segment {seg}: substeps {seg * every}..{seg * every + every - 1}")
        t = bp.Solver(sseg, prm)
        Ls, _, gv0s, gths = gpu_gradient(t, target, n, every, every)
        t.close()
        Lhs, dLs = orc.rollout_tangent(bp.SceneArrays(sseg), prm, every, DT, target, n, dirs)
        worst = max(worst, _compare(f"segment {seg}", Ls, gv0s, gths, Lhs, dLs, dirs))
        assert rel(Ls, Lhs) < 1e-5
    assert worst <= ADJ_TOL, f"worst relative gradient error over the segments {worst:.3e}"
