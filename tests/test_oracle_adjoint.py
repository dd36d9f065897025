"""CPU: pins of the FP64 host derivative of a rollout (or_rollout_tangent,
forward-mode dual numbers over the oracle's reference-order FP64 rollout),
which the GPU adjoint is checked against.

- closed form: a free-falling particle, x_T = x_0 + K dt v_0 + dt^2 K(K+1)/2 g;
- central finite differences of the FP64 rollout with eps = 1e-4 (relative
  for XSPH c, whose value is 2e-5) for every parameter and for initial
  velocities along random directions over 8 substeps;
- over 32 substeps the fluid's rollout is only piecewise smooth (the
  reference's density sums over partners with a higher slot index, so a
  particle changing cell re-sorts the slots and the density jumps); there the
  differences are taken with eps = 1e-8, inside the smooth piece."""
import numpy as np

import brasspan as bp
import oracle_py as orc
import scenes

DT = scenes.ADJOINT_SUBDT


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-30)


def small():
    scs, prms = scenes.adjoint_scenes(1, seed=5, fluid_dims=(8, 8, 8), cloth_n=12, box=3.2, grid=(32, 32, 32))
    return scs[0], prms[0]


def theta_of(prm):
    return np.array([prm.wind[0], prm.wind[1], prm.wind[2], prm.drag_coeff, prm.lift_coeff, prm.xsph_c],
                    np.float32).astype(np.float64)


def test_tangent_closed_form_free_fall(have_oracle):
    sc = scenes.Scene(particle_radius=0.025, fluid_kernel_radius=0.1, cell_size=0.1, grid_dims=(16, 16, 16))
    sc.add_granulars_block((1, 1, 1), (0.8, 1.2, 0.8), 0.05, 1.0)
    prm = bp.default_params(wind_enabled=0)
    arr = bp.SceneArrays(sc)
    K = 5
    x0 = np.asarray(sc.positions, np.float64).reshape(-1, 3)
    target = x0 + np.array([[0.1, -0.2, 0.05]])
    u = np.array([[0.3, -1.0, 2.0]])
    L, dL = orc.rollout_tangent(arr, prm, K, DT, target, 1.0, [u])
    dt = np.float32(DT)
    g = np.array([0.0, float(np.float32(-9.8)), 0.0])
    xT = x0 + g * float(dt) ** 2 * K * (K + 1) / 2
    assert abs(L - float(np.sum((xT - target) ** 2))) < 1e-9
    expect = 2.0 * float(np.sum((xT - target) * u)) * K * float(dt)
    assert rel(dL[6], expect) < 1e-6
    assert np.all(dL[:6] == 0.0)


def test_tangent_matches_finite_differences_8_substeps(have_oracle):
    sc, prm = small()
    arr = bp.SceneArrays(sc)
    n = arr.n
    rng = np.random.default_rng(0)
    target = np.asarray(sc.positions, np.float64).reshape(-1, 3) + rng.normal(0, 0.05, (n, 3))
    dirs = [rng.normal(0, 1, (n, 3)) for _ in range(2)]
    K = 8
    L, dL = orc.rollout_tangent(arr, prm, K, DT, target, n, dirs)
    th = theta_of(prm)
    for k in range(6):
        eps = 1e-4 * abs(th[k]) if k == 5 else 1e-4
        tp, tm = th.copy(), th.copy()
        tp[k] += eps
        tm[k] -= eps
        fd = (orc.rollout_loss(arr, prm, K, DT, target, n, tp) - orc.rollout_loss(arr, prm, K, DT, target, n, tm)) / (2 * eps)
        print(k, fd, dL[k], rel(dL[k], fd))
        assert rel(dL[k], fd) < 1e-5
    v0 = np.zeros((n, 3))
    for j, d in enumerate(dirs):
        fd = (orc.rollout_loss(arr, prm, K, DT, target, n, None, v0 + 1e-4 * d)
              - orc.rollout_loss(arr, prm, K, DT, target, n, None, v0 - 1e-4 * d)) / 2e-4
        print("v0", j, fd, dL[6 + j], rel(dL[6 + j], fd))
        assert rel(dL[6 + j], fd) < 1e-3


def test_tangent_matches_finite_differences_32_substeps(have_oracle):
    sc, prm = small()
    arr = bp.SceneArrays(sc)
    n = arr.n
    rng = np.random.default_rng(1)
    target = np.asarray(sc.positions, np.float64).reshape(-1, 3) + rng.normal(0, 0.05, (n, 3))
    dirs = [rng.normal(0, 1, (n, 3))]
    K = 32
    L, dL = orc.rollout_tangent(arr, prm, K, DT, target, n, dirs)
    th = theta_of(prm)
    for k in range(5):  # wind, drag, lift act on the cloth, whose rollout is smooth
        tp, tm = th.copy(), th.copy()
        tp[k] += 1e-4
        tm[k] -= 1e-4
        fd = (orc.rollout_loss(arr, prm, K, DT, target, n, tp) - orc.rollout_loss(arr, prm, K, DT, target, n, tm)) / 2e-4
        assert rel(dL[k], fd) < 1e-4
    v0 = np.zeros((n, 3))
    eps = 1e-8
    fd = (orc.rollout_loss(arr, prm, K, DT, target, n, None, v0 + eps * dirs[0])
          - orc.rollout_loss(arr, prm, K, DT, target, n, None, v0 - eps * dirs[0])) / (2 * eps)
    assert rel(dL[6], fd) < 1e-4
    tp, tm = th.copy(), th.copy()
    tp[5] += 1e-12
    tm[5] -= 1e-12
    fd = (orc.rollout_loss(arr, prm, K, DT, target, n, tp) - orc.rollout_loss(arr, prm, K, DT, target, n, tm)) / 2e-12
    assert rel(dL[5], fd) < 1e-3
