"""Scene construction (brasstower/scene.h, scene.cpp) and the benchmark
workloads: topology counts, rest density, PRNG stream."""
import numpy as np
import pytest

import scenes


def test_splitmix64_known_values():
    # published splitmix64 reference stream for seed 0
    out = scenes.splitmix64(0, 3)
    assert [int(v) for v in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    u = scenes.uniform01(1, 1000)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.05


@pytest.mark.parametrize("n", [4, 32, 128])
def test_cloth_topology_counts(n):
    sc = scenes.Scene(particle_radius=0.025, fluid_kernel_radius=0.1, cell_size=0.1, grid_dims=(64, 64, 64))
    sc.add_fluid_block((2, 2, 2), (0, 0, 0), 0.05, 1.0)  # offset check
    s, e = sc.add_cloth((0, 0, 0), (0.05, 0, 0), (0, 0.05, 0), n, n, 1.0, 1.0, 0.5)
    assert (s, e) == (8, 8 + n * n)
    nd = len(sc.distance_pairs)
    # structural 2n(n-1) + shear 2(n-1)^2 + hack bending 6(n-2)^2 (scene.cpp:47-106)
    assert nd == 2 * n * (n - 1) + 2 * (n - 1) ** 2 + 6 * (n - 2) ** 2
    assert len(sc.faces) == 2 * (n - 1) ** 2
    pairs = np.asarray(sc.distance_pairs)
    assert pairs.min() >= s and pairs.max() < e
    # rest lengths equal the initial distances for structural / shear pairs
    pos = sc.positions
    prm = np.asarray(sc.distance_params)
    d = np.linalg.norm(pos[pairs[:, 0]] - pos[pairs[:, 1]], axis=1)
    assert np.allclose(d, prm[:, 0], atol=1e-6)
    assert len(set(sc.phases[s:e].tolist())) == n * n  # self-collidable: one phase per particle
    assert np.all(sc.masses[s:e] == 1.0)


def test_rigid_box_rest_positions_centered():
    sc = scenes.Scene(grid_dims=(64, 64, 64))
    sc.add_fluid_block((2, 2, 2), (0, 0, 0), 0.05, 1.0)
    s, e = sc.add_rigid_box((4, 4, 4), (1, 1, 1), 0.05, 1.0, rotation=(0, 0, 0.38268343, 0.92387953))
    rest = sc.rigidbody_initial_positions
    assert len(rest) == e and np.all(rest[:s] == 0)
    assert np.abs(rest[s:e].astype(np.float64).mean(0)).max() < 1e-7
    pos = sc.positions[s:e].astype(np.float64)
    # same shape: pairwise distances preserved by the rotation
    a = np.linalg.norm(pos[:, None] - pos[None], axis=2)
    b = np.linalg.norm(rest[s:e, None].astype(np.float64) - rest[None, s:e], axis=2)
    assert np.allclose(a, b, atol=1e-6)
    assert len(set(sc.phases[s:e].tolist())) == 1 and sc.phases[s] > 0


def test_workload_sizes():
    assert scenes.dam_break().num_particles() == 16384
    mp = scenes.mixed_phase()
    assert mp.num_particles() == 262144 + 64 * 64 + 128 * 128
    assert len(mp.distance_pairs) == 160026 and len(mp.faces) == 32258
    assert len(mp.immovable_constraints) == 128
    assert len(mp.rigidbody_particle_id_ranges) == 64
    assert mp.rigidbody_particle_id_ranges[0] == (0, 64)  # rigid bodies first


def test_lattice_rest_density_matches_mean_half_sum():
    # direct lattice sum over a large unjittered block interior
    h, sp = 0.1, 0.05
    k = 2
    offs = np.array([(a, b, c) for a in range(-k, k + 1) for b in range(-k, k + 1) for c in range(-k, k + 1)], float) * sp
    r2 = (offs ** 2).sum(1)
    w = scenes.poly6_f64(r2[r2 <= h * h], h)
    full = w.sum()
    w0 = scenes.poly6_f64(0.0, h)
    assert scenes.lattice_rest_density(sp, h, 1.0) == pytest.approx(w0 + 0.5 * (full - w0), rel=1e-6)
