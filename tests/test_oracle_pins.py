"""Pins of the CPU oracle (oracle/bp_oracle.cpp) against closed forms,
invariants and brute force.  The reference ships no tests or golden vectors
(its only check is the opt-in CHECK_NAN macro, kernel/cubrasstower.cuh:15),
so every oracle function is pinned here by mathematics that a plausible
mistake (dropped term, wrong sign or index, transposed operand) would break.
CPU only."""
import math

import numpy as np
import pytest

import brasspan as bp
import oracle_py as orc
import scenes

PI_REF = float(np.float32(3.141592))  # the reference's pi literal as an FP32 constant (fluid.cuh:12-16)
F32 = lambda v: float(np.float32(v))  # scene and parameter values enter as FP32


def quiet_params(**kw):
    """No gravity, planes off by scene, no post passes unless asked."""
    p = dict(gravity=(0.0, 0.0, 0.0), iterations=1, stabilization_passes=0, vorticity_eps=0.0, xsph_c=0.0,
             akinci=0, fluid_rest_density=1000.0, mass_scaling_k=0.0)
    p.update(kw)
    return bp.default_params(**p)


def make_scene(grid=(16, 16, 16), h=0.1):
    return scenes.Scene(particle_radius=0.025, fluid_kernel_radius=h, cell_size=h, grid_dims=grid)


def run_one(sc, prm, fp64=True, mode=orc.SCATTER, substeps=1, dt=1.0 / 120.0):
    o = orc.Oracle(bp.SceneArrays(sc), prm, fp64=fp64, mode=mode)
    x0, v0 = o.state_original_order()
    o.step(substeps, dt)
    return o, x0, v0


# ---------------------------------------------------------------- SPH kernels
def test_poly6_closed_form(have_oracle):
    for h in (0.1, 0.115, 0.23):
        w0 = 315.0 / (64.0 * PI_REF * h ** 3)
        assert orc.poly6(0.0, h) == pytest.approx(w0, rel=1e-12)
        assert orc.poly6(h * h, h) == 0.0
        assert orc.poly6(0.0, h, fp64=False) == pytest.approx(w0, rel=2e-6)
        r2 = 0.3 * h * h
        assert orc.poly6(r2, h) == pytest.approx(315.0 / (64 * PI_REF * h ** 9) * (h * h - r2) ** 3, rel=1e-12)


def test_poly6_integrates_to_one(have_oracle):
    h = 0.1
    r = np.linspace(0.0, h, 4001)
    w = np.array([orc.poly6(x * x, h) for x in r])
    integral = 4.0 * math.pi * np.trapezoid(w * r * r, r)
    assert integral == pytest.approx(math.pi / PI_REF, rel=1e-6)


def test_spiky_gradient_is_derivative_of_spiky(have_oracle):
    h = 0.1
    spiky = lambda r: 15.0 / (PI_REF * h ** 6) * (h - r) ** 3
    for r in (0.01, 0.05, 0.08, 0.0999):
        v = np.array([r, 0.0, 0.0])
        g = orc.grad_spiky(v, r * r, h)
        e = 1e-7
        fd = (spiky(r + e) - spiky(r - e)) / (2 * e)
        assert g[0] == pytest.approx(fd, rel=1e-6)
        assert g[1] == 0.0 and g[2] == 0.0
    # direction along v, zero at the support radius
    d = np.array([0.03, -0.04, 0.0])
    g = orc.grad_spiky(d, float(d @ d), h)
    assert np.allclose(np.cross(g, d), 0.0, atol=1e-9) and g @ d < 0
    assert np.allclose(orc.grad_spiky(np.array([h, 0, 0]), h * h, h), 0.0)


def test_akinci_spline_reference_form(have_oracle):
    h = 0.1
    c3 = 32.0 / (PI_REF * h ** 9)
    for r in (0.01, 0.03, 0.049, 0.05, 0.07, 0.0999):
        t3 = ((h - r) * r) ** 3
        # akinciSplineC (fluid.cuh:59-73): the h^6/64 term is not scaled by 32/(pi h^9)
        expect = c3 * t3 if r >= h / 2 else 2.0 * c3 * t3 - h ** 6 / 64.0
        assert orc.akinci_c(r, h) == pytest.approx(expect, rel=1e-12)


# ---------------------------------------------------------------- neighbour grid
def numpy_keys(xs, cs, dims):
    xs = np.asarray(xs, np.float32)
    g = ((xs + np.float32(50000.0)) / np.asarray(cs, np.float32)).astype(np.int64)
    m = np.asarray(dims) - 1
    g = g & m
    return (g[:, 0] * dims[1] * dims[2] + g[:, 1] * dims[2] + g[:, 2]).astype(np.int32)


@pytest.mark.parametrize("dims", [(16, 16, 16), (64, 8, 4), (4, 4, 4)])
def test_neighbor_grid_brute_force(have_oracle, dims):
    rng = np.random.default_rng(7)
    n = 5000
    cs = (0.1, 0.07, 0.13)
    xs = rng.uniform(-1.0, 3.0, size=(n, 3)).astype(np.float32)
    xs[:50] = xs[50:100]  # duplicates test stability
    keys, perm, st, en = orc.neighbor_grid(xs, cs, dims)
    k_np = numpy_keys(xs, cs, dims)
    p_np = np.argsort(k_np, kind="stable")
    assert np.array_equal(perm, p_np)
    assert np.array_equal(keys, k_np[p_np])
    assert np.all(np.diff(keys) >= 0)
    # ranges partition [0, n): every occupied cell maps to its run, others -1
    occ = np.unique(keys)
    assert np.all(st[occ] >= 0)
    assert np.sum(en[occ] - st[occ]) == n
    for c in occ[:200]:
        assert np.all(keys[st[c]:en[c]] == c)
    empty = np.setdiff1d(np.arange(np.prod(dims)), occ)
    assert np.all(st[empty] == -1) and np.all(en[empty] == -1)


def test_grid_hash_quantisation_quirk(have_oracle):
    # (p + 50000) is rounded to 2^-8 in FP32 before the division (grid.cuh:12):
    # a point 0.001 below a cell boundary hashes into the next cell.
    cs, dims = (0.1, 0.1, 0.1), (64, 64, 64)
    p = np.array([0.3 - 0.001, 0.0, 0.0], np.float32)
    exact = int((50000.0 + 0.299) / 0.1)
    a = orc.cell_address(p, cs, dims)
    assert (a // (64 * 64)) == (numpy_keys(p[None], cs, dims)[0] // (64 * 64))
    assert (a // (64 * 64)) == ((exact + 1) & 63)


# ---------------------------------------------------------------- fluid
def lattice_scene(n=6, jitter=0.004, seed=11, h=0.1, grid=(16, 16, 16), solid_layer=False):
    sc = make_scene(grid, h)
    sc.add_fluid_block((n, n, n), (0.5, 0.5, 0.5), 0.05, 1.0, jitter=jitter, seed=seed)
    if solid_layer:  # solids 0.06 apart and 0.065 below the fluid: inside h, outside contact 2r
        sc.add_granulars_block((n, 1, n), (0.5, 0.435, 0.5), (0.06, 0.06, 0.06), 1.0)
    return sc


def brute_density(xs_sorted, phases_sorted, masses_sorted, cs, dims, h, sds):
    """fluidLambda density (fluid.cuh:78-159) by brute force over all pairs."""
    keys = numpy_keys(xs_sorted, cs, dims)
    d = np.asarray(dims)
    coords = np.stack([keys // (d[1] * d[2]), (keys // d[2]) % d[1], keys % d[2]], 1)
    g = ((xs_sorted.astype(np.float32) + np.float32(50000)) / np.float32(cs[0])).astype(np.int64) & (d - 1)
    n = len(xs_sorted)
    out = np.zeros(n)
    for i in range(n):
        diffc = (coords - g[i] + 1) % d
        cand = np.all(diffc <= 2, axis=1)
        cand[:i] = False  # j >= i only
        r2 = np.sum((xs_sorted[i] - xs_sorted) ** 2, axis=1)
        sel = cand & (r2 <= h * h)
        w = 315.0 / (64 * PI_REF * h ** 9) * (h * h - r2[sel]) ** 3 * masses_sorted[sel]
        w = np.where(phases_sorted[sel] < 0, w, sds * w)
        out[i] = w.sum()
    return out


def test_density_half_sum_brute_force(have_oracle):
    sc = lattice_scene(solid_layer=True)
    prm = quiet_params()
    o, x0, _ = run_one(sc, prm, dt=1e-30)  # x* = x: the lambda pass sees the initial positions
    pos, _, mp = o.get_state()
    dens = o.debug(0)
    ph = sc.phases[mp]
    ms = sc.masses[mp].astype(np.float64)
    xs_sorted = x0[mp]
    ref = brute_density(xs_sorted, ph, ms, sc.cell_size, sc.grid_dims, F32(0.1), 2.0)
    fl = ph < 0
    assert np.allclose(dens[fl], ref[fl], rtol=1e-12, atol=0)
    lam = o.debug(1)
    assert np.allclose(lam[fl], -(ref[fl] / 1000.0 - 1.0) / 15000.0, rtol=1e-12, atol=1e-18)


def test_lambda_akinci_clamp_and_mode1(have_oracle):
    sc = lattice_scene(jitter=0.0)
    o, _, _ = run_one(sc, quiet_params(akinci=1, fluid_rest_density=1e9), dt=1e-30)
    assert np.all(o.debug(1) == 0.0)  # C clamped at 0 (fluid.cuh:150)
    o2, _, _ = run_one(sc, quiet_params(lambda_mode=1), dt=1e-30)
    o1, _, _ = run_one(sc, quiet_params(), dt=1e-30)
    c = o1.debug(0) / 1000.0 - 1.0
    lam2 = o2.debug(1)
    # mode 1 (fluid.cuh:155): -C/(sum|grad C|^2 + eps) has the sign of -C and |.| < |C|/eps
    assert np.all(np.sign(lam2[np.abs(c) > 1e-9]) == -np.sign(c[np.abs(c) > 1e-9]))
    assert np.all(np.abs(lam2) <= np.abs(c) / 300.0 + 1e-18)


def test_position_pass_conserves_momentum(have_oracle):
    sc = lattice_scene(jitter=0.006)
    for mode in (orc.SCATTER, orc.GATHER):
        o, x0, _ = run_one(sc, quiet_params(), mode=mode, dt=1e-30)
        x1, _ = o.state_original_order()
        dx = x1 - x0
        assert np.abs(dx).max() > 1e-6  # non-trivial correction
        p = (sc.masses[:, None] * dx).sum(0)
        assert np.abs(p).max() < 1e-12 * max(1.0, np.abs(dx).sum())


def test_scatter_and_gather_agree_without_cell_changes(have_oracle):
    sc = lattice_scene(jitter=0.006, solid_layer=True)
    prm = quiet_params(iterations=1)
    a, _, _ = run_one(sc, prm, mode=orc.SCATTER, dt=1e-30)
    b, _, _ = run_one(sc, prm, mode=orc.GATHER, dt=1e-30)
    xa, _ = a.state_original_order()
    xb, _ = b.state_original_order()
    assert np.abs(xa - xb).max() < 1e-14


def test_two_particle_position_correction(have_oracle):
    h, m, rho0, r = F32(0.1), 1.0, 1000.0, F32(0.04)
    sc = make_scene()
    sc.add_fluid_block((2, 1, 1), (0.5, 0.5, 0.5), (r, 0, 0), m)
    prm = quiet_params(fluid_rest_density=rho0)
    o, x0, _ = run_one(sc, prm, dt=1e-30)
    x1, _ = o.state_original_order()
    w = lambda r2: 315.0 / (64 * PI_REF * h ** 9) * (h * h - r2) ** 3
    dens = np.array([m * (w(0) + w(r * r)), m * w(0)])  # slot 0 sees both (j >= i)
    lam = -(dens / rho0 - 1.0) / 15000.0
    wdq = w((0.03 * h) ** 2)
    scorr = -1e-4 * (w(r * r) / wdq) ** 4
    diff = np.array([-r, 0.0, 0.0])  # x0 - x1
    grad = -45.0 / (PI_REF * h ** 6) * (h - r) ** 2 * diff / r
    delta = (lam[0] + lam[1] + scorr) * grad
    ln = np.linalg.norm(delta)
    if ln >= 5.0:  # clamp (fluid.cuh:214-216)
        delta = delta * 5.0 / ln
    assert np.allclose(x1[0] - x0[0], m * delta / rho0, rtol=1e-10, atol=1e-16)
    assert np.allclose(x1[1] - x0[1], -m * delta / rho0, rtol=1e-10, atol=1e-16)


def test_xsph_uniform_velocity_unchanged(have_oracle):
    sc = lattice_scene(jitter=0.004)
    v = np.tile(np.array([0.3, -0.2, 0.1], np.float32), (sc.num_particles(), 1))
    sc._vel = [v]
    sc._has_vel = True
    prm = quiet_params(xsph_c=0.01, lambda_denominator=1e30, scorr_k=0.0)
    o, _, _ = run_one(sc, prm, fp64=True)
    _, v1 = o.state_original_order()
    assert np.abs(v1 - v).max() < 1e-8  # (x+v dt - x)/dt with an FP32 1/dt only


def test_akinci_normal_direction_and_tension_momentum(have_oracle):
    """Akinci surface tension (fluid.cuh:387-489).  Normal: n_i = h sum_j
    (1/rho_j) gradW(x_i - x_j), and gradW(v) is a negative multiple of v, so a
    particle of the top face of a block (all neighbours level or below) has a
    normal pointing down and the bottom face one pointing up.  Tension: k_ij is
    symmetric and both the cohesion term C(r)(x_i - x_j)/r and the curvature
    term n_i - n_j change sign under i <-> j, so on a block at rest (no
    pressure correction, no gravity: the pair sets stay symmetric) the forces
    cancel in pairs and the total momentum does not change."""
    sc = lattice_scene(n=6, jitter=0.002)
    dt = 1.0 / 120.0
    prm = quiet_params(akinci=1, tension_gamma=0.5, lambda_denominator=1e30, scorr_k=0.0)
    o, x0, v0 = run_one(sc, prm, dt=dt)
    _, v1 = o.state_original_order()
    dv = v1 - v0
    assert np.abs(dv).max() > 1e-4  # the tension acts
    dp = (sc.masses[:, None] * dv).sum(0)
    assert np.abs(dp).max() < 1e-12 * (sc.masses[:, None] * np.abs(dv)).sum()
    nrm = o.debug(5)
    _, _, mp = o.get_state()
    y = sc.positions[mp][:, 1]
    top, bottom = y > y.max() - 0.01, y < y.min() + 0.01
    assert np.all(nrm[top][:, 1] < 0.0) and np.all(nrm[bottom][:, 1] > 0.0)


def test_vorticity_rigid_rotation_closed_form(have_oracle):
    h = 0.1
    sc = lattice_scene(n=7, jitter=0.0)
    x = sc.positions.astype(np.float64)
    c = x.mean(0)
    omega = np.array([0.0, 0.0, 2.0])
    v = np.cross(omega, x - c)
    sc._vel = [v.astype(np.float32)]
    sc._has_vel = True
    prm = quiet_params(vorticity_eps=1e-4, lambda_denominator=1e30, scorr_k=0.0)
    o, _, _ = run_one(sc, prm, dt=1e-6)
    om = o.debug(4)
    _, _, mp = o.get_state()
    centre = np.argmin(np.linalg.norm(x[mp] - c, axis=1))
    # interior lattice particle: omega_i = (2/3) Omega sum_j c(r_j) r_j^2, c(r) = C2 (h-r)^2 / r
    offs = np.array([(a, b, d) for a in range(-2, 3) for b in range(-2, 3) for d in range(-2, 3)
                     if 0 < a * a + b * b + d * d <= 4], float) * 0.05
    rr = np.linalg.norm(offs, axis=1)
    c2 = -45.0 / (PI_REF * h ** 6)
    expect = (2.0 / 3.0) * omega * np.sum(c2 * (h - rr) ** 2 / rr * rr * rr)
    assert np.allclose(om[centre], expect, rtol=1e-4, atol=1e-6 * np.abs(expect).max())


# ---------------------------------------------------------------- solids, cloth, boundary
def test_plane_stabilization_pushes_to_radius(have_oracle):
    sc = make_scene()
    sc.add_granulars_block((1, 1, 1), (0.5, -0.1, 0.5), (0.05, 0.05, 0.05), 1.0)
    sc.add_plane((0, 0, 0), (0, 1, 0))
    prm = quiet_params(iterations=0, stabilization_passes=2)
    o, _, _ = run_one(sc, prm)
    x1, _ = o.state_original_order()
    assert x1[0, 1] == pytest.approx(F32(0.025), abs=1e-15)


def test_plane_collision_static_friction(have_oracle):
    sc = make_scene()
    sc.add_granulars_block((1, 1, 1), (0.5, 0.03, 0.5), (0.05, 0.05, 0.05), 1.0)
    sc._vel = [np.array([[0.3, -6.0, 0.0]], np.float32)]
    sc._has_vel = True
    sc.add_plane((0, 0, 0), (0, 1, 0))
    dt = F32(1.0 / 60.0)
    # moving into the plane, diffNormal < 0, so the static test (particleplane.cuh:51)
    # never passes; with dynamic friction 0 the tangential move is kept in full
    o, x0, _ = run_one(sc, quiet_params(iterations=1, friction_static=0.5, friction_dynamic=0.0), dt=dt)
    x1, _ = o.state_original_order()
    assert x1[0, 1] == pytest.approx(F32(0.025), abs=1e-12)
    assert x1[0, 0] == pytest.approx(x0[0, 0] + F32(0.3) * dt, abs=1e-12)
    # dynamic friction saturating at 1 (min(mu_k |dn| / |dt|, 1)) removes it
    o, x0, _ = run_one(sc, quiet_params(iterations=1, friction_static=0.5, friction_dynamic=10.0), dt=dt)
    x1, _ = o.state_original_order()
    assert x1[0, 1] == pytest.approx(F32(0.025), abs=1e-12)
    assert x1[0, 0] == pytest.approx(x0[0, 0], abs=1e-12)


def rot_quat(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.array([*(axis * math.sin(ang / 2)), math.cos(ang / 2)])


def test_shape_matching_recovers_rigid_motion(have_oracle):
    q = rot_quat((1, 2, 3), 0.7)
    sc = make_scene()
    sc.add_rigid_box((3, 3, 3), (1.0, 1.0, 1.0), 0.05, 1.0, rotation=q)
    o, x0, _ = run_one(sc, quiet_params(iterations=1))
    x1, _ = o.state_original_order()
    qq, cm = o.rigid_state()
    # the extraction stops once |omega|^2 <= 1e-9 (shapematching.cuh:83): ~3e-5 rad
    assert np.abs(x1 - x0).max() < 5e-6
    assert min(np.abs(qq[0] - q).max(), np.abs(qq[0] + q).max()) < 1e-4
    assert np.allclose(cm[0], x0.mean(0), atol=1e-12)


def test_rotation_extraction_matches_polar_decomposition(have_oracle):
    rng = np.random.default_rng(3)
    for _ in range(5):
        R0 = scenes.quat_to_matrix(rot_quat(rng.normal(size=3), rng.uniform(0, 1.2)))
        S = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
        S = 0.5 * (S + S.T) + 0.5 * np.eye(3)
        A = R0 @ S
        U, _, Vt = np.linalg.svd(A)
        Rp = U @ Vt
        q = orc.extract_rotation(A.T.reshape(9), np.array([0, 0, 0, 1.0]), iterations=400)
        # converged to the stopping threshold |omega|^2 <= 1e-9 (shapematching.cuh:83)
        assert np.allclose(scenes.quat_to_matrix(q), Rp, atol=1e-4)
        assert not np.allclose(scenes.quat_to_matrix(q), R0, atol=1e-3) or np.allclose(Rp, R0, atol=1e-3)


def test_distance_constraint_restores_rest_length(have_oracle):
    sc = make_scene()
    s, e = sc.add_rope((0.5, 0.5, 0.5), (0.2, 0.0, 0.0), 2, 1.0)
    sc.distance_params[-1] = (0.15, 1.0)
    o, x0, _ = run_one(sc, quiet_params(iterations=1))
    x1, _ = o.state_original_order()
    assert np.linalg.norm(x1[1] - x1[0]) == pytest.approx(F32(0.15), abs=1e-12)
    assert np.allclose(x1.mean(0), x0.mean(0), atol=1e-12)  # equal masses: centre kept


def bending_scene(angle):
    sc = make_scene()
    # hinge along x through p1 p2, wings p3, p4 at dihedral `angle`
    pts = np.array([[0.5, 0.5, 0.5], [0.6, 0.5, 0.5],
                    [0.55, 0.5 + 0.1, 0.5],
                    [0.55, 0.5 + 0.1 * math.cos(angle), 0.5 + 0.1 * math.sin(angle)]])
    for p in pts:
        sc.add_granulars_block((1, 1, 1), p, (0.05, 0.05, 0.05), 1.0)
    sc.bending_constraints.append((0, 1, 2, 3))
    return sc, pts


def dihedral(p):
    n1 = np.cross(p[1] - p[0], p[2] - p[0])
    n2 = np.cross(p[1] - p[0], p[3] - p[0])
    return math.acos(np.clip(n1 @ n2 / np.linalg.norm(n1) / np.linalg.norm(n2), -1, 1))


def test_bending_flat_is_fixed_point_and_bent_relaxes(have_oracle):
    prm = quiet_params(iterations=1, friction_static=0.0)
    sc, pts = bending_scene(math.pi)
    o, x0, _ = run_one(sc, prm)
    x1, _ = o.state_original_order()
    assert np.abs(x1 - x0).max() < 1e-12
    # bent hinge: dihedral angle between normals (acos d, bending.cuh:61) moves toward pi
    sc, pts = bending_scene(1.2)
    o, x0, _ = run_one(sc, prm)
    x1, _ = o.state_original_order()
    assert dihedral(x1) > dihedral(x0) + 1e-3


def wind_scene(normal_along_flow: bool, v=(0.0, 0.0, -4.0)):
    sc = make_scene()
    if normal_along_flow:
        tri = np.array([[0.5, 0.5, 0.5], [0.75, 0.5, 0.5], [0.5, 0.75, 0.5]])
    else:
        tri = np.array([[0.5, 0.5, 0.5], [0.75, 0.5, 0.5], [0.5, 0.5, 0.75]])
    for p in tri:
        sc.add_granulars_block((1, 1, 1), p, (0.05, 0.05, 0.05), 2.0)
    sc.faces.append((0, 1, 2))
    sc._vel = [np.tile(np.asarray(v, np.float32), (3, 1))]
    sc._has_vel = True
    return sc, tri


def test_wind_drag_perpendicular_face(have_oracle):
    sc, tri = wind_scene(True)
    dt = F32(1.0 / 120.0)
    o, x0, v0 = run_one(sc, quiet_params(iterations=0, drag_coeff=10.0, lift_coeff=100.0), dt=dt)
    x1, _ = o.state_original_order()
    move = x1 - x0 - v0 * dt
    area = 0.5 * 0.25 * 0.25
    vv = np.array([0.0, 0.0, -4.0])
    force = -10.0 * (vv @ vv) * area * vv / np.linalg.norm(vv)  # lift is zero: n parallel to v
    assert np.allclose(move, np.tile(force * dt * dt / 2.0, (3, 1)), rtol=1e-9, atol=1e-15)


def test_wind_parallel_face_and_wind_vector(have_oracle):
    sc, _ = wind_scene(False)
    dt = F32(1.0 / 120.0)
    o, x0, v0 = run_one(sc, quiet_params(iterations=0), dt=dt)
    x1, _ = o.state_original_order()
    assert np.abs(x1 - x0 - v0 * dt).max() < 1e-15
    # air moving with the particles: zero relative velocity, zero force
    sc, _ = wind_scene(True)
    o, x0, v0 = run_one(sc, quiet_params(iterations=0, wind=(0.0, 0.0, -4.0)), dt=dt)
    x1, _ = o.state_original_order()
    assert np.abs(x1 - x0 - v0 * dt).max() < 1e-15


def test_particle_particle_contact_separates_to_two_radii(have_oracle):
    sc = make_scene()
    sc.add_granulars_block((2, 1, 1), (0.5, 0.5, 0.5), (0.03, 0.05, 0.05), 1.0)
    o, x0, _ = run_one(sc, quiet_params(iterations=1))
    x1, _ = o.state_original_order()
    assert np.linalg.norm(x1[1] - x1[0]) == pytest.approx(2 * F32(0.025), abs=1e-12)


def test_immovable_and_pick(have_oracle):
    sc = make_scene()
    sc.add_granulars_block((3, 1, 1), (0.5, 0.5, 0.5), (0.2, 0.05, 0.05), 1.0)
    sc.make_immovable(1)
    sc._vel = [np.tile(np.array([1.0, 0, 0], np.float32), (3, 1))]
    sc._has_vel = True
    o, x0, _ = run_one(sc, quiet_params(iterations=2))
    x1, v1 = o.state_original_order()
    assert np.array_equal(x1[1], x0[1]) and np.all(v1[1] == 0.0)
    assert x1[0, 0] > x0[0, 0]
    # picking (particlesolver.cuh:790-794,1066-1071): pinned at the picked position, given velocity
    o = orc.Oracle(bp.SceneArrays(sc), quiet_params(iterations=2), fp64=True)
    o.step(2, 1 / 60, picked=2, pos=(1.0, 1.0, 1.0), vel=(0.5, 0.0, 0.0))
    x2, v2 = o.state_original_order()
    assert np.allclose(x2[2], (1.0, 1.0, 1.0)) and np.allclose(v2[2], (0.5, 0.0, 0.0))


# ---------------------------------------------------------------- known answers for single terms
# Each test below sets up a few particles where one term of the substep acts
# alone, and derives the expected motion in numpy from the mathematics the
# term implements.  The docstrings name the slips in the oracle each one was
# checked to catch (the slip planted in a local copy of bp_oracle.cpp, CPU).
E_REF = float(np.float32(2.7182818284))  # the reference's e literal as an FP32 constant (particlesolver.cuh:94)
DT_POW2 = 1.0 / 128.0  # 1/dt exact in FP32: (x* - x) * (1/dt) returns v to FP64 rounding


def _set_particles(sc, pos=None, vel=None, mass=None):
    if pos is not None:
        sc._pos = [np.asarray(pos, np.float32).reshape(-1, 3)]
    if vel is not None:
        sc._vel = [np.asarray(vel, np.float32).reshape(-1, 3)]
        sc._has_vel = True
    if mass is not None:
        sc._mass = [np.asarray(mass, np.float32).reshape(-1)]


def _f64(a):
    return np.asarray(a, np.float32).astype(np.float64)


def test_mass_scaling_splits_distance_correction(have_oracle):
    """computeInvScaledMasses (particlesolver.cuh:84-98) feeding one distance
    constraint (distance.cuh:3-41).  The shock-propagation weight is
    w_i = 1 / (e^(-k y_i) m_i) with y the height of x (the position at the
    start of the substep, particlesolver.cuh:812-816), not of x*.  A stretched
    pair of rest length L moves each end along the pair by w_i / (w_1 + w_2)
    of (d - L), the first toward the second and the second toward the first.
    Masses 1 and 3 at heights 0.3 and 0.75, k = 2, moving vertically so that
    y(x*) differs from y(x).  Catches: the exponent's sign flipped, the
    exponent dropped (w = 1/m), the mass dropped, x* used for the height,
    either end's weight swapped."""
    k = 2.0
    sc = make_scene()
    sc.add_rope((0.5, 0.3, 0.5), (0.12, 0.45, 0.0), 2, 1.0)
    sc.distance_params[-1] = (0.3, 1.0)
    x0 = np.array([[0.5, 0.3, 0.5], [0.62, 0.75, 0.5]])
    v0 = np.array([[0.0, 6.0, 0.0], [0.0, -3.0, 0.0]])
    m = np.array([1.0, 3.0])
    _set_particles(sc, x0, v0, m)
    o, _, _ = run_one(sc, quiet_params(iterations=1, mass_scaling_k=k), dt=DT_POW2)
    x1, _ = o.state_original_order()
    x0, v0, m = _f64(x0), _f64(v0), _f64(m)
    xs = x0 + v0 * DT_POW2
    w = E_REF ** (k * x0[:, 1]) / m
    d = np.linalg.norm(xs[0] - xs[1])
    n = (xs[0] - xs[1]) / d
    corr = (d - F32(0.3)) * n
    expect = xs + np.stack([-w[0] / w.sum() * corr, w[1] / w.sum() * corr])
    assert np.abs(x1 - expect).max() < 1e-12
    assert abs(w[0] / w.sum() - 0.75) > 0.1  # far from the unscaled 1/m split (0.75 : 0.25)


def test_wind_lift_oblique_face(have_oracle):
    """applyWindForce (wind.cuh:3-50) with the build's wind vector W: a face at
    45 degrees to the relative air velocity v = mean vertex velocity - W.
    With n the unit face normal turned so n . v > 0, v^ = v/|v| and A the
    area, the force is drag cd |v|^2 A (n.v^)(-v^) plus lift
    cl |v|^2 A (n.v^) u, u = (n x v^) x v^ = (n.v^) v^ - n (the vector
    identity, written here without cross products), and every vertex moves
    F dt^2 / m_vertex.  Lift is perpendicular to v and dominates (cl = 100).
    Catches: the lift's sign (v^ x (n x v^)), the lift or drag dropped,
    the (n.v^) factor dropped from the lift, + W instead of - W, one vertex's
    mass used for all three, F/3 per vertex."""
    sc = make_scene()
    c45 = math.cos(math.pi / 4)
    tri = np.array([[0.5, 0.5, 0.5], [0.75, 0.5, 0.5], [0.5, 0.5 + 0.25 * c45, 0.5 + 0.25 * c45]])
    for p in tri:
        sc.add_granulars_block((1, 1, 1), p, (0.05, 0.05, 0.05), 1.0)
    sc.faces.append((0, 1, 2))
    vel = np.tile(np.array([0.0, 0.0, -4.0]), (3, 1))
    m = np.array([1.0, 2.0, 4.0])
    _set_particles(sc, vel=vel, mass=m)
    wind = (1.0, 0.0, 0.5)
    cd, cl = 10.0, 100.0
    o, x0, v0 = run_one(sc, quiet_params(iterations=0, drag_coeff=cd, lift_coeff=cl, wind=wind), dt=DT_POW2)
    x1, _ = o.state_original_order()
    xs = x0 + v0 * DT_POW2
    vrel = v0.mean(0) - _f64(wind)
    vhat = vrel / np.linalg.norm(vrel)
    nrm = np.cross(xs[1] - xs[0], xs[2] - xs[0])
    area = 0.5 * np.linalg.norm(nrm)
    nrm /= np.linalg.norm(nrm)
    if nrm @ vrel < 0:
        nrm = -nrm
    nv = nrm @ vhat
    assert 0.5 < nv < 0.9  # oblique: neither parallel nor perpendicular
    q = vrel @ vrel
    drag = cd * q * area * nv * (-vhat)
    lift = cl * q * area * nv * (nv * vhat - nrm)
    assert abs(lift @ vhat) < 1e-12 * np.linalg.norm(lift) and np.linalg.norm(lift) > 3 * np.linalg.norm(drag)
    expect = (drag + lift)[None, :] * DT_POW2 * DT_POW2 / _f64(m)[:, None]
    assert np.allclose(x1 - xs, expect, rtol=1e-9, atol=1e-16)


def _brute_eta(xs, ph, om, h, cs, dims):
    """eta_i = sum_{j != i, same phase, 0 < r^2 <= h^2} |omega_j| gradW(x_i - x_j),
    gradW(v) = -45/(pi h^6) (h - r)^2 v / r: the SPH estimate of grad |omega|
    (Macklin and Mueller 2013, eq. 16), over all pairs whose hashed cells are
    within one cell of each other (the 27-cell search of fluid.cuh:298-322;
    the FP32 hash can place a particle near a cell face one cell over)."""
    c2 = -45.0 / (PI_REF * h ** 6)
    d = xs[:, None, :] - xs[None, :, :]
    r2 = (d ** 2).sum(-1)
    dd = np.asarray(dims)
    g = ((xs.astype(np.float32) + np.float32(50000)) / np.asarray(cs, np.float32)).astype(np.int64) & (dd - 1)
    near = np.all((g[None, :, :] - g[:, None, :] + 1) % dd <= 2, axis=2)
    sel = (r2 > 0) & (r2 <= h * h) & (ph[:, None] == ph[None, :]) & near
    r = np.sqrt(np.where(sel, r2, 1.0))
    coef = np.where(sel, c2 * (h - r) ** 2 / r, 0.0) * np.linalg.norm(om, axis=1)[None, :]
    return (coef[:, :, None] * d).sum(1)


def test_vorticity_confinement_update_and_cutoff(have_oracle):
    """fluidVorticity (fluid.cuh:281-336): with eta the SPH gradient of
    |omega| (fluid.cuh:304-326), a particle whose |eta|^2 exceeds 1e-3
    (fluid.cuh:329) gets v += eps (eta/|eta| x omega_i) dt, any other keeps v.
    A 7^3 lattice in slow rigid rotation about z.  omega comes from the
    oracle's omega pass (pinned by test_vorticity_rigid_rotation_closed_form),
    eta from the brute-force sum.  On the unjittered lattice the direction of
    eta for the top-face centre particle is fixed by geometry: every
    neighbour is level or below, so eta points down.  On a jittered lattice,
    with the spin scaled so that the median |eta|^2 sits at the cut-off,
    particles on both sides of it (down to |eta|^2 = 1e-5) are checked.
    Catches: the cross product's operands swapped, eta not normalised, eps or
    dt dropped, the cut-off removed or applied to |eta| instead of |eta|^2,
    |omega_j| replaced by |omega_i|."""
    h = 0.1
    eps = 1e4  # a free parameter: large, so the update stands far above the rounding of v

    def run(spin, jitter):
        sc = lattice_scene(n=7, jitter=jitter)
        x = _f64(sc.positions)
        v = np.cross(np.array([0.0, 0.0, spin]), x - x.mean(0))
        _set_particles(sc, vel=v)
        o, x0, v0 = run_one(sc, quiet_params(vorticity_eps=eps, lambda_denominator=1e30, scorr_k=0.0),
                            dt=DT_POW2)
        pos, vel, mp = o.get_state()
        om = o.debug(4)
        eta = _brute_eta(pos, sc.phases[mp], om, F32(h), sc.cell_size, sc.grid_dims)
        e2 = (eta ** 2).sum(1)
        nh = eta / np.sqrt(np.maximum(e2, 1e-300))[:, None]
        return x0[mp], vel - v0[mp], e2, nh, eps * np.cross(nh, om) * DT_POW2

    def check(dv, e2, expect):
        above, below = e2 > 1.05e-3, e2 < 0.95e-3
        scale = np.abs(expect[above]).max()
        assert scale > 1e-9
        assert np.abs(dv[above] - expect[above]).max() < 1e-6 * scale
        assert np.abs(dv[below]).max() < 1e-9 * scale
        return above.sum(), (below & (e2 > 1e-5)).sum()

    for jitter in (0.0, 0.004):
        e2 = run(1.0, jitter)[2]
        q = 1.0 if jitter == 0.0 else 0.5  # top of the range, or its median
        spin = math.sqrt(0.05 if jitter == 0.0 else 1e-3) / math.sqrt(np.quantile(e2, q))
        x0s, dv, e2, nh, expect = run(spin, jitter)
        n_above, n_below = check(dv, e2, expect)
        assert n_above >= 20
        if jitter == 0.0:
            top = np.argmin(np.abs(x0s - np.array([x0s[:, 0].mean(), x0s[:, 1].max(), x0s[:, 2].mean()])).sum(1))
            assert np.allclose(nh[top], (0.0, -1.0, 0.0), atol=1e-5)  # FP32 lattice: symmetric to ~1e-6
        else:
            assert n_below >= 20


def test_xsph_two_velocities_and_coincident_pair(have_oracle):
    """fluidXSph (fluid.cuh:338-385): v_i' = v_i + c sum_j (v_j - v_i) W(r^2)
    over distinct same-phase j with r^2 <= h^2, r = 0 included, no density
    weighting.  A and B of one fluid block meet at one point after the
    prediction (r = 0, W(0) = 315 / (64 pi h^3)), D of the same block sits
    0.06 away, C of another fluid block 0.03 away (excluded by phase).
    Expected values from the poly6 closed form.  Catches: (v_i - v_j)
    instead of (v_j - v_i), c dropped, the r = 0 pair dropped (d2 > 0),
    the phase filter dropped, an m/rho weight, the spiky kernel for poly6."""
    h, c = F32(0.1), 0.01
    sc = make_scene()
    sc.add_fluid_block((3, 1, 1), (0.5, 0.5, 0.5), (0.04, 0.0, 0.0), 1.0)  # A, B, D: one phase
    sc.add_fluid_block((1, 1, 1), (0.5, 0.53, 0.5), 0.05, 1.0)  # C: another phase
    p = np.array([0.5, 0.5, 0.5])
    dtv = np.array([DT_POW2, 0.0, 0.0])
    pos = np.array([p - dtv, p + dtv, p + np.array([0.0, 0.06, 0.0]), p + np.array([0.0, 0.0, 0.03])])
    vel = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 0.5, 0.0], [0.3, 0.3, 0.3]])
    _set_particles(sc, pos, vel)
    o, x0, v0 = run_one(sc, quiet_params(xsph_c=c, lambda_denominator=1e30, scorr_k=0.0), dt=DT_POW2)
    x1, v1 = o.state_original_order()
    assert np.array_equal(x1[0], x1[1])  # A and B coincide when XSPH runs
    W = lambda r2: 315.0 / (64.0 * PI_REF * h ** 9) * (h * h - r2) ** 3
    r2 = ((x1[0] - x1[2]) ** 2).sum()
    cF = F32(c)
    expect = np.array([
        v0[0] + cF * ((v0[1] - v0[0]) * W(0.0) + (v0[2] - v0[0]) * W(r2)),
        v0[1] + cF * ((v0[0] - v0[1]) * W(0.0) + (v0[2] - v0[1]) * W(r2)),
        v0[2] + cF * ((v0[0] - v0[2]) * W(r2) + (v0[1] - v0[2]) * W(r2)),
        v0[3],
    ])
    assert np.allclose(v1, expect, rtol=1e-9, atol=1e-12)


def _contact_scene(pos, vel, mass):
    sc = make_scene()
    for q in pos:
        sc.add_granulars_block((1, 1, 1), q, (0.05, 0.05, 0.05), 1.0)
    _set_particles(sc, pos, vel, mass)
    return sc


def _tangential(v, n):
    return v - (v @ n) * n


def test_contact_static_friction_stops_relative_slip(have_oracle):
    """particleParticleCollisionConstraint (particleparticle.cuh:5-101),
    static branch, one contact between unequal masses (1 and 3).  Each
    particle writes only its own correction, w_i/(w_i + w_j) of the push to
    2r plus, when the slip is small (|tangential slip| <= mu_s |own push|),
    minus w_i/(w_i + w_j) of the tangential relative displacement.  The two
    friction shares add to the whole slip, so after the pass the particles
    have no relative motion tangential to the contact normal (Macklin et al.
    2014, eq. 24), and since w = 1/m the corrections carry no momentum.
    Catches: the friction's sign, its w_i/(w_i + w_j) weight dropped or
    swapped, the static branch skipped."""
    r, mu_s = F32(0.025), 0.5
    pos = np.array([[0.5, 0.54, 0.5], [0.5, 0.5, 0.5]])
    vel = np.array([[0.1, 0.0, 0.0], [0.0, 0.0, 0.0]])
    sc = _contact_scene(pos, vel, (1.0, 3.0))
    o, x0, v0 = run_one(sc, quiet_params(iterations=1, friction_static=mu_s, friction_dynamic=0.0), dt=DT_POW2)
    x1, _ = o.state_original_order()
    xs = x0 + v0 * DT_POW2
    d = np.linalg.norm(xs[0] - xs[1])
    n = (xs[0] - xs[1]) / d
    slip = _tangential((xs[0] - x0[0]) - (xs[1] - x0[1]), n)
    # both sides take the static branch: |slip| <= mu_s w_i/(w_i + w_j) (2r - d) for w = 1 and 1/3
    assert 5e-4 < np.linalg.norm(slip) < mu_s * 0.25 * (2 * r - d)
    rel = _tangential((x1[0] - x0[0]) - (x1[1] - x0[1]), n)
    assert np.abs(rel).max() < 1e-15
    assert np.abs(1.0 * (x1[0] - xs[0]) + 3.0 * (x1[1] - xs[1])).max() < 1e-15


@pytest.mark.parametrize("mu_k", [0.3, 50.0])
def test_contact_dynamic_friction_reduces_slip(have_oracle, mu_k):
    """particleParticleCollisionConstraint, dynamic branch (slip above
    mu_s times the own push), equal masses: each side removes
    min(mu_k |own push| / |slip|, 1) of its half of the tangential slip, so
    the relative tangential motion after the pass is (1 - f) times the slip
    with f = min(mu_k (pen / 2) / |slip|, 1), pen = 2r - d (particleparticle
    .cuh:88-92; the reference's |own push| is half the penetration here).
    mu_k = 50 saturates f at 1.  Catches: the dynamic factor inverted
    (|slip| / |push|), the min with 1 dropped, mu_s used for mu_k."""
    r = F32(0.025)
    pos = np.array([[0.5, 0.54, 0.5], [0.5, 0.5, 0.5]])
    vel = np.array([[2.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    sc = _contact_scene(pos, vel, (1.0, 1.0))
    o, x0, v0 = run_one(sc, quiet_params(iterations=1, friction_static=0.5, friction_dynamic=mu_k), dt=DT_POW2)
    x1, _ = o.state_original_order()
    xs = x0 + v0 * DT_POW2
    d = np.linalg.norm(xs[0] - xs[1])
    n = (xs[0] - xs[1]) / d
    slip = _tangential((xs[0] - x0[0]) - (xs[1] - x0[1]), n)
    pen = 2 * r - d
    assert np.linalg.norm(slip) > 0.5 * 0.5 * pen  # dynamic branch
    f = min(F32(mu_k) * 0.5 * pen / np.linalg.norm(slip), 1.0)
    assert (f < 0.5) if mu_k < 1 else (f == 1.0)
    rel = _tangential((x1[0] - x0[0]) - (x1[1] - x0[1]), n)
    assert np.abs(rel - (1.0 - f) * slip).max() < 1e-14


def test_contact_friction_averaged_over_contacts(have_oracle):
    """particleParticleCollisionConstraint: the pushes of all contacts add,
    the friction terms are averaged over the contact count
    (particleparticle.cuh:92-100, "averaging constraints is very important").
    Particle i slides in +x over j and k, placed as mirror images in z
    (M = diag(1, 1, -1)), all masses 1.  Run once with j alone: i's own
    correction is the push P_j = 1/2 (x*_i - x*_j)(2r/d - 1) plus friction
    F_j.  With j and k, symmetry gives i's correction as
    P_j + M P_j + (F_j + M F_j) / 2.  Catches: the average dropped (friction
    summed), the pushes averaged too, the count taken over all candidates."""
    r = F32(0.025)
    # offsets of 2^-5: exact in FP32, so j and k are exact mirror images about z = 0.5
    pi_, pj, pk = [0.5, 0.53125, 0.5], [0.5, 0.5, 0.46875], [0.5, 0.5, 0.53125]
    vi = [2.0, 0.0, 0.0]
    prm = quiet_params(iterations=1, friction_static=0.5, friction_dynamic=0.3)
    a = _contact_scene(np.array([pi_, pj]), np.array([vi, [0, 0, 0]]), (1.0, 1.0))
    oa, xa0, va0 = run_one(a, prm, dt=DT_POW2)
    b = _contact_scene(np.array([pi_, pj, pk]), np.array([vi, [0, 0, 0], [0, 0, 0]]), (1.0, 1.0, 1.0))
    ob, xb0, vb0 = run_one(b, prm, dt=DT_POW2)
    xa1, _ = oa.state_original_order()
    xb1, _ = ob.state_original_order()
    xs = xa0 + va0 * DT_POW2
    diff = xs[0] - xs[1]
    d = np.linalg.norm(diff)
    assert d < 2 * r and np.linalg.norm(np.array(pj) - np.array(pk)) > 2 * r
    push = 0.5 * diff * (2 * r / d - 1.0)
    fric = (xa1[0] - xs[0]) - push
    assert np.linalg.norm(fric) > 1e-5
    M = np.array([1.0, 1.0, -1.0])
    expect = push + M * push + 0.5 * (fric + M * fric)
    assert np.abs((xb1[0] - (xb0[0] + vb0[0] * DT_POW2)) - expect).max() < 1e-14


def _hinge_d(p):
    """n1 . n2 of the hinge (p1, p2 edge; wings p3, p4), bending.cuh:24-37."""
    a, b, c = p[1] - p[0], p[2] - p[0], p[3] - p[0]
    n1, n2 = np.cross(a, b), np.cross(a, c)
    return float(n1 @ n2 / np.linalg.norm(n1) / np.linalg.norm(n2))


@pytest.mark.parametrize("angle,shift", [(1.2, 0.0), (2.0, 0.03), (2.8, -0.02)])
def test_bending_projection_is_pbd_step_of_dihedral_angle(have_oracle, angle, shift):
    """bendingConstraints (bending.cuh:3-76), Mueller et al. 2007: the
    constraint C = acos(d) - pi, d = n1 . n2, projected by one PBD step
    dp_i = -k s w_i grad_i C, s = C / sum_j w_j |grad_j C|^2.  With
    grad C = -grad d / sqrt(1 - d^2) this is
    dp_i = k w_i C sqrt(1 - d^2) grad_i d / sum_j w_j |grad_j d|^2.
    grad d is taken here by central differences of d(p) in FP64, not from the
    closed-form gradients the kernel uses.  Masses 1..4, wings at unequal
    positions along the edge.  Catches: two of the gradient vectors swapped,
    a sign slip in any of them, kBend or a mass dropped, sqrt(1 - d^2)
    dropped, the rest angle changed."""
    kb = 0.8
    sc, pts = bending_scene(angle)
    pts = pts.copy()
    pts[3, 0] += shift
    _set_particles(sc, pts, mass=(1.0, 2.0, 3.0, 4.0))
    o, x0, _ = run_one(sc, quiet_params(iterations=1, friction_static=0.0, bend_stiffness=kb))
    x1, _ = o.state_original_order()
    p = x0.copy()
    d = _hinge_d(p)
    g = np.zeros((4, 3))
    e = 1e-7
    for i in range(4):
        for a in range(3):
            pp, pm = p.copy(), p.copy()
            pp[i, a] += e
            pm[i, a] -= e
            g[i, a] = (_hinge_d(pp) - _hinge_d(pm)) / (2 * e)
    w = 1.0 / np.array([1.0, 2.0, 3.0, 4.0])
    C = math.acos(d) - F32(3.141592)
    dp = F32(kb) * w[:, None] * C * math.sqrt(1 - d * d) * g / (w * (g ** 2).sum(1)).sum()
    assert np.abs(dp).max() > 1e-3
    assert np.allclose(x1 - x0, dp, rtol=1e-6, atol=1e-9)
