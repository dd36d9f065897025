"""Scene description API and the benchmark scenes.

``Scene`` mirrors brass-pan's ``struct Scene`` (brasstower/scene.h:455-536,
brasstower/scene.cpp) with the same member names in snake case.  Differences:

* ``add_fluid_block``, ``add_rigid_box``, ``add_rope`` and ``make_immovable``
  are declared but not implemented in the reference (scene.h:470-496); they
  are implemented here.
* ``add_cloth`` follows scene.cpp:4-128 (structural, shear and "hack bending"
  distance constraints, two faces per quad) but indexes particles in the order
  they are stored and offsets them by the particles already in the scene, and
  pushes masses and group ids (the reference pushes only positions and phases).
* rigid rest positions are stored by original particle id for any insertion
  order (the reference indexes them by global id but sizes them per body,
  particlesolver.cuh:262,454, so rigid bodies had to come first).
* grid cell size and per-axis power-of-two grid dimensions are explicit
  (the reference hard-codes cellSize = 2.3 r and a 512^3 grid,
  particlesolver.cuh:150-151).

Random numbers: splitmix64, float = (u64 >> 40) * 2^-24, so host and GPU box
generate the same scenes.
"""
from __future__ import annotations

import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """First `count` outputs of splitmix64 started at `seed` (uint64)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, count: int) -> np.ndarray:
    """Uniform floats in [0, 1): (u64 >> 40) * 2^-24, as float64."""
    return (splitmix64(seed, count) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)


class Rng:
    """Sequential splitmix64 stream."""

    def __init__(self, seed: int):
        self.seed = int(seed)
        self.used = 0

    def uniform(self, count: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            start = (np.uint64(self.seed) + np.uint64(self.used) * GOLDEN)
        self.used += count
        return uniform01(int(start), count)


def quat_to_matrix(q):
    x, y, z, w = q
    return np.array([
        [1 - 2 * y * y - 2 * z * z, 2 * x * y - 2 * w * z, 2 * x * z + 2 * w * y],
        [2 * x * y + 2 * w * z, 1 - 2 * x * x - 2 * z * z, 2 * y * z - 2 * w * x],
        [2 * x * z - 2 * w * y, 2 * y * z + 2 * w * x, 1 - 2 * x * x - 2 * y * y],
    ])


class Scene:
    """Host scene description (scene.h:455-536)."""

    def __init__(self, particle_radius: float = 0.05, fluid_kernel_radius: float | None = None,
                 cell_size=None, grid_dims=(512, 512, 512)):
        self.particle_radius = float(particle_radius)
        # scene.h:501: fluidKernelRadius = 2.3 * particleRadius
        self.fluid_kernel_radius = float(fluid_kernel_radius) if fluid_kernel_radius else 2.3 * particle_radius
        # particlesolver.cuh:150: cellSize = particleRadius * 2.3f
        cs = cell_size if cell_size is not None else np.float32(particle_radius) * np.float32(2.3)
        self.cell_size = tuple(float(c) for c in np.broadcast_to(np.float32(cs), (3,)))
        self.grid_dims = tuple(int(d) for d in grid_dims)
        self._pos, self._vel, self._mass, self._phase, self._group = [], [], [], [], []
        self._n = 0
        self.planes = []
        self.faces = []
        self.rigidbody_particle_id_ranges = []
        self._rigid_rest = {}
        self.distance_pairs = []
        self.distance_params = []
        self.bending_constraints = []
        self.immovable_constraints = []
        self.group_id_counter = 1
        self.solid_phase_counter = 1
        self.fluid_phase_counter = -1
        self._has_vel = False

    # ------------------------------------------------------------ particle data
    def num_particles(self) -> int:
        return self._n

    def _push(self, pos, mass, phases, group, vel=None):
        pos = np.asarray(pos, np.float32).reshape(-1, 3)
        n = len(pos)
        start = self._n
        self._pos.append(pos)
        if vel is not None:
            self._has_vel = True
            self._vel.append(np.asarray(vel, np.float32).reshape(-1, 3))
        else:
            self._vel.append(np.zeros((n, 3), np.float32))
        self._mass.append(np.broadcast_to(np.float32(mass), (n,)).astype(np.float32))
        self._phase.append(np.asarray(phases, np.int32).reshape(n))
        self._group.append(np.full(n, group, np.int32))
        self._n += n
        return start, self._n

    @property
    def positions(self) -> np.ndarray:
        return np.concatenate(self._pos) if self._pos else np.zeros((0, 3), np.float32)

    @property
    def velocities(self):
        return np.concatenate(self._vel) if self._has_vel else None

    @property
    def masses(self) -> np.ndarray:
        return np.concatenate(self._mass) if self._mass else np.zeros(0, np.float32)

    @property
    def phases(self) -> np.ndarray:
        return np.concatenate(self._phase) if self._phase else np.zeros(0, np.int32)

    @property
    def group_ids(self) -> np.ndarray:
        return np.concatenate(self._group) if self._group else np.zeros(0, np.int32)

    @property
    def rigidbody_initial_positions(self) -> np.ndarray:
        """Rest offsets by original id, covering ids [0, last rigid particle]."""
        if not self._rigid_rest:
            return np.zeros((0, 3), np.float32)
        end = max(s + len(r) for s, r in self._rigid_rest.items())
        out = np.zeros((end, 3), np.float32)
        for s, r in self._rigid_rest.items():
            out[s:s + len(r)] = r
        return out

    # ------------------------------------------------------------ builders
    def add_plane(self, origin, normal):
        n = np.asarray(normal, np.float64)
        n = n / np.linalg.norm(n)
        self.planes.append((np.asarray(origin, np.float32), n.astype(np.float32)))

    def add_box_planes(self, lo, hi):
        """Six inward-facing planes of the box [lo, hi]."""
        lo = np.asarray(lo, np.float32)
        hi = np.asarray(hi, np.float32)
        self.box = (lo, hi)
        for a in range(3):
            n = np.zeros(3, np.float32)
            n[a] = 1.0
            self.add_plane(lo, n)
            self.add_plane(hi, -n)

    def add_granulars_block(self, dimension, start_position, step, mass_per_particle):
        """scene.cpp:130-146: one solid phase per particle."""
        dim = [int(d) for d in dimension]
        idx = np.indices(dim).reshape(3, -1).T.astype(np.float32)
        pos = np.asarray(start_position, np.float32) + idx * np.asarray(step, np.float32)
        n = len(pos)
        phases = np.arange(self.solid_phase_counter, self.solid_phase_counter + n, dtype=np.int32)
        self.solid_phase_counter += n
        group = self.group_id_counter
        self.group_id_counter += 1
        return self._push(pos, mass_per_particle, phases, group)

    def add_fluid_block(self, dimension, start_position, step, mass_per_particle, jitter=0.0, seed=0,
                        rng: Rng | None = None):
        """Fluid lattice (declared at scene.h:475), x-major like addGranularsBlock.
        One negative phase per block (particlesolver.cuh:501-505).  `jitter`
        adds U(-jitter, jitter) per axis from the splitmix64 stream."""
        dim = [int(d) for d in dimension]
        idx = np.indices(dim).reshape(3, -1).T.astype(np.float64)
        pos = np.asarray(start_position, np.float64) + idx * np.asarray(step, np.float64)
        if jitter:
            r = rng if rng is not None else Rng(seed)
            u = r.uniform(3 * len(pos)).reshape(-1, 3)
            pos = pos + (2.0 * u - 1.0) * jitter
        phase = self.fluid_phase_counter
        self.fluid_phase_counter -= 1
        group = self.group_id_counter
        self.group_id_counter += 1
        return self._push(pos.astype(np.float32), mass_per_particle, np.full(len(pos), phase), group)

    def add_rigid_box(self, dimension, start_position, step, mass_per_particle, rotation=(0, 0, 0, 1)):
        """Shape-matched box (declared at scene.h:487): lattice of `dimension`
        particles rotated by quaternion `rotation` (x,y,z,w) about its centre,
        which sits at start_position + (dimension-1)*step/2.  Rest offsets are
        relative to the centre of mass (scene.h:521, particlesolver.cuh:444-448)."""
        dim = [int(d) for d in dimension]
        idx = np.indices(dim).reshape(3, -1).T.astype(np.float64)
        local = idx * np.asarray(step, np.float64)
        local -= local.mean(axis=0)
        centre = np.asarray(start_position, np.float64) + (np.asarray(dim) - 1) * np.asarray(step) / 2.0
        pos = local @ quat_to_matrix(rotation).T + centre
        rest = local.astype(np.float32)
        rest -= rest.mean(axis=0, dtype=np.float64).astype(np.float32)
        phase = self.solid_phase_counter
        self.solid_phase_counter += 1
        group = self.group_id_counter
        self.group_id_counter += 1
        s, e = self._push(pos.astype(np.float32), mass_per_particle, np.full(len(pos), phase), group)
        self.rigidbody_particle_id_ranges.append((s, e))
        self._rigid_rest[s] = rest
        return s, e

    def add_rope(self, start_position, step, num_joint, mass_per_particle, k_stiffness=1.0):
        """Chain of distance constraints (declared at scene.h:491)."""
        idx = np.arange(num_joint, dtype=np.float32)[:, None]
        pos = np.asarray(start_position, np.float32) + idx * np.asarray(step, np.float32)
        phases = np.arange(self.solid_phase_counter, self.solid_phase_counter + num_joint, dtype=np.int32)
        self.solid_phase_counter += num_joint
        group = self.group_id_counter
        self.group_id_counter += 1
        s, e = self._push(pos, mass_per_particle, phases, group)
        seg = float(np.linalg.norm(np.asarray(step, np.float32)))
        for i in range(num_joint - 1):
            self.distance_pairs.append((s + i, s + i + 1))
            self.distance_params.append((seg, k_stiffness))
        return s, e

    def add_cloth(self, start_position, step_x, step_y, num_joint_x, num_joint_y, mass_per_particle,
                  k_stiffness, k_bending, is_self_collidable=True, dihedral_bending=False):
        """scene.cpp:4-128 with storage-order indices and global offsets.
        dihedral_bending adds one hinge per quad, (p3, p2, p1, p4) or
        (p1, p4, p3, p2) by the parity of x + y, as the reference's
        commented-out generator (scene.cpp:108-123)."""
        sx = np.asarray(step_x, np.float32)
        sy = np.asarray(step_y, np.float32)
        nx, ny = int(num_joint_x), int(num_joint_y)
        lx = float(np.sqrt(np.float32(np.dot(sx, sx))))
        ly = float(np.sqrt(np.float32(np.dot(sy, sy))))
        sxy = sx + sy
        ld = float(np.sqrt(np.float32(np.dot(sxy, sxy))))
        xs, ys = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        pos = (np.asarray(start_position, np.float32) + sy * ys.reshape(-1, 1).astype(np.float32)
               + sx * xs.reshape(-1, 1).astype(np.float32))
        n = nx * ny
        if is_self_collidable:
            phases = np.arange(self.solid_phase_counter, self.solid_phase_counter + n, dtype=np.int32)
            self.solid_phase_counter += n
        else:
            phases = np.full(n, self.solid_phase_counter, np.int32)
            self.solid_phase_counter += 1
        group = self.group_id_counter
        self.group_id_counter += 1
        base, end = self._push(pos, mass_per_particle, phases, group)
        P = lambda x, y: base + x * ny + y
        for x in range(nx):
            for y in range(ny):
                p1, p2, p3, p4 = P(x, y), P(x + 1, y), P(x, y + 1), P(x + 1, y + 1)
                if x < nx - 1 and y < ny - 1:
                    self.faces.append((p1, p2, p3))
                    self.faces.append((p2, p3, p4))
                if x < nx - 1:
                    self.distance_pairs.append((p1, p2))
                    self.distance_params.append((lx, k_stiffness))
                if y < ny - 1:
                    self.distance_pairs.append((p1, p3))
                    self.distance_params.append((ly, k_stiffness))
                if x < nx - 1 and y < ny - 1:
                    self.distance_pairs.append((p1, p4))
                    self.distance_params.append((ld, k_stiffness))
                    self.distance_pairs.append((p2, p3))
                    self.distance_params.append((ld, k_stiffness))
                if 0 < x < nx - 1 and 0 < y < ny - 1:
                    for j in (-1, 0, 1):
                        self.distance_pairs.append((P(x - 1, y + j), P(x + 1, y + j)))
                        self.distance_params.append((lx * 2, k_bending))
                    for i in (-1, 0, 1):
                        self.distance_pairs.append((P(x + i, y - 1), P(x + i, y + 1)))
                        self.distance_params.append((ly * 2, k_bending))
                if dihedral_bending and x < nx - 1 and y < ny - 1:
                    self.bending_constraints.append((p3, p2, p1, p4) if (x + y) % 2 else (p1, p4, p3, p2))
        return base, end

    def make_immovable(self, particle_id: int):
        """declared at scene.h:496"""
        self.immovable_constraints.append(int(particle_id))


# ---------------------------------------------------------------- SPH helpers
def poly6_f64(r2, h):
    return 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r2) ** 3


def lattice_rest_density(spacing: float, h: float, mass: float) -> float:
    """Mean density an interior particle of an unjittered cubic lattice gets
    from fluidLambda's half sum (j >= i only, fluid.cuh:117): every pair is
    counted once, so the mean is m (W(0) + 1/2 sum_{j != i} W(r_ij^2))."""
    k = int(math.ceil(h / spacing))
    total = 0.0
    for a in range(-k, k + 1):
        for b in range(-k, k + 1):
            for c in range(-k, k + 1):
                if a == b == c == 0:
                    continue
                r2 = (a * a + b * b + c * c) * spacing * spacing
                if r2 <= h * h:
                    total += poly6_f64(r2, h)
    return float(np.float32(mass * (poly6_f64(0.0, h) + 0.5 * total)))


# ---------------------------------------------------------------- workloads
H = 0.1
SPACING = 0.05
RADIUS = 0.025


def dam_break(dims=(32, 32, 16), seed=1, grid=(64, 64, 64), box=6.4, jitter=0.001):
    """Plain PBF dam break: a jittered cubic lattice at spacing 0.5h (h=0.1) in
    a cubic box of `box` with one plane per face and `grid` cells of h.  The
    default is the 16,384-particle dam break."""
    sc = Scene(particle_radius=RADIUS, fluid_kernel_radius=H, cell_size=H, grid_dims=grid)
    start = (SPACING / 2,) * 3
    sc.add_fluid_block(dims, start, SPACING, 1.0, jitter=jitter, seed=seed)
    sc.add_box_planes((0, 0, 0), (box, box, box))
    return sc


def dam_break_params(**kw):
    """4 iterations, XSPH c = 0.01, no vorticity, no Akinci (sCorr on)."""
    import brasspan as bp
    p = dict(iterations=4, xsph_c=0.01, vorticity_eps=0.0, akinci=0,
             fluid_rest_density=lattice_rest_density(SPACING, H, 1.0))
    p.update(kw)
    return bp.default_params(**p)


def large_block(dims=(256, 128, 128), seed=3, grid=(256, 128, 128), box=(25.6, 12.8, 12.8)):
    """Plain large single scene: 4,194,304 particles at spacing 0.5h."""
    sc = Scene(particle_radius=RADIUS, fluid_kernel_radius=H, cell_size=H, grid_dims=grid)
    sc.add_fluid_block(dims, (SPACING / 2,) * 3, SPACING, 1.0, jitter=0.001, seed=seed)
    sc.add_box_planes((0, 0, 0), box)
    return sc


def large_block_params(**kw):
    """Akinci tension gamma 0.5, vorticity 1e-4, 4 iterations (1 substep)."""
    import brasspan as bp
    p = dict(iterations=4, akinci=1, tension_gamma=0.5, vorticity_eps=1e-4,
             fluid_rest_density=lattice_rest_density(SPACING, H, 1.0))
    p.update(kw)
    return bp.default_params(**p)


def mixed_phase(fluid_dims=(64, 64, 64), num_boxes=64, box_dim=4, cloth_n=128, seed=2,
                grid=(128, 128, 128), box=12.8):
    """Plain mixed-phase scene: rigid boxes first (shape matching), then a PBF
    fluid block, then a cloth flag pinned along one edge."""
    sc = Scene(particle_radius=RADIUS, fluid_kernel_radius=H, cell_size=H, grid_dims=grid)
    rng = Rng(seed)
    for _ in range(num_boxes):
        u = rng.uniform(7)
        start = (4.0 + 4.0 * u[0], 1.0 + 3.0 * u[1], 1.0 + 3.0 * u[2])
        axis = u[3:6] * 2.0 - 1.0
        axis = axis / max(np.linalg.norm(axis), 1e-6)
        ang = u[6] * 2.0 * math.pi
        q = (*(axis * math.sin(ang / 2)), math.cos(ang / 2))
        sc.add_rigid_box((box_dim,) * 3, start, SPACING, 1.0, rotation=q)
    sc.add_fluid_block(fluid_dims, (0.5, SPACING / 2, 0.5), SPACING, 1.0, jitter=0.001, rng=rng)
    # stiffness 0.2 / 0.1 as the reference's cloth preset (scene.h:207-208)
    cs, ce = sc.add_cloth((4.0, 5.0, 8.0), (SPACING, 0, 0), (0, SPACING, 0), cloth_n, cloth_n, 0.5,
                          0.2, 0.1, True)
    for y in range(cloth_n):  # pin the edge x = 0 of the flag
        sc.make_immovable(cs + y)
    sc.add_box_planes((0, 0, 0), (box, box, box))
    return sc


def mixed_phase_params(**kw):
    """Wind (5,0,1), friction 0.3 (static and dynamic), 4 iterations, 2
    substeps.  Akinci tension is off: the scene does not ask for it, and the
    reference's unweighted tension (fluid.cuh:436-489) diverges at spacing
    0.5h.  Every other parameter keeps the reference literal."""
    import brasspan as bp
    p = dict(iterations=4, friction_static=0.3, friction_dynamic=0.3, wind=(5.0, 0.0, 1.0),
             fluid_rest_density=lattice_rest_density(SPACING, H, 1.0), akinci=0)
    p.update(kw)
    return bp.default_params(**p)


def next_pow2(v):
    p = 4
    while p < v:
        p *= 2
    return p


def batch_member(rng: Rng, n_target: int):
    """One scene of the batched workload: a fluid block, 1-8 rigid 4x4x4 boxes
    and one square cloth (32..96 per side, at most half the particles) under a
    random horizontal wind of 0..8 m/s, with its own kernel radius h and a box
    and grid sized to it."""
    u = rng.uniform(12)
    h = (0.08, 0.1, 0.12)[min(2, int(u[0] * 3))]
    sp, rad = 0.5 * h, 0.25 * h
    cloth_n = 32 + int(u[1] * 65)
    while cloth_n > 32 and cloth_n * cloth_n > n_target // 2:
        cloth_n -= 1
    nboxes = 1 + min(7, int(u[2] * 8))
    fluid_n = max(512, n_target - cloth_n * cloth_n - 64 * nboxes)
    a = max(4, int(round(fluid_n ** (1.0 / 3.0))))
    b = max(1, fluid_n // (a * a))
    # box: fluid at the floor, boxes beside it, cloth above
    fluid_w = a * sp
    cloth_w = cloth_n * sp
    width = max(2.0 * fluid_w + 1.0, cloth_w + 2.0)
    height = b * sp + cloth_w + 2.0
    dims = (next_pow2(width / h), next_pow2(height / h), next_pow2(width / h))
    sc = Scene(particle_radius=rad, fluid_kernel_radius=h, cell_size=h, grid_dims=dims)
    for k in range(nboxes):
        v = rng.uniform(7)
        start = (fluid_w + 0.5 + v[0] * (width - fluid_w - 1.0), 0.2 + v[1] * 1.0, 0.3 + v[2] * (width - 0.8))
        axis = v[3:6] * 2.0 - 1.0
        axis = axis / max(np.linalg.norm(axis), 1e-6)
        ang = v[6] * 2.0 * math.pi
        sc.add_rigid_box((4, 4, 4), start, sp, 1.0, rotation=(*(axis * math.sin(ang / 2)), math.cos(ang / 2)))
    sc.add_fluid_block((a, b, a), (sp / 2, sp / 2, sp / 2), sp, 1.0, jitter=0.01 * h, rng=rng)
    cs, _ = sc.add_cloth((0.5, b * sp + 1.0, width / 2), (sp, 0, 0), (0, sp, 0), cloth_n, cloth_n, 0.5, 0.2,
                         0.1, True)
    for y in range(cloth_n):
        sc.make_immovable(cs + y)
    sc.add_box_planes((0, 0, 0), (width, height, width))
    wmag = 8.0 * u[3]
    wang = 2.0 * math.pi * u[4]
    import brasspan as bp
    prm = bp.default_params(iterations=4, fluid_rest_density=lattice_rest_density(sp, h, 1.0), akinci=0,
                            friction_static=0.3, friction_dynamic=0.3,
                            wind=(wmag * math.cos(wang), 0.0, wmag * math.sin(wang)))
    return sc, prm


def batched_scenes(num=512, seed=4, lo=2048, hi=65536):
    """`num` scenes with particle counts drawn log-uniform in [lo, hi]."""
    rng = Rng(seed)
    targets = np.exp(np.log(lo) + rng.uniform(num) * (np.log(hi) - np.log(lo))).astype(int)
    out = [batch_member(rng, int(t)) for t in targets]
    return [o[0] for o in out], [o[1] for o in out]


def adjoint_member(rng: Rng, fluid_dims=(32, 16, 32), cloth_n=64, box=6.4, grid=(64, 64, 64)):
    """One scene of the differentiable workload: a jittered PBF fluid block
    (32x16x32 = 16,384 by default) on the floor of a frictionless box and a
    cloth flag of cloth_n^2 particles (one solid phase, no self-contact)
    pinned along its left edge, high enough above the fluid that the two never
    touch, under a seeded wind of 2..6 m/s."""
    u = rng.uniform(3)
    sc = Scene(particle_radius=RADIUS, fluid_kernel_radius=H, cell_size=H, grid_dims=grid)
    # 1 cm clear of the floor and walls: a particle started exactly at contact
    # distance sits on the kink of the plane projection, where the rollout
    # has no derivative
    sc.add_fluid_block(fluid_dims, (SPACING / 2 + 0.01,) * 3, SPACING, 1.0, jitter=0.01 * H, rng=rng)
    top = fluid_dims[1] * SPACING + 0.01
    x0 = fluid_dims[0] * SPACING * 0.25
    y0 = top + 1.2
    cs, _ = sc.add_cloth((x0, y0, box * 0.6), (SPACING, 0, 0), (0, SPACING, 0), cloth_n, cloth_n, 0.5, 0.2, 0.1,
                         False)
    for y in range(cloth_n):
        sc.make_immovable(cs + y)
    sc.add_box_planes((0, 0, 0), (box, box, box))
    wmag = 2.0 + 4.0 * u[0]
    wang = 0.5 * math.pi * (u[1] - 0.5)
    import brasspan as bp
    prm = bp.default_params(iterations=4, fluid_rest_density=lattice_rest_density(SPACING, H, 1.0), akinci=0,
                            vorticity_eps=0.0, xsph_c=2e-5, friction_static=0.0, friction_dynamic=0.0,
                            wind=(wmag * math.cos(wang), 0.0, wmag * math.sin(wang)))
    return sc, prm


def adjoint_scenes(num=64, seed=5, **kw):
    """The differentiable workload: `num` scenes of adjoint_member."""
    rng = Rng(seed)
    out = [adjoint_member(rng, **kw) for _ in range(num)]
    return [o[0] for o in out], [o[1] for o in out]


ADJOINT_SUBSTEPS = 32
ADJOINT_SUBDT = 1.0 / 120.0
ADJOINT_EVERY = 8


WORKLOADS = {
    # name: (builder, params builder, dt, substeps per frame)
    "dam_break_16k": (dam_break, dam_break_params, 1.0 / 60.0, 2),
    "mixed_phase_283k": (mixed_phase, mixed_phase_params, 1.0 / 60.0, 2),
    "large_block_4m": (large_block, large_block_params, 1.0 / 60.0, 1),
}
