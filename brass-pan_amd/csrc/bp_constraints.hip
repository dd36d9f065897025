// bp_constraints.hip -- boundary, solid and cloth passes on gfx950: plane
// collisions, particle-particle contact, distance and dihedral bending
// constraints, aerodynamic wind, rigid shape matching and pins.
// Constraint groups keep the reference's Jacobi semantics (Δx from the same
// x*, then x* += Δx, particlesolver.cuh:887-969) in gather form: a compute
// kernel per constraint writes its corrections, then one thread per affected
// particle sums them in constraint-index order (CSR built on the host) and
// applies them.  No floating-point atomics.
#include "bp_kernels.h"

namespace bp {

// particlePlaneCollisionConstraint over all planes (particleplane.cuh:21-58,
// particlesolver.cuh:872-880), in place.
template <class SP>
__global__ __launch_bounds__(kBlock) void k_plane_collide(float4 *__restrict__ xs, const float4 *__restrict__ x,
                                                          int n, SP sp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sp.with_scene_ref(i, [&](const SceneConst &S) {
        const float4 a = xs[i];
        const f3 out = plane_collide_one(xyz(a), xyz(x[i]), S.pl, S.radius, S.friction_static, S.friction_dynamic);
        xs[i] = f4(out, a.w);
    });
}

// Cells whose 27-cell block holds a solid this substep (by the grid build's
// cells): cell_stamp[cell] == *stamp, so a particle learns whether its walk
// can meet a solid from one entry.  The relation is symmetric (a solid's cell
// lies in the block around c iff c lies in the block around it), so each
// solid marks the 27 cells around its own.  Cells holding a fluid particle:
// fluid_stamp[cell] == *stamp.  The stamp advances once per substep
// (k_stamp_next), so no table needs clearing.
__global__ void k_stamp_next(uint32_t *__restrict__ stamp) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *stamp = *stamp + 1u;
}
template <class SP>
__global__ __launch_bounds__(kBlock) void k_mark_cells(const int *__restrict__ sorted_keys,
                                                       const int *__restrict__ ph, int n,
                                                       const uint32_t *__restrict__ stamp,
                                                       uint32_t *__restrict__ cell_stamp,
                                                       uint32_t *__restrict__ fluid_stamp, SP sp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t st = *stamp;
    const int key = sorted_keys[i];
    const bool fluid = ph[i] <= 0;
    // the previous slot in the same cell and of the same kind (fluid or solid)
    // writes the same stamps: only the first of such a run writes them
    if (i > 0 && sorted_keys[i - 1] == key && (ph[i - 1] <= 0) == fluid) return;
    if (fluid) {
        fluid_stamp[key] = st;
        return;
    }
    sp.with_scene_ref(i, [&](const SceneConst &S) {
        const GridDesc &G = S.G;
        int g[3];
        unpack_cell(G.pack_addr(key), g);
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz) cell_stamp[G.addr(g[0] + dx, g[1] + dy, g[2] + dz)] = st;
    });
}

// particleParticleCollisionConstraint (particleparticle.cuh:5-101) + Δx apply.
// A fluid particle can only touch solids (the pair needs a positive phase),
// so it skips the walk when none of its 27 cells holds one (one cell_stamp
// entry, k_mark_cells).  A partner's x* and x are fetched together, ahead of
// the pair term.  (The staged walk, with both records in LDS, measured
// slower here: 0.323 -> 0.337 ms per launch on the 512-scene batch.)  Five
// waves per SIMD: the batched kernel fits 96 VGPRs without spilling (it took
// 101, four waves; batch 0.318 -> 0.299 ms per launch).
template <class SP>
__global__ __launch_bounds__(kMaxBlock) kPassWaves(SP, 5) void k_particle_particle(const float4 *__restrict__ xs,
                                                                 const float4 *__restrict__ x,
                                                                 const int2 *__restrict__ ranges,
                                                                 float4 *__restrict__ xs_out, int n, SP sp,
                                                                 int *__restrict__ contacts,
                                                                 const uint32_t *__restrict__ cell_stamp,
                                                                 const uint32_t *__restrict__ stamp, LambdaIn L,
                                                                 const float *__restrict__ m) {
    const int i = pass_slot();
    if (i >= n) return;
    sp.with_scene(i, [&](const SceneConst &S) {
        const GridDesc &G = S.G;
        const float r = S.radius, fs = S.friction_static, fd = S.friction_dynamic;
        const float4 a = xs[i];
        const f3 xip = xyz(a);
        const int pi = phase_of(a);
        f3 sumD = mk3(0.f, 0.f, 0.f), sumF = mk3(0.f, 0.f, 0.f);
        int cnt = 0;
        int g[3];
        G.pos(xip, g);
        if (cell_stamp && pi <= 0 && cell_stamp[G.addr(g[0], g[1], g[2])] != *stamp) {
            xs_out[i] = a;
            if (L.planes) store_lambda_in(L, i, xip, m[i], pi);
            return;
        }
        // (x, w) only for the particles that walk
        const float4 xa = x[i];
        const f3 xi = xyz(xa);
        const float wi = xa.w;
        const float r2x4 = r * r * 4.0f;
        const float thr = r * r * 0.001f * 0.001f;
        auto pick = [&](int, float4 b) {  // pi != pj also excludes i itself
            const int pj = phase_of(b);
            const float d2 = dist2(xip, xyz(b));
            return (pi != pj) & ((pi > 0) | (pj > 0)) & (d2 < r2x4) & (d2 > 1e-5f);
        };
        // partner records: b = (x*, phase), xb = (x, w)
        auto body = [&](int, const StagedRec<2> &R) {
            const float4 b = R.r[0], xb = R.r[1];
            const f3 xjp = xyz(b);
            const f3 diff = xip - xjp;
            const float dist = fsqrt(dist2(xip, xjp));
            const float idist = frcp(dist);
            const float wj = xb.w;
            const float iw = frcp(wi + wj);
            const float w1 = wi * iw;
            const float w2 = wj * iw;
            const f3 proj = diff * (2.0f * r * idist - 1.0f);
            const f3 dxi = w1 * proj;
            const f3 xistar = dxi + xip;
            sumD += dxi;
            const f3 dxj = -w2 * proj;
            const float l2 = length2(dxi);
            if (l2 > thr) {
                cnt += 1;
                const f3 xjstar = dxj + xjp;
                const f3 term1 = (xistar - xi) - (xjstar - xyz(xb));
                const f3 nn = diff * idist;
                const f3 tan = term1 - dot(term1, nn) * nn;
                const float tl2 = length2(tan);
                if (tl2 <= (fs * fs) * l2)
                    sumF -= w1 * tan;
                else
                    sumF -= w1 * tan * fminf(fd * fsqrt(l2 * frcp(tl2)), 1.0f);
            }
        };
        for_each_neighbor_nb<1>(
            G, ranges, g, xs, 0, pick, [&](int j) { return StagedRec<2>{{ld32(xs, j), ld32(x, j)}}; }, body,
            AllowAll{});
        const f3 d = (cnt == 0) ? sumD : sumD + sumF / (float)cnt;
        xs_out[i] = f4(xip + d, a.w);
        // the lambda pass, which follows, reads x* also as LambdaIn
        if (L.planes) store_lambda_in(L, i, xip + d, m[i], pi);
        // the adjoint treats this pass as the identity: it counts the particles
        // whose contact set is not empty
        if (contacts && (sumD.x != 0.0f || sumD.y != 0.0f || sumD.z != 0.0f || cnt > 0)) atomicAdd(contacts, 1);
    });
}

// distanceConstraints (distance.cuh:3-41): k of each pair
__global__ __launch_bounds__(kBlock) void k_distance_compute(const float4 *__restrict__ xs,
                                                             const float4 *__restrict__ x,
                                                             const int2 *__restrict__ pairs,
                                                             const float2 *__restrict__ params,
                                                             const int *__restrict__ map_to_new,
                                                             float4 *__restrict__ kbuf, int nc) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int2 pr = pairs[c];
    const float2 pa = params[c];
    const int id1 = map_to_new[pr.x], id2 = map_to_new[pr.y];
    const f3 x12 = xyz(xs[id1]) - xyz(xs[id2]);
    const float d = length(x12);
    f3 k = mk3(0.f, 0.f, 0.f);
    float valid = 0.0f;
    if (!(d <= 1e-5f)) {
        const f3 nn = x12 / d;
        const float w1 = x[id1].w, w2 = x[id2].w;
        const float sw = w1 + w2;
        k = nn * ((d - pa.x) / sw) * pa.y;
        valid = 1.0f;
    }
    kbuf[c] = f4(k, valid);
}
// Δx of each constrained particle: -w1 k on the first end, +w2 k on the second
__global__ __launch_bounds__(kBlock) void k_distance_apply(float4 *__restrict__ xs, const float4 *__restrict__ x,
                                                           const int *__restrict__ verts,
                                                           const int *__restrict__ offs,
                                                           const int *__restrict__ ents,
                                                           const float4 *__restrict__ kbuf,
                                                           const int *__restrict__ map_to_new, int nv) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;
    const int slot = map_to_new[verts[t]];
    const float w = x[slot].w;
    f3 dx = mk3(0.f, 0.f, 0.f);
    gather_entries(
        ents, offs[t], offs[t + 1], [&](int en) { return kbuf[en >> 1]; },
        [&](int en, float4 k4) {
            if (k4.w == 0.0f) return;
            const f3 k = xyz(k4);
            if (en & 1)
                dx += w * k;
            else
                dx += -w * k;
        });
    const float4 a = xs[slot];
    xs[slot] = f4(xyz(a) + dx, a.w);
}

// Double-precision 3-vectors for the dihedral hinge below.
struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator-(d3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ d3 operator/(d3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 dcross(d3 a, d3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ d3 dwide(float4 a) { return {(double)a.x, (double)a.y, (double)a.z}; }
__device__ __forceinline__ float4 dnarrow(d3 a) { return make_float4((float)a.x, (float)a.y, (float)a.z, 1.0f); }

// bendingConstraints (bending.cuh:3-76): the four corrections of each dihedral
// hinge, evaluated in FP64.  Near a flat hinge (d -> -1) both
// sqrt(1 - d^2)(acos d - pi) and sum w|Q|^2 vanish, and in FP32 their ratio
// is set by rounding (FP32 and FP64 evaluations of a rippled flag differ by
// 2e-3 per substep), so the reference's FP32 arithmetic does not define the
// result there; the host oracle evaluates the hinge in FP64 as well.  At most
// a few thousand hinges per scene: the FP64 rate is not a concern.
__global__ __launch_bounds__(kBlock) void k_bending_compute(const float4 *__restrict__ xs,
                                                            const float4 *__restrict__ x,
                                                            const int4 *__restrict__ bends,
                                                            const int *__restrict__ map_to_new,
                                                            float4 *__restrict__ bbuf, int nb, float kb) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nb) return;
    const int4 ids = bends[c];
    const int id1 = map_to_new[ids.x], id2 = map_to_new[ids.y], id3 = map_to_new[ids.z], id4 = map_to_new[ids.w];
    for (int a = 0; a < 4; ++a) bbuf[4 * c + a] = make_float4(0.f, 0.f, 0.f, 0.f);
    const d3 P1 = dwide(xs[id1]);
    const d3 P2 = dwide(xs[id2]) - P1, P3 = dwide(xs[id3]) - P1, P4 = dwide(xs[id4]) - P1;
    const d3 c23 = dcross(P2, P3), c24 = dcross(P2, P4);
    const double l23 = sqrt(ddot(c23, c23));
    if (l23 <= (double)1e-5f) return;
    const double l24 = sqrt(ddot(c24, c24));
    if (l24 <= (double)1e-5f) return;
    const d3 N1 = c23 * (1.0 / l23);
    const d3 N2 = c24 * (1.0 / l24);
    double d = ddot(N1, N2);
    d = fmin(fmax(d, -1.0), 1.0);
    const d3 c2n1 = dcross(P2, N1), c2n2 = dcross(P2, N2);
    const d3 c3n1 = dcross(P3, N1), c3n2 = dcross(P3, N2);
    const d3 c4n1 = dcross(P4, N1), c4n2 = dcross(P4, N2);
    const d3 Q3 = (c2n2 - c2n1 * d) / l23;
    const d3 Q4 = (c2n1 - c2n2 * d) / l24;
    const d3 Q2 = (-(c3n2 - c3n1 * d)) / l23 - (c4n1 - c4n2 * d) / l24;
    const d3 Q1 = -Q2 - Q3 - Q4;
    const double w1 = x[id1].w, w2 = x[id2].w, w3 = x[id3].w, w4 = x[id4].w;
    const double den = w1 * ddot(Q1, Q1) + w2 * ddot(Q2, Q2) + w3 * ddot(Q3, Q3) + w4 * ddot(Q4, Q4);
    const double num = sqrt(1.0 - d * d) * (acos(d) - (double)3.141592f);
    if (den <= (double)1e-5f) return;
    const double sc = num / den, kbd = (double)kb;
    bbuf[4 * c + 0] = dnarrow(Q1 * (-w1) * sc * kbd);
    bbuf[4 * c + 1] = dnarrow(Q2 * (-w2) * sc * kbd);
    bbuf[4 * c + 2] = dnarrow(Q3 * (-w3) * sc * kbd);
    bbuf[4 * c + 3] = dnarrow(Q4 * (-w4) * sc * kbd);
}
// sum of per-constraint corrections listed in a CSR, applied in place
__global__ __launch_bounds__(kBlock) void k_gather_apply(float4 *__restrict__ xs, const int *__restrict__ verts,
                                                         const int *__restrict__ offs, const int *__restrict__ ents,
                                                         const float4 *__restrict__ buf,
                                                         const int *__restrict__ map_to_new,
                                                         const float *__restrict__ m, int divide_by_mass, int nv) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;
    const int slot = map_to_new[verts[t]];
    const float mass = divide_by_mass ? m[slot] : 1.0f;
    f3 dx = mk3(0.f, 0.f, 0.f);
    gather_entries(
        ents, offs[t], offs[t + 1], [&](int en) { return buf[en]; },
        [&](int, float4 b) {
            if (b.w == 0.0f) return;
            if (divide_by_mass)
                dx += xyz(b) / mass;
            else
                dx += xyz(b);
        });
    const float4 a = xs[slot];
    xs[slot] = f4(xyz(a) + dx, a.w);
}

// applyWindForce (wind.cuh:3-50) with a wind vector: relative air velocity is
// the face-average particle velocity minus `wind` (0 reproduces the reference).
template <class SP>
__global__ __launch_bounds__(kBlock) void k_wind_compute(const float4 *__restrict__ xs, const float4 *__restrict__ v,
                                                         const int4 *__restrict__ faces,
                                                         const int *__restrict__ map_to_new,
                                                         float4 *__restrict__ fbuf, int nf, SP sp, float dt) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int4 ids = faces[f];
    const int id1 = map_to_new[ids.x], id2 = map_to_new[ids.y], id3 = map_to_new[ids.z];
    const SceneConst &S = sp.at(id1);
    const f3 wind = S.wind;
    const float cd = S.drag_coeff, cl = S.lift_coeff;
    fbuf[f] = make_float4(0.f, 0.f, 0.f, 0.f);
    const f3 p1 = xyz(xs[id1]), p2 = xyz(xs[id2]), p3 = xyz(xs[id3]);
    const f3 a = p2 - p1, b = p3 - p1;
    f3 vel = (xyz(v[id1]) + xyz(v[id2]) + xyz(v[id3])) / 3.0f;
    vel = vel - wind;
    const float lv = length(vel);
    if (lv <= 1e-5f) return;
    const f3 vhat = vel / lv;
    f3 fn = cross(a, b);
    const float lfn = length(fn);
    if (lfn <= 1e-5f) return;
    fn = fn / lfn;
    fn = (dot(fn, vel) > 0.0f) ? fn : -fn;
    const float area = 0.5f * lfn;
    const f3 drag = cd * (lv * lv) * area * dot(fn, vhat) * (-vhat);
    const f3 uhat = cross(cross(fn, vhat), vhat);
    const f3 lift = cl * (lv * lv) * area * dot(fn, vhat) * uhat;
    const f3 force = drag + lift;
    const f3 move = force * dt * dt;
    fbuf[f] = f4(move, 1.0f);
}

// shapeMatchingAlphaOne (shapematching.cuh:8-99): one 64-lane wave per body.
// Sums are taken in particle order (every lane walks the body in LDS), so all
// lanes hold the same centre of mass and Apq; the Müller 2016 rotation
// extraction (shapematching.cuh:70-88) runs uniformly on every lane.
// Blocks past the last body apply the pins (immovableConstraints) when the
// caller has found no pinned particle inside a rigid body (Pins.n > 0 only
// then): the two launches then touch disjoint particles.
struct Pins {
    const float4 *x;
    const int *ids;
    int n;
};
__device__ __forceinline__ void pin_one(float4 *xs, const float4 *x, const int *ids, const int *map_to_new, int t) {
    const int slot = map_to_new[ids[t]];
    const float4 a = xs[slot];
    const float4 b = x[slot];
    xs[slot] = make_float4(b.x, b.y, b.z, a.w);
}

__global__ __launch_bounds__(64) void k_shape_matching(float4 *__restrict__ xs, const int *__restrict__ map_to_new,
                                                       const float4 *__restrict__ rest,
                                                       const int2 *__restrict__ ranges, float4 *__restrict__ rot,
                                                       float4 *__restrict__ cms, int iters, int nr, Pins pins) {
    __shared__ float sp[kMaxRigidParticles][3];
    __shared__ float sq[kMaxRigidParticles][3];
    const int b = blockIdx.x, t = threadIdx.x;
    if (b >= nr) {
        const int i = (b - nr) * 64 + t;
        if (i < pins.n) pin_one(xs, pins.x, pins.ids, map_to_new, i);
        return;
    }
    const int2 rg = ranges[b];
    const int np = rg.y - rg.x;
    int slot = 0;
    f3 q = mk3(0.f, 0.f, 0.f);
    float pw = 0.f;
    if (t < np) {
        slot = map_to_new[rg.x + t];
        const float4 a = xs[slot];
        pw = a.w;
        sp[t][0] = a.x; sp[t][1] = a.y; sp[t][2] = a.z;
        const float4 r4 = rest[rg.x + t];
        q = xyz(r4);
        sq[t][0] = r4.x; sq[t][1] = r4.y; sq[t][2] = r4.z;
    }
    __syncthreads();
    f3 sum = mk3(0.f, 0.f, 0.f);
    for (int u = 0; u < np; ++u) sum += mk3(sp[u][0], sp[u][1], sp[u][2]);
    const f3 cm = sum / (float)np;
    f3 A0 = mk3(0.f, 0.f, 0.f), A1 = A0, A2 = A0;
    for (int u = 0; u < np; ++u) {
        const f3 p = mk3(sp[u][0], sp[u][1], sp[u][2]) - cm;
        A0 += p * sq[u][0];
        A1 += p * sq[u][1];
        A2 += p * sq[u][2];
    }
    float4 qq = rot[b];
    for (int it = 0; it < iters; ++it) {
        const float x2 = qq.x * qq.x, y2 = qq.y * qq.y, z2 = qq.z * qq.z;
        const f3 R0 = mk3(1.0f - 2.0f * y2 - 2.0f * z2, 2.0f * qq.x * qq.y + 2.0f * qq.w * qq.z,
                          2.0f * qq.x * qq.z - 2.0f * qq.w * qq.y);
        const f3 R1 = mk3(2.0f * qq.x * qq.y - 2.0f * qq.w * qq.z, 1.0f - 2.0f * x2 - 2.0f * z2,
                          2.0f * qq.y * qq.z + 2.0f * qq.w * qq.x);
        const f3 R2 = mk3(2.0f * qq.x * qq.z + 2.0f * qq.w * qq.y, 2.0f * qq.y * qq.z - 2.0f * qq.w * qq.x,
                          1.0f - 2.0f * x2 - 2.0f * y2);
        const f3 num = cross(R0, A0) + cross(R1, A1) + cross(R2, A2);
        const float den = 1.0f / fabsf(dot(R0, A0) + dot(R1, A1) + dot(R2, A2)) + 1e-9f;
        const f3 om = num * den;
        const float w2 = length2(om);
        if (w2 <= 1e-9f) break;
        const float w = sqrtf(w2);
        const f3 dir = om / w;
        const float sh = sinf(w * 0.5f), ch = cosf(w * 0.5f);
        const float4 aa = make_float4(dir.x * sh, dir.y * sh, dir.z * sh, ch);
        float4 r;  // mul(aa, qq) (cudamatrix.cuh:112-121)
        r.x = aa.w * qq.x + aa.x * qq.w + aa.y * qq.z - aa.z * qq.y;
        r.y = aa.w * qq.y - aa.x * qq.z + aa.y * qq.w + aa.z * qq.x;
        r.z = aa.w * qq.z + aa.x * qq.y - aa.y * qq.x + aa.z * qq.w;
        r.w = aa.w * qq.w - aa.x * qq.x - aa.y * qq.y - aa.z * qq.z;
        const float inv = 1.0f / sqrtf(r.x * r.x + r.y * r.y + r.z * r.z + r.w * r.w);
        qq = make_float4(r.x * inv, r.y * inv, r.z * inv, r.w * inv);
    }
    const float x2 = qq.x * qq.x, y2 = qq.y * qq.y, z2 = qq.z * qq.z;
    const f3 R0 = mk3(1.0f - 2.0f * y2 - 2.0f * z2, 2.0f * qq.x * qq.y + 2.0f * qq.w * qq.z,
                      2.0f * qq.x * qq.z - 2.0f * qq.w * qq.y);
    const f3 R1 = mk3(2.0f * qq.x * qq.y - 2.0f * qq.w * qq.z, 1.0f - 2.0f * x2 - 2.0f * z2,
                      2.0f * qq.y * qq.z + 2.0f * qq.w * qq.x);
    const f3 R2 = mk3(2.0f * qq.x * qq.z + 2.0f * qq.w * qq.y, 2.0f * qq.y * qq.z - 2.0f * qq.w * qq.x,
                      1.0f - 2.0f * x2 - 2.0f * y2);
    if (t == 0) {
        rot[b] = qq;
        cms[b] = f4(cm, 0.0f);
    }
    if (t < np) {
        f3 o = R0 * q.x;  // matrix3 * float3 (cudamatrix.cuh:87-95)
        o += R1 * q.y;
        o += R2 * q.z;
        xs[slot] = f4(o + cm, pw);
    }
}

// immovableConstraints (immovable.cuh:3-14)
__global__ __launch_bounds__(kBlock) void k_immovable(float4 *__restrict__ xs, const float4 *__restrict__ x,
                                                      const int *__restrict__ ids,
                                                      const int *__restrict__ map_to_new, int ni) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni) return;
    pin_one(xs, x, ids, map_to_new, t);
}

// setParticle (particlesolver.cuh:388-393) on the picked particle, found by
// original id through the inverse map (queryNewParticleId, :360-366).
__global__ void k_pick(float4 *__restrict__ pos, float4 *__restrict__ v, const int *__restrict__ map_to_new,
                       int orig, f3 p, int keep_pos, f3 vel) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int slot = map_to_new[orig];
    const float4 a = pos[slot];
    if (!keep_pos) pos[slot] = make_float4(p.x, p.y, p.z, a.w);
    v[slot] = make_float4(vel.x, vel.y, vel.z, 0.0f);
}

// updateVelocity + updatePositions (particlesolver.cuh:71-81,121-137)
__global__ __launch_bounds__(kBlock) void k_finish(const float4 *__restrict__ xs, float4 *__restrict__ x,
                                                   float4 *__restrict__ v, const float *__restrict__ vw, int n,
                                                   float inv_dt, float eps2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = xs[i];
    const float4 x4 = x[i];
    const f3 out = xyz(a), xp = xyz(x4);
    v[i] = f4((out - xp) * inv_dt, (vw && phase_of(a) < 0) ? vw[i] : 0.0f);  // v.w as in the position epilogue
    const float d2 = length2(out - xp);
    if (d2 >= eps2 || phase_of(a) < 0) x[i] = f4(out, x4.w);
}

// ------------------------------------------------------------ launchers
template <class SP>
void launch_plane_collide(float4 *xs, const float4 *x, int n, const SP &sp, hipStream_t st) {
    hipLaunchKernelGGL(k_plane_collide<SP>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, xs, x, n, sp);
}
template <class SP>
void launch_particle_particle(const float4 *xs, const float4 *x, const int2 *ranges, float4 *xs_out, int n,
                              const SP &sp, int block, int *contacts, const uint32_t *cell_stamp,
                              const uint32_t *stamp, const LambdaIn &L, const float *m, hipStream_t st) {
    hipLaunchKernelGGL(k_particle_particle<SP>, dim3(grid_blocks(n, block)), dim3(block), 0, st, xs, x, ranges,
                       xs_out, n, sp, contacts, cell_stamp, stamp, L, m);
}
template <class SP>
void launch_mark_cells(const int *sorted_keys, const int *ph, int n, uint32_t *stamp, uint32_t *cell_stamp,
                       uint32_t *fluid_stamp, const SP &sp, hipStream_t st) {
    hipLaunchKernelGGL(k_stamp_next, dim3(1), dim3(64), 0, st, stamp);
    hipLaunchKernelGGL(k_mark_cells<SP>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, sorted_keys, ph, n, stamp,
                       cell_stamp, fluid_stamp, sp);
}
void launch_distance(float4 *xs, const float4 *x, const int2 *pairs, const float2 *params, const int *map_to_new,
                     float4 *kbuf, int nc, const int *verts, const int *offs, const int *ents, int nv,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_distance_compute, dim3(grid_blocks(nc)), dim3(kBlock), 0, st, xs, x, pairs, params,
                       map_to_new, kbuf, nc);
    hipLaunchKernelGGL(k_distance_apply, dim3(grid_blocks(nv)), dim3(kBlock), 0, st, xs, x, verts, offs, ents, kbuf,
                       map_to_new, nv);
}
void launch_bending(float4 *xs, const float4 *x, const int4 *bends, const int *map_to_new, float4 *bbuf, int nb,
                    float kb, const int *verts, const int *offs, const int *ents, int nv, hipStream_t st) {
    hipLaunchKernelGGL(k_bending_compute, dim3(grid_blocks(nb)), dim3(kBlock), 0, st, xs, x, bends, map_to_new,
                       bbuf, nb, kb);
    hipLaunchKernelGGL(k_gather_apply, dim3(grid_blocks(nv)), dim3(kBlock), 0, st, xs, verts, offs, ents, bbuf,
                       map_to_new, (const float *)nullptr, 0, nv);
}
template <class SP>
void launch_wind(float4 *xs, const float4 *v, const float *m, const int4 *faces, const int *map_to_new,
                 float4 *fbuf, int nf, const SP &sp, float dt, const int *verts, const int *offs, const int *ents,
                 int nv, hipStream_t st) {
    hipLaunchKernelGGL(k_wind_compute<SP>, dim3(grid_blocks(nf)), dim3(kBlock), 0, st, xs, v, faces, map_to_new,
                       fbuf, nf, sp, dt);
    hipLaunchKernelGGL(k_gather_apply, dim3(grid_blocks(nv)), dim3(kBlock), 0, st, xs, verts, offs, ents, fbuf,
                       map_to_new, m, 1, nv);
}
void launch_shape_matching(float4 *xs, const int *map_to_new, const float4 *rest, const int2 *ranges, float4 *rot,
                           float4 *cms, int nr, int iters, hipStream_t st, const float4 *pin_x, const int *pin_ids,
                           int npins) {
    const Pins pins{pin_x, pin_ids, npins};
    hipLaunchKernelGGL(k_shape_matching, dim3(nr + (npins + 63) / 64), dim3(64), 0, st, xs, map_to_new, rest, ranges,
                       rot, cms, iters, nr, pins);
}
void launch_immovable(float4 *xs, const float4 *x, const int *ids, const int *map_to_new, int ni, hipStream_t st) {
    hipLaunchKernelGGL(k_immovable, dim3(grid_blocks(ni)), dim3(kBlock), 0, st, xs, x, ids, map_to_new, ni);
}
void launch_pick(float4 *pos, float4 *v, const int *map_to_new, int orig, f3 p, int keep_pos, f3 vel,
                 hipStream_t st) {
    hipLaunchKernelGGL(k_pick, dim3(1), dim3(64), 0, st, pos, v, map_to_new, orig, p, keep_pos, vel);
}
void launch_finish(const float4 *xs, float4 *x, float4 *v, const float *vw, int n, float inv_dt, float eps2,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_finish, dim3(grid_blocks(n)), dim3(kBlock), 0, st, xs, x, v, vw, n, inv_dt, eps2);
}

#define BP_INSTANTIATE_CONSTRAINTS(SP)                                                                          \
    template void launch_plane_collide<SP>(float4 *, const float4 *, int, const SP &, hipStream_t);              \
    template void launch_mark_cells<SP>(const int *, const int *, int, uint32_t *, uint32_t *, uint32_t *,       \
                                        const SP &, hipStream_t);                                                \
    template void launch_particle_particle<SP>(const float4 *, const float4 *, const int2 *, float4 *, int,      \
                                               const SP &, int, int *, const uint32_t *, const uint32_t *,       \
                                               const LambdaIn &, const float *, hipStream_t);                    \
    template void launch_wind<SP>(float4 *, const float4 *, const float *, const int4 *, const int *, float4 *,  \
                                  int, const SP &, float, const int *, const int *, const int *, int, hipStream_t);
BP_INSTANTIATE_CONSTRAINTS(OneScene)
BP_INSTANTIATE_CONSTRAINTS(SceneTable)

}  // namespace bp
