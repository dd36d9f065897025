// bp_internal.h -- shared device helpers and launch interfaces of the gfx950
// particle solver.  Everything here is HIP C++ for gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bp {

constexpr int kMaxPlanes = 16;
constexpr int kMaxRigidParticles = 64;  // NUM_MAX_PARTICLE_PER_RIGID_BODY (cubrasstower.cuh:22)
constexpr int kBlock = 256;      // streaming kernels
constexpr int kMaxBlock = 512;   // neighbour passes: block size chosen at run time (autotuner)
// Minimum waves per SIMD a neighbour pass asks the compiler for (the smaller
// of SP::kWaves and the pass's own W; 6 = at most 80 VGPRs, 5 = 96).  The
// passes wait on memory, so occupancy pays where the walk fits the budget
// with few or no spills.  1 leaves the register budget open.
#define kPassWaves(SP, W) __attribute__((amdgpu_waves_per_eu((SP::kWaves < (W)) ? SP::kWaves : (W))))

// ------------------------------------------------------------ small vectors
// Own 3-vector type: component-wise operators with the left-to-right order of
// the reference's float3 operators (ext/helper_math.h), so FP32 results match
// the host oracle bit-for-bit wherever the same expression is evaluated.
struct f3 {
    float x, y, z;
};
__host__ __device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__host__ __device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__host__ __device__ __forceinline__ f3 &operator+=(f3 &a, f3 b) { a = a + b; return a; }
__host__ __device__ __forceinline__ f3 &operator-=(f3 &a, f3 b) { a = a - b; return a; }
__host__ __device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ float length2(f3 a) { return dot(a, a); }
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
__host__ __device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 xyz(float4 a) { return f3{a.x, a.y, a.z}; }

// Pair-term arithmetic.  Squared distances use FMA; square roots and
// reciprocals in per-pair terms use the 1-ulp hardware instructions
// (v_sqrt_f32 / v_rcp_f32).  The grid hash keeps IEEE division (GridDesc::pos).
__device__ __forceinline__ float dist2(f3 a, f3 b) {
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
// x^n for a small non-negative integer n (sCorr exponent, fluid.cuh:208)
__device__ __forceinline__ float ipowf(float x, int n) {
    if (n == 4) {  // the loop's own product for n = 4, without the loop
        const float x2 = x * x;
        return x2 * x2;
    }
    float r = 1.0f;
    while (n > 0) {
        if (n & 1) r *= x;
        x *= x;
        n >>= 1;
    }
    return r;
}
__device__ __forceinline__ float4 f4(f3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
// p[j] addressed by a 32-bit byte offset from the (uniform) base, so the load
// takes the scalar-base form with no 64-bit address arithmetic per lane.
// Slot arrays hold at most kMaxSlots records (checked at create).
constexpr int kMaxSlots = 1 << 27;
template <class T> __device__ __forceinline__ T ld32(const T *__restrict__ p, int j) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + (uint32_t)j * (uint32_t)sizeof(T));
}

// ------------------------------------------------------------ parameters
struct Planes {
    int n;
    float o[kMaxPlanes][3];
    float nrm[kMaxPlanes][3];
};

// Uniform hash grid (grid.cuh:6-28), per-axis power-of-two dims.
struct GridDesc {
    float cs[3];
    int dims[3];
    uint32_t mask[3];
    uint32_t stride_x;  // dims[1]*dims[2]
    uint32_t shift_y;   // log2(dims[2])
    uint32_t shift_x;   // log2(dims[1]*dims[2])
    uint32_t base;      // first cell id of this scene (scenes of a batch own disjoint cell ranges)
    // cell id -> packed wrapped coordinates
    __device__ __forceinline__ uint32_t pack_addr(int a) const {
        const uint32_t u = (uint32_t)a - base;
        return ((u >> shift_x) << 20) | (((u >> shift_y) & mask[1]) << 10) | (u & mask[2]);
    }
    __device__ __forceinline__ void pos(f3 p, int g[3]) const {
        g[0] = (int)((p.x + 50000.0f) / cs[0]);
        g[1] = (int)((p.y + 50000.0f) / cs[1]);
        g[2] = (int)((p.z + 50000.0f) / cs[2]);
    }
    __device__ __forceinline__ int addr(int x, int y, int z) const {
        return (int)(base + ((uint32_t)x & mask[0]) * stride_x + ((uint32_t)y & mask[1]) * (uint32_t)dims[2] +
                     ((uint32_t)z & mask[2]));
    }
    // wrapped coordinates packed 10:10:10 (dims <= 1024 per axis)
    __device__ __forceinline__ uint32_t pack(const int g[3]) const {
        return (((uint32_t)g[0] & mask[0]) << 20) | (((uint32_t)g[1] & mask[1]) << 10) | ((uint32_t)g[2] & mask[2]);
    }
    // true iff packed cell `c` lies in the 27-cell block around packed cell `g`
    __device__ __forceinline__ bool in_block(uint32_t g, uint32_t c) const {
        uint32_t dx = (((c >> 20) & 1023u) - ((g >> 20) & 1023u) + 1u) & mask[0];
        uint32_t dy = (((c >> 10) & 1023u) - ((g >> 10) & 1023u) + 1u) & mask[1];
        uint32_t dz = ((c & 1023u) - (g & 1023u) + 1u) & mask[2];
        return dx <= 2u && dy <= 2u && dz <= 2u;
    }
};

// SPH kernel constants, SetKernelRadius (fluid.cuh:11-27), computed on the host.
struct SphConst {
    float c1, c2, c3, c4, h, h2, halfh, wdq;
    float inv_wdq;  // 1/wdq rounded to float, or 0 where over_wdq must divide
    __device__ __forceinline__ float poly6(float r2) const {
        float t = h2 - r2;
        return c1 * t * t * t;
    }
    // a / wdq (the sCorr ratio, fluid.cuh:208) for a = poly6(r^2) >= 0.  The
    // product with the rounded reciprocal, corrected by one FMA step
    // (Markstein), is the correctly rounded quotient when the residual does
    // not underflow: it equals IEEE division for every float a in
    // [2^-100, W(0)] at kernel radii 0.033 to 1 (a host check over all such
    // floats); the first mismatches lie below 2^-104.  Smaller a, and a wdq
    // outside [2^-20, 2^20] (inv_wdq = 0), take the division.
    __device__ __forceinline__ float over_wdq(float a) const {
        if ((inv_wdq != 0.0f) & (a >= 0x1p-100f)) {
            const float q0 = a * inv_wdq;
            return fmaf(fmaf(-q0, wdq, a), inv_wdq, q0);
        }
        return a / wdq;
    }
    __device__ __forceinline__ f3 grad(f3 v, float r2) const {
        float r = sqrtf(r2);
        float t = h - r;
        float s = c2 * t * t;
        return (s * v) / r;
    }
    // gradientSpikyKernel(v, r^2) = grad_coef(r, 1/r) * v
    __device__ __forceinline__ float grad_coef(float r, float ir) const {
        const float t = h - r;
        return c2 * t * t * ir;
    }
    __device__ __forceinline__ float akinci(float r) const {
        float t = (h - r) * r;
        float t3 = t * t * t;
        if (r >= halfh) return c3 * t3;
        return 2.0f * c3 * t3 - c4;
    }
};

struct FluidParams {
    float rest_density, inv_rest_density, solid_density_scale;
    int lambda_mode;
    float lambda_denominator, lambda_relaxation;
    int akinci;
    float scorr_k, delta_clamp;
    int scorr_n;
    float vorticity_eps, tension_gamma, xsph_c;
};

// Epilogue fused into the last full-particle Jacobi pass of an iteration.
// What the lambda pass reads besides x*, rewritten whenever x* changes before
// it: x* as three planes x[], y[], z[] (`pstride` floats apart, kWalkPad slots
// of slack) for its distance screen, and (x*, m) records for its pair bodies,
// with the sign of m set for a non-fluid particle (phase >= 0).
struct LambdaIn {
    float *planes = nullptr;
    float4 *xm = nullptr;
    int pstride = 0;
};

struct Epilogue {
    int planes_next;      // apply the next iteration's plane collisions
    int finish;           // last iteration: updateVelocity + updatePositions
    float inv_dt;
    float sleep_eps2;
    // with finish: v.w = vw[k] for fluid particles (the density, read by the
    // omega/normal pass; the vorticity/tension pass sets it back to 0), else 0
    const float *vw = nullptr;
    // x* out also as the lambda pass's inputs (LambdaIn) when the next lambda
    // pass reads this x* (nothing else moves it in between), else null
    LambdaIn lin;
    const float *m = nullptr;  // masses, with lin
};

// Everything a kernel needs to know about the scene a particle belongs to.
struct SceneConst {
    GridDesc G;
    SphConst K;
    FluidParams F;
    float radius, friction_static, friction_dynamic;
    f3 wind;
    float drag_coeff, lift_coeff;
    Planes pl;
};

// Parameter sources: one scene by value (plain solver), or a per-particle
// lookup in a scene table (batched solver: scenes concatenated slot-wise).
// with_scene(slot, f) runs f(scene of slot); in a scene table it runs once per
// distinct scene among the wave's active lanes (scenes are contiguous slot
// ranges, so nearly always once), with that scene's id made wave-uniform so
// its constants are scalar loads rather than one load per lane.  f gets a
// local copy of the entry: for passes that read scalar fields only.  A pass
// that indexes the plane arrays with a loop variable keeps at(slot), since
// the copy would then live in scratch memory.
struct OneScene {
    static constexpr int kWaves = 6;
    SceneConst c;
    __device__ __forceinline__ const SceneConst &at(int) const { return c; }
    template <class F> __device__ __forceinline__ void with_scene(int, F &&f) const { f(c); }
    template <class F> __device__ __forceinline__ void with_scene_ref(int, F &&f) const { f(c); }
};
struct SceneTable {
    // 96 VGPRs: the omega/normal pass spills 2 and runs 10% faster than
    // uncapped at 98 (4 waves); lambda and XSPH fit without spills
    static constexpr int kWaves = 5;
    const SceneConst *__restrict__ table;
    const int *__restrict__ sid;  // slot -> scene
    __device__ __forceinline__ const SceneConst &at(int slot) const { return table[sid[slot]]; }
    template <class F> __device__ __forceinline__ void with_scene(int slot, F &&f) const {
        const int s = sid[slot];
        // a wave-uniform loop over the distinct scenes of the active lanes
        uint64_t todo = __builtin_amdgcn_ballot_w64(true);
        while (todo) {
            const int s0 = __builtin_amdgcn_readlane(s, __builtin_ctzll(todo));
            // an opaque copy of s0 in a scalar register: inside the branch
            // below the compiler would otherwise substitute the per-lane s
            int u = s0;
            asm volatile("" : "+s"(u));
            const bool mine = s == s0;
            if (mine) {
                // the table is read-only for the whole launch: read the entry
                // through the constant address space, as scalar loads
                typedef const __attribute__((address_space(4))) uint32_t cword;
                static_assert(sizeof(SceneConst) % 4 == 0, "scene records are whole words");
                const cword *src = (const cword *)(table + u);
                SceneConst sc;
                uint32_t *dst = reinterpret_cast<uint32_t *>(&sc);
#pragma unroll
                for (int w = 0; w < (int)(sizeof(SceneConst) / 4); ++w) dst[w] = src[w];
                f(sc);
            }
            todo &= ~__builtin_amdgcn_ballot_w64(mine);
        }
    }
    // As with_scene, but f gets a reference to the wave-uniform table entry
    // (read through the constant address space) instead of a local copy: for
    // a pass that reads many fields, such as the plane arrays, whose copy
    // would be placed in scratch memory.
    template <class F> __device__ __forceinline__ void with_scene_ref(int slot, F &&f) const {
        const int s = sid[slot];
        uint64_t todo = __builtin_amdgcn_ballot_w64(true);
        while (todo) {
            const int s0 = __builtin_amdgcn_readlane(s, __builtin_ctzll(todo));
            int u = s0;
            asm volatile("" : "+s"(u));
            const bool mine = s == s0;
            if (mine) {
                typedef const __attribute__((address_space(4))) SceneConst cscene;
                f(*(const SceneConst *)((const cscene *)table + u));
            }
            todo &= ~__builtin_amdgcn_ballot_w64(mine);
        }
    }
};

// ------------------------------------------------------------ device helpers
__device__ __forceinline__ int phase_of(float4 a) { return __float_as_int(a.w); }
// packed 10:10:10 cell (GridDesc::pack) -> grid position
__device__ __forceinline__ void unpack_cell(uint32_t c, int g[3]) {
    g[0] = (int)(c >> 20);
    g[1] = (int)((c >> 10) & 1023u);
    g[2] = (int)(c & 1023u);
}

// particlePlaneCollisionConstraint (particleplane.cuh:21-58) for one particle.
// The loop is unrolled over the kMaxPlanes slots, so a wave-uniform copy of
// the planes (with_scene) is indexed by constants and stays in scalar
// registers; the planes are applied in order, as before.
__device__ __forceinline__ f3 plane_collide_one(f3 xs, f3 pos, const Planes &pl, float r, float fs, float fd) {
#pragma unroll
    for (int p = 0; p < kMaxPlanes; ++p) {
        if (p >= pl.n) break;
        f3 o = mk3(pl.o[p][0], pl.o[p][1], pl.o[p][2]);
        f3 n = mk3(pl.nrm[p][0], pl.nrm[p][1], pl.nrm[p][2]);
        f3 o2p = o - xs;
        float d = dot(o2p, n) + r;
        if (d <= 0.0f) continue;
        f3 diff = xs - pos;
        float dn = dot(diff, n);
        f3 dtan = diff - dn * n;
        float dtl = length(dtan);
        f3 resolved = d * n + xs;
        f3 dX = resolved - pos;
        f3 tdx = dX - dot(dX, n) * n;
        if (dtl < fs * dn)
            xs = resolved - tdx;
        else
            xs = resolved - tdx * fminf(fd * -dn / dtl, 1.0f);
    }
    return xs;
}

// four ints at 8-byte alignment (two adjacent cell ranges in one load)
typedef int int4_a8 __attribute__((ext_vector_type(4), aligned(8)));

// Iterates the reference's 27-cell candidate set (fluid.cuh:104-115) around
// grid position g: the 9 (x,y) columns in order, each visiting the cells
// z-1, z, z+1 (wrapped) in that order.  Their particles are consecutive slots
// except across the wrap, so a column is one run of slots, or two when it
// wraps.  The next column's three cell ranges are loaded while the current
// one is walked.  (Grid dimensions are at least 4, so the cells differ.)
template <class F>
__device__ __forceinline__ void for_each_run(const GridDesc &G, const int2 *__restrict__ ranges, const int g[3],
                                             F &&f) {
    const uint32_t zc[3] = {(uint32_t)(g[2] - 1) & G.mask[2], (uint32_t)g[2] & G.mask[2],
                            (uint32_t)(g[2] + 1) & G.mask[2]};
    const bool brk01 = zc[1] != zc[0] + 1u, brk12 = zc[2] != zc[1] + 1u;  // the wrap point, if any
    auto column = [&](int ci) {
        const int dy = ci / 3 - 1, dx = ci % 3 - 1;
        return G.base + ((uint32_t)(g[0] + dx) & G.mask[0]) * G.stride_x +
               ((uint32_t)(g[1] + dy) & G.mask[1]) * (uint32_t)G.dims[2];
    };
    // a column's three ranges: two loads where its cells are adjacent
    const bool adjacent = !brk01 && !brk12;
    auto load_column = [&](uint32_t col, int2 r[3]) {
        if (adjacent) {
            const int4_a8 a = *reinterpret_cast<const int4_a8 *>(ranges + col + zc[0]);
            r[0] = make_int2(a.x, a.y);
            r[1] = make_int2(a.z, a.w);
            r[2] = ranges[col + zc[2]];
        } else {
#pragma unroll
            for (int q = 0; q < 3; ++q) r[q] = ranges[col + zc[q]];
        }
    };
    int2 nr[3];
    load_column(column(0), nr);
    for (int ci = 0; ci < 9; ++ci) {
        int2 r[3] = {nr[0], nr[1], nr[2]};
        if (ci + 1 < 9) load_column(column(ci + 1), nr);
        int s = -1, e = -1;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if ((q == 1 && brk01) || (q == 2 && brk12)) {
                if (s >= 0) f(s, e);
                s = -1;
            }
            if (r[q].x >= 0) {
                if (s < 0) s = r[q].x;
                e = r[q].y;
            }
        }
        if (s >= 0) f(s, e);
    }
}

// Index rules of a walk, as masks over a chunk of 32 consecutive slots (bit t
// stands for slot base + t), so the per-candidate screen holds no index
// comparisons.
struct AllowAll {
    __device__ __forceinline__ uint32_t operator()(int) const { return ~0u; }
};
__device__ __forceinline__ uint32_t chunk_bit(int k, int base) {  // bit of slot k, 0 if outside
    const uint32_t d = (uint32_t)(k - base);
    return d < 32u ? (1u << d) : 0u;
}
__device__ __forceinline__ uint32_t chunk_below(int k, int base) {  // bits of slots < k
    const int d = k - base;
    return d <= 0 ? 0u : (d >= 32 ? ~0u : (1u << d) - 1u);
}
struct AllowNotSelf {
    int k;
    __device__ __forceinline__ uint32_t operator()(int base) const { return ~chunk_bit(k, base); }
};

// Screen record: the candidate masks one walk computed, replayed by a later
// walk over the same positions, cell and ranges instead of screening again
// (lambda -> position of one iteration; omega/normal -> vorticity/tension and
// XSPH).  Chunk c (in walk order) of slot `slot` is m[c * n + slot], for the
// first kRecChunks chunks; later chunks are screened by both walks.
constexpr int kRecChunks = 16;
struct ScreenRec {
    uint32_t *out = nullptr;       // record this walk's masks (every chunk is screened)
    const uint32_t *in = nullptr;  // replay recorded masks
    int slot = 0, n = 0;
    // Optional copy of the positions as three planes x[], y[], z[], `pstride`
    // floats apart, for a screen that tests positions only: 8 candidates are
    // then 6 loads of 4 floats rather than 8 records (pick sees w = 0).
    const float *planes = nullptr;
    int pstride = 0;
};

// True iff one of the 27 cells around g carries stamp s in `table` (a cell
// table filled once per substep from the grid build's keys).
__device__ __forceinline__ bool stamp_near(const GridDesc &G, const int g[3], const uint32_t *__restrict__ table,
                                           uint32_t s) {
    bool any = false;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const uint32_t col = G.base + ((uint32_t)(g[0] + dx) & G.mask[0]) * G.stride_x +
                                 ((uint32_t)(g[1] + dy) & G.mask[1]) * (uint32_t)G.dims[2];
#pragma unroll
            for (int dz = -1; dz <= 1; ++dz) any |= table[col + ((uint32_t)(g[2] + dz) & G.mask[2])] == s;
        }
    return any;
}

// Two-phase neighbour walk: for every run, chunks of up to 32 candidates are
// first screened with the cheap test `pick(j, b)` (b = xs[j]) into a bitmask,
// which `allow(base)` then restricts by slot index; `consume(base, mask)`
// receives each chunk's partners.  With rec.out every chunk is screened (even
// one `allow` empties) and the first kRecChunks masks are stored before
// `allow`; with rec.in those chunks take their masks from the record instead.
//
// Candidates are loaded in batches of 8 from one base address with immediate
// offsets.  A batch may read up to 7 slots past the end of its run, so every
// x* buffer (and plane) a walk reads holds kWalkPad slots beyond the last
// particle; the bits of those slots are cleared once per chunk.
constexpr int kWalkPad = 8;
// four floats at 4-byte alignment (one 16-byte load from any slot of a plane)
typedef float float4_a4 __attribute__((ext_vector_type(4), aligned(4)));
// Stores slot i of the lambda pass's inputs (x* at p, mass m, phase ph): the
// planes if L has them, the (x*, m) record if L has those.
__device__ __forceinline__ void store_lambda_in(const LambdaIn &L, int i, f3 p, float m, int ph) {
    if (L.planes) {
        L.planes[i] = p.x;
        L.planes[L.pstride + i] = p.y;
        L.planes[2 * (size_t)L.pstride + i] = p.z;
    }
    if (L.xm) L.xm[i] = make_float4(p.x, p.y, p.z, ph < 0 ? m : -m);
}
// Screens the SB candidates at slots base + t0 .. base + t0 + SB - 1 with
// `pick(j, b)` and returns their bits (bit u for slot base + t0 + u).  The
// candidates are read as x* records from xs or, where rec has them, as 3 x
// SB/4 loads from the position planes (pick then sees w = 0).  A batch may
// read past the end of its run (kWalkPad).  SB is 8, or 4 in the passes that
// screen only the rare chunks past their screen record, where the batch's
// registers would otherwise set the pass's register budget.
template <int SB, class Pick>
__device__ __forceinline__ uint32_t screen_batch(const float4 *__restrict__ xs, const ScreenRec &rec, int base,
                                                 int t0, Pick &&pick) {
    static_assert(SB == 4 || SB == 8, "screen batches of 4 or 8 candidates");
    constexpr int QH = SB / 4;  // float4_a4 loads per coordinate plane
    float4 bb[SB];
    if (rec.planes) {
        const float *__restrict__ px = rec.planes + (base + t0);
        float4_a4 q[3 * QH];  // x, y, z of candidates 0-3 (and 4-7)
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int h = 0; h < QH; ++h)
                q[QH * a + h] = *reinterpret_cast<const float4_a4 *>(px + (size_t)a * rec.pstride + 4 * h);
#pragma unroll
        for (int u = 0; u < SB; ++u)
            bb[u] = make_float4(q[u >> 2][u & 3], q[QH + (u >> 2)][u & 3], q[2 * QH + (u >> 2)][u & 3], 0.0f);
    } else {
        const float4 *__restrict__ p = xs + (base + t0);
#pragma unroll
        for (int u = 0; u < SB; ++u) bb[u] = p[u];
    }
    uint32_t m = 0u;  // constant shifts within the batch
#pragma unroll
    for (int u = 0; u < SB; ++u) m |= (uint32_t)pick(base + t0 + u, bb[u]) << u;
    return m;
}
// SB: candidates per screen batch (screen_batch).
template <int SB = 8, class Pick, class Allow, class Consume>
__device__ __forceinline__ void walk_chunks(const GridDesc &G, const int2 *__restrict__ ranges, const int g[3],
                                            const float4 *__restrict__ xs, int jmin, Pick &&pick, Allow &&allow,
                                            ScreenRec rec, Consume &&consume) {
    int c = 0;  // chunk number in walk order
    // a replayed chunk's mask is loaded one chunk ahead
    uint32_t ahead = rec.in ? rec.in[rec.slot] : 0u;
    for_each_run(G, ranges, g, [&](int s, int e) {
        if (s < jmin) s = jmin;
        for (int base = s; base < e; base += 32, ++c) {
            const int cnt = (e - base) < 32 ? (e - base) : 32;
            const uint32_t valid = (cnt >= 32) ? 0xffffffffu : ((1u << cnt) - 1u);
            const uint32_t keep = allow(base) & valid;
            uint32_t mask = 0u;
            if (rec.in && c < kRecChunks) {
                mask = ahead;
                if (c + 1 < kRecChunks) ahead = rec.in[(size_t)(c + 1) * rec.n + rec.slot];
                if (keep == 0u) continue;
            } else {
                if (keep == 0u && !rec.out) continue;
                for (int t0 = 0; t0 < cnt; t0 += SB) mask |= screen_batch<SB>(xs, rec, base, t0, pick) << t0;
                mask &= valid;
                if (rec.out && c < kRecChunks) rec.out[(size_t)c * rec.n + rec.slot] = mask;
            }
            mask &= keep;
            if (mask) consume(base, mask);
        }
    });
}

// Marks every word of a fetched record as used here, so the compiler keeps its
// loads where they were issued instead of sinking each into the branch of the
// pair term that reads it (which costs one memory round trip per branch).
template <class T> __device__ __forceinline__ void pin_loaded(const T &v) {
    static_assert(sizeof(T) % 4 == 0, "records are whole words");
    const uint32_t *w = reinterpret_cast<const uint32_t *>(&v);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) asm volatile("" ::"v"(w[i]));
}

// Pair bodies, NB partners at a time: `fetch(j)` loads everything the pair
// term needs, for NB partners, as independent loads that complete together;
// then `use(j, rec)` runs for each in ascending j.  next(u, ok) yields the
// u-th partner and whether it exists (a missing one fetches the first partner
// again and is not used).
template <int NB, class Next, class Fetch, class Use>
__device__ __forceinline__ void run_bodies(Next &&next, Fetch &&fetch, Use &&use) {
    int j[NB];
    bool ok[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        j[u] = next(u, ok[u]);
        if (u > 0 && !ok[u]) j[u] = j[0];
    }
    decltype(fetch(0)) rv[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) rv[u] = fetch(j[u]);
#pragma unroll
    for (int u = 0; u < NB; ++u) pin_loaded(rv[u]);
#pragma unroll
    for (int u = 0; u < NB; ++u)
        if (ok[u]) use(j[u], rv[u]);
}

// The walk with its pair bodies, run per chunk: a wave executes the body
// sum-over-chunks(max over lanes of the chunk's partners) times, in ascending
// j per particle.
constexpr int kBodyBatch = 2;
// Screen batch of the passes that replay a screen record and screen only the
// chunks past it (position, vorticity/tension, XSPH): 4 candidates, so the
// rare screen does not set their register budget.  Their VGPRs fall by 14-16
// and one to two more waves per SIMD fit (batch position 1.27 -> 1.16 ms,
// vorticity/tension 0.71 -> 0.59 ms per launch).
constexpr int kReplayScreen = 4;
template <int NB, int SB = 8, class Pick, class Fetch, class Use, class Allow>
__device__ __forceinline__ void for_each_neighbor_nb(const GridDesc &G, const int2 *__restrict__ ranges,
                                                     const int g[3], const float4 *__restrict__ xs, int jmin,
                                                     Pick &&pick, Fetch &&fetch, Use &&use, Allow &&allow,
                                                     ScreenRec rec = ScreenRec{}) {
    walk_chunks<SB>(G, ranges, g, xs, jmin, pick, allow, rec, [&](int base, uint32_t mask) {
        while (mask) {
            run_bodies<NB>(
                [&](int, bool &ok) {
                    ok = mask != 0u;
                    const int j = base + __ffs(mask) - 1;
                    mask &= mask - 1u;
                    return j;
                },
                fetch, use);
        }
    });
}
template <class Pick, class Body, class Allow>
__device__ __forceinline__ void for_each_neighbor(const GridDesc &G, const int2 *__restrict__ ranges,
                                                  const int g[3], const float4 *__restrict__ xs, int jmin,
                                                  Pick &&pick, Body &&body, Allow &&allow,
                                                  ScreenRec rec = ScreenRec{}) {
    for_each_neighbor_nb<1>(
        G, ranges, g, xs, jmin, pick, [&](int j) { return ld32(xs, j); }, body, allow, rec);
}
template <class Pick, class Body>
__device__ __forceinline__ void for_each_neighbor(const GridDesc &G, const int2 *__restrict__ ranges,
                                                  const int g[3], const float4 *__restrict__ xs, int jmin,
                                                  Pick &&pick, Body &&body) {
    for_each_neighbor(G, ranges, g, xs, jmin, pick, body, AllowAll{});
}

// ------------------------------------------------------------ flat walk
// The walk with its pair bodies run over the whole walk rather than per
// chunk.  A wave runs a chunk's bodies as often as its busiest lane has
// partners in that chunk, so the per-chunk form costs the sum over chunks of
// that maximum; the flat form costs the maximum over lanes of all partners.
// The walk appends each chunk's (base, mask) to a list of kListCap entries per
// lane in LDS, and the bodies then drain the list in order (chunks in walk
// order, partners ascending within a chunk): the pairs and their summation
// order are those of for_each_neighbor_nb, so results are bit-identical.
// The wave drains when one of its lanes' lists is full and after the walk.
// A column's runs share one chunk loop, so the chunk code appears once in the
// kernel.  Launch with flat_lds_bytes(block) bytes of dynamic LDS.
constexpr int kListCap = 12;
inline size_t flat_lds_bytes(int block) { return (size_t)block * kListCap * 8; }
typedef unsigned int u2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) u2v lds_u2;
extern __shared__ u2v bp_flat_lds[];

template <int NB, int SB = 8, class Pick, class Fetch, class Use, class Allow>
__device__ __forceinline__ void for_each_neighbor_flat(const GridDesc &G, const int2 *__restrict__ ranges,
                                                       const int g[3], const float4 *__restrict__ xs, int jmin,
                                                       Pick &&pick, Fetch &&fetch, Use &&use, Allow &&allow,
                                                       ScreenRec rec) {
    lds_u2 *const list = (lds_u2 *)bp_flat_lds + threadIdx.x;  // entry e at list[e * stride]
    const int stride = (int)blockDim.x;
    const uint32_t zc[3] = {(uint32_t)(g[2] - 1) & G.mask[2], (uint32_t)g[2] & G.mask[2],
                            (uint32_t)(g[2] + 1) & G.mask[2]};
    const bool brk01 = zc[1] != zc[0] + 1u, brk12 = zc[2] != zc[1] + 1u;
    const bool adjacent = !brk01 && !brk12;
    // cells [0, kw) and [kw, 3) of a column are one run each (kw = 3: no wrap)
    const int kw = brk01 ? 1 : (brk12 ? 2 : 3);
    auto column = [&](int ci) {
        const int dy = ci / 3 - 1, dx = ci % 3 - 1;
        return G.base + ((uint32_t)(g[0] + dx) & G.mask[0]) * G.stride_x +
               ((uint32_t)(g[1] + dy) & G.mask[1]) * (uint32_t)G.dims[2];
    };
    auto load_column = [&](uint32_t col, int2 r[3]) {
        if (adjacent) {
            const int4_a8 a = *reinterpret_cast<const int4_a8 *>(ranges + col + zc[0]);
            r[0] = make_int2(a.x, a.y);
            r[1] = make_int2(a.z, a.w);
            r[2] = ranges[col + zc[2]];
        } else {
#pragma unroll
            for (int q = 0; q < 3; ++q) r[q] = ranges[col + zc[q]];
        }
    };
    // a column's runs, [s0, e0) then [s1, e1), from jmin on; an empty run has s >= e
    int s0, e0, s1, e1;
    auto runs_of = [&](const int2 r[3]) {
        s0 = -1, e0 = -1, s1 = -1, e1 = -1;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (r[q].x >= 0) {
                if (q < kw) {
                    if (s0 < 0) s0 = r[q].x;
                    e0 = r[q].y;
                } else {
                    if (s1 < 0) s1 = r[q].x;
                    e1 = r[q].y;
                }
            }
        }
        s0 = s0 < jmin ? jmin : s0;
        s1 = s1 < jmin ? jmin : s1;
    };

    // the drain: every listed partner, NB at a time
    int cnt = 0;
    auto drain = [&]() {
        int e = 0, b = 0;
        uint32_t m = 0u;
        u2v nxt = cnt > 0 ? list[0] : u2v{0u, 0u};
        while (__builtin_amdgcn_ballot_w64((m != 0u) | (e < cnt))) {
            if ((m != 0u) | (e < cnt)) {
                run_bodies<NB>(
                    [&](int, bool &ok) {
                        if ((m == 0u) & (e < cnt)) {
                            b = (int)nxt.x;
                            m = nxt.y;
                            ++e;
                            if (e < cnt) nxt = list[e * stride];
                        }
                        ok = m != 0u;
                        const int j = b + __ffs(m) - 1;
                        m &= m - 1u;
                        return j;
                    },
                    fetch, use);
            }
        }
        cnt = 0;
    };

    int c = 0;  // chunk number in walk order
    uint32_t ahead = rec.in ? rec.in[rec.slot] : 0u;
    int2 nr[3];
    load_column(column(0), nr);
    for (int ci = 0; ci < 9; ++ci) {
        const int2 r[3] = {nr[0], nr[1], nr[2]};
        if (ci + 1 < 9) load_column(column(ci + 1), nr);
        runs_of(r);
#pragma unroll 1
        for (int rr = 0; rr < 2; ++rr) {
            const int end = rr ? e1 : e0;
            for (int base = rr ? s1 : s0; base < end; base += 32, ++c) {
                const int cn = (end - base) < 32 ? (end - base) : 32;
                const uint32_t valid = (cn >= 32) ? 0xffffffffu : ((1u << cn) - 1u);
                const uint32_t keep = allow(base) & valid;
                uint32_t mask = 0u;
                if (rec.in && c < kRecChunks) {
                    mask = ahead;
                    if (c + 1 < kRecChunks) ahead = rec.in[(size_t)(c + 1) * rec.n + rec.slot];
                } else if (keep != 0u || rec.out) {
                    for (int t0 = 0; t0 < cn; t0 += SB)
                        mask |= screen_batch<SB>(xs, rec, base, t0, pick) << t0;
                    mask &= valid;
                    if (rec.out && c < kRecChunks) rec.out[(size_t)c * rec.n + rec.slot] = mask;
                }
                mask &= keep;
                if (mask) {
                    list[cnt * stride] = u2v{(uint32_t)base, mask};
                    ++cnt;
                }
                if (__builtin_amdgcn_ballot_w64(cnt == kListCap) != 0ull) drain();
            }
        }
    }
    drain();
}

// ------------------------------------------------------------ staged walk
// The per-chunk walk with its candidates read from LDS.  A wave's lanes hold
// consecutive sorted slots, so they search few distinct (x, y) cell columns
// and overlapping z windows: for each of the 9 column offsets, the union of
// the lanes' candidate runs is a few contiguous slot segments (one per group
// of lanes sharing a current column).  The wave loads those segments' records
// into its LDS once, with coalesced loads spread over its lanes, and each lane
// then screens its runs and fetches its partners from LDS, in the walk's
// order: same pairs, same summation order, same results.  A column whose
// segments exceed Cap records, and a walk whose lanes fall in more than
// kStageGroups columns or whose z window wraps, reads global memory as the
// per-chunk walk does (the x* planes, or srcs[0], for the screen; srcs for
// the bodies).  srcs[0] holds x* (the screen's records); the pair body gets
// srcs[0..NR-1][j].  Launch with staged_lds_bytes(block, NR, Cap) bytes of
// dynamic LDS.
//
// Cap, the region's size in records per wave and column offset (about eight
// per occupied cell), is a tuning knob: a larger region stages the wider
// segments of waves that span several columns, a smaller one leaves more LDS
// for resident waves.  kStageCaps lists the sizes built.
constexpr int kStageCaps[3] = {128, 256, 512};
constexpr int kStageCapDefault = 256;
constexpr int kStageGroups = 4;  // lane groups (distinct current columns) per wave
constexpr int kStagePad = 8;     // records after the region: a screen batch reads up to 7 past a run
constexpr int kCopyBatch = 2;    // records a lane copies into the region per round of loads
inline bool stage_cap_valid(int cap) { return cap == 128 || cap == 256 || cap == 512; }
// nr staged arrays per wave of `cap` records each
inline size_t staged_lds_bytes(int block, int nr, int cap) {
    return (size_t)(block / 64) * nr * (cap + kStagePad) * 16;
}
// The records a staged walk's pair body receives: srcs[0..NR-1][j]
template <int NR> struct StagedRec {
    float4 r[NR];
};
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) f4v lds_f4;
__device__ __forceinline__ float4 as_float4(f4v v) { return make_float4(v.x, v.y, v.z, v.w); }
__device__ __forceinline__ f4v as_f4v(float4 a) { return f4v{a.x, a.y, a.z, a.w}; }

template <int NB, int NR, int Cap, int SB = 8, class Pick, class Use, class Allow>
__device__ __forceinline__ void for_each_neighbor_staged(const GridDesc &G, const int2 *__restrict__ ranges,
                                                         const int g[3], const float4 *const (&srcs)[NR], int jmin,
                                                         Pick &&pick, Use &&use, Allow &&allow, ScreenRec rec) {
    static_assert(Cap == 128 || Cap == 256 || Cap == 512, "staging region sizes are those of kStageCaps");
    constexpr int kRegion = Cap + kStagePad;
    lds_f4 *const lds = (lds_f4 *)bp_flat_lds + (threadIdx.x >> 6) * (NR * kRegion);
    const uint64_t act = __builtin_amdgcn_ballot_w64(true);
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
    const int nact = __builtin_popcountll(act);
    // ---- lane groups: lanes sharing a current (x, y) column.  A group's z
    // window spans its first and last lanes' z cells +- 2 (every lane of the
    // group must lie within one cell of those two, else no staging).  Its wn
    // cells' ranges are loaded by the active lanes of ranks wl .. wl + wn - 1
    // (lane mask wmask).
    const uint32_t ckey = (((uint32_t)g[0] & G.mask[0]) << 10) | ((uint32_t)g[1] & G.mask[1]);
    uint64_t wmask[kStageGroups];
    int ng = 0, wl = 0, grp = 0;
    int wq_cx = 0, wq_cy = 0, wq_z = 0;  // the window cell this lane loads (wq)
    bool wq = false, stage = true;
    uint64_t todo = act;
#pragma unroll
    for (int q = 0; q < kStageGroups; ++q) {
        wmask[q] = 0ull;
        if (todo == 0ull || !stage) continue;
        const int lead = __builtin_ctzll(todo);
        const uint32_t k0 = __builtin_amdgcn_readlane(ckey, lead);
        const bool mine = ckey == k0;
        const uint64_t m = __builtin_amdgcn_ballot_w64(mine);
        const int za = __builtin_amdgcn_readlane(g[2], __builtin_ctzll(m));
        const int zb = __builtin_amdgcn_readlane(g[2], 63 - __builtin_clzll(m));
        const int zlo = (za < zb ? za : zb) - 1, zhi = (za < zb ? zb : za) + 1;
        const int wn = zhi - zlo + 3;  // window cells zlo - 1 .. zhi + 1
        const uint32_t wz0 = (uint32_t)(zlo - 1) & G.mask[2];
        const bool outside = mine && (g[2] < zlo || g[2] > zhi);
        if (__builtin_amdgcn_ballot_w64(outside) != 0ull || wl + wn > nact ||
            wz0 + (uint32_t)wn > (uint32_t)G.dims[2]) {
            stage = false;
            continue;
        }
        if (mine) grp = q;
        const bool inq = rank >= wl && rank < wl + wn;
        wmask[q] = __builtin_amdgcn_ballot_w64(inq);
        if (inq) {
            wq = true;
            wq_cx = (int)(k0 >> 10);
            wq_cy = (int)(k0 & 1023u);
            wq_z = (int)wz0 + (rank - wl);
        }
        wl += wn;
        ng = q + 1;
        todo &= ~m;
    }
    if (todo != 0ull) stage = false;  // more columns than groups

    const uint32_t zc[3] = {(uint32_t)(g[2] - 1) & G.mask[2], (uint32_t)g[2] & G.mask[2],
                            (uint32_t)(g[2] + 1) & G.mask[2]};
    const bool brk01 = zc[1] != zc[0] + 1u, brk12 = zc[2] != zc[1] + 1u;
    const bool adjacent = !brk01 && !brk12;
    const int kw = brk01 ? 1 : (brk12 ? 2 : 3);
    auto column = [&](int ci) {
        const int dy = ci / 3 - 1, dx = ci % 3 - 1;
        return G.base + ((uint32_t)(g[0] + dx) & G.mask[0]) * G.stride_x +
               ((uint32_t)(g[1] + dy) & G.mask[1]) * (uint32_t)G.dims[2];
    };
    auto load_column = [&](uint32_t col, int2 r[3]) {
        if (adjacent) {
            const int4_a8 a = *reinterpret_cast<const int4_a8 *>(ranges + col + zc[0]);
            r[0] = make_int2(a.x, a.y);
            r[1] = make_int2(a.z, a.w);
            r[2] = ranges[col + zc[2]];
        } else {
#pragma unroll
            for (int q = 0; q < 3; ++q) r[q] = ranges[col + zc[q]];
        }
    };
    int c = 0;  // chunk number in walk order
    uint32_t ahead = rec.in ? rec.in[rec.slot] : 0u;
    int2 nr[3];
    load_column(column(0), nr);
    for (int ci = 0; ci < 9; ++ci) {
        const int dy = ci / 3 - 1, dx = ci % 3 - 1;
        const int2 r[3] = {nr[0], nr[1], nr[2]};
        if (ci + 1 < 9) load_column(column(ci + 1), nr);
        // ---- this column offset's segments, staged when they fit
        bool staged = false;
        int ldsbase = 0;  // LDS index of slot j: j + ldsbase (this lane's group)
        if (stage) {
            int2 wr = make_int2(-1, -1);
            if (wq)
                wr = ranges[G.base + ((uint32_t)(wq_cx + dx) & G.mask[0]) * G.stride_x +
                            ((uint32_t)(wq_cy + dy) & G.mask[1]) * (uint32_t)G.dims[2] + (uint32_t)wq_z];
            const uint64_t ne = __builtin_amdgcn_ballot_w64(wr.x >= 0);
            int tot = 0, seg0[kStageGroups], sbase[kStageGroups];
#pragma unroll
            for (int q = 0; q < kStageGroups; ++q) {
                seg0[q] = tot;
                sbase[q] = 0;
                if (q < ng) {
                    const uint64_t n1 = ne & wmask[q];
                    const int s1 = n1 ? __builtin_amdgcn_readlane(wr.x, __builtin_ctzll(n1)) : 0;
                    const int e1 = n1 ? __builtin_amdgcn_readlane(wr.y, 63 - __builtin_clzll(n1)) : 0;
                    sbase[q] = tot - s1;  // LDS index = slot + sbase
                    tot += e1 - s1;
                }
            }
            if (tot <= Cap) {
                staged = true;
                // kCopyBatch records per lane in flight before their LDS stores
                // (one at a time, each store waited for its load)
                for (int t0 = rank; t0 < tot; t0 += kCopyBatch * nact) {
                    f4v buf[kCopyBatch][NR];
#pragma unroll
                    for (int u = 0; u < kCopyBatch; ++u) {
                        const int t = t0 + u * nact;
                        int b = sbase[0];
#pragma unroll
                        for (int q = 1; q < kStageGroups; ++q)
                            if (q < ng && t >= seg0[q]) b = sbase[q];
                        if (t < tot) {
#pragma unroll
                            for (int a = 0; a < NR; ++a) buf[u][a] = as_f4v(ld32(srcs[a], t - b));
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kCopyBatch; ++u) {
                        const int t = t0 + u * nact;
                        if (t < tot) {
#pragma unroll
                            for (int a = 0; a < NR; ++a) lds[a * kRegion + t] = buf[u][a];
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                for (int q = 0; q < kStageGroups; ++q)
                    if (q == grp) ldsbase = sbase[q];
            }
        }
        // ---- this lane's runs of the column, chunk by chunk (as walk_chunks).
        // The staged and global forms are separate instances of the chunk
        // loop: merged into one loop, the staged form's LDS reads waited for
        // every global load in flight (the two forms' loads shared registers),
        // including the next column's prefetched ranges.
        auto run = [&](auto lds_c, int s, int e) {
            constexpr bool kLds = decltype(lds_c)::value;
            if (s < jmin) s = jmin;
            for (int base = s; base < e; base += 32, ++c) {
                const int cnt = (e - base) < 32 ? (e - base) : 32;
                const uint32_t valid = (cnt >= 32) ? 0xffffffffu : ((1u << cnt) - 1u);
                const uint32_t keep = allow(base) & valid;
                uint32_t mask = 0u;
                if (rec.in && c < kRecChunks) {
                    mask = ahead;
                    if (c + 1 < kRecChunks) ahead = rec.in[(size_t)(c + 1) * rec.n + rec.slot];
                    if (keep == 0u) continue;
                } else {
                    if (keep == 0u && !rec.out) continue;
                    for (int t0 = 0; t0 < cnt; t0 += SB) {
                        if constexpr (kLds) {
                            float4 bb[SB];
                            const lds_f4 *q = lds + (base + t0 + ldsbase);
#pragma unroll
                            for (int u = 0; u < SB; ++u) bb[u] = as_float4(q[u]);
                            uint32_t m8 = 0u;
#pragma unroll
                            for (int u = 0; u < SB; ++u) m8 |= (uint32_t)pick(base + t0 + u, bb[u]) << u;
                            mask |= m8 << t0;
                        } else {
                            mask |= screen_batch<SB>(srcs[0], rec, base, t0, pick) << t0;
                        }
                    }
                    mask &= valid;
                    if (rec.out && c < kRecChunks) rec.out[(size_t)c * rec.n + rec.slot] = mask;
                }
                mask &= keep;
                while (mask) {
                    run_bodies<NB>(
                        [&](int, bool &ok) {
                            ok = mask != 0u;
                            const int j = base + __ffs(mask) - 1;
                            mask &= mask - 1u;
                            return j;
                        },
                        [&](int j) {
                            StagedRec<NR> R;
#pragma unroll
                            for (int a = 0; a < NR; ++a) {
                                if constexpr (kLds)
                                    R.r[a] = as_float4(lds[a * kRegion + j + ldsbase]);
                                else
                                    R.r[a] = ld32(srcs[a], j);
                            }
                            return R;
                        },
                        use);
                }
            }
        };
        // the column's runs: cells [0, kw) and [kw, 3) (kw = 3 without a z wrap)
        int s0 = -1, e0 = -1, s1 = -1, e1 = -1;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (r[q].x >= 0) {
                if (q < kw) {
                    if (s0 < 0) s0 = r[q].x;
                    e0 = r[q].y;
                } else {
                    if (s1 < 0) s1 = r[q].x;
                    e1 = r[q].y;
                }
            }
        }
#pragma unroll 1
        for (int rr = 0; rr < 2; ++rr) {
            const int rs = rr ? s1 : s0, re = rr ? e1 : e0;
            if (rs >= 0) {
                if (staged)
                    run(std::true_type{}, rs, re);
                else
                    run(std::false_type{}, rs, re);
            }
        }
        if (staged) {  // every lane is done with the region before the next fill
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

// The walk in its per-chunk (Flat = false) or flat form; same pairs, same order.
template <bool Flat, int NB, int SB = 8, class... A> __device__ __forceinline__ void walk_form(A &&...a) {
    if constexpr (Flat)
        for_each_neighbor_flat<NB, SB>(static_cast<A &&>(a)...);
    else
        for_each_neighbor_nb<NB, SB>(static_cast<A &&>(a)...);
}

// Walks the CSR entries [e0, e1) in order, eight at a time: the entry
// indices, then the eight records they name (`load(en)`), are fetched as
// independent loads before `use(en, rec)` consumes them in entry order.
template <class Load, class Use>
__device__ __forceinline__ void gather_entries(const int *__restrict__ ents, int e0, int e1, Load &&load, Use &&use) {
    for (int b = e0; b < e1; b += 8) {
        int en[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) en[u] = (b + u < e1) ? ents[b + u] : ents[b];
        decltype(load(0)) rec[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) rec[u] = load(en[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (b + u < e1) use(en[u], rec[u]);
    }
}

// The slot a neighbour-pass thread owns.  Workgroups are dealt to the 8 XCDs
// round-robin (workgroup b runs on XCD b % 8), so consecutive slot blocks,
// which hold neighbouring cells, spread over all XCDs.  Renumbering blocks so
// that each XCD walks one contiguous slot range (its L2 holding its cells)
// measured slower once these passes became bound by vector-memory issue and
// by the imbalance of long walks rather than by HBM misses.
__device__ __forceinline__ int pass_slot() { return (int)blockIdx.x * (int)blockDim.x + (int)threadIdx.x; }
// The reverse (adjoint) neighbour passes, which fetch three to four records
// per partner, keep the renumbering: each XCD walks one contiguous range of
// sorted slots.  Bijective for any grid size.
__device__ __forceinline__ int xcd_range_slot() {
    constexpr int kNumXcd = 8;
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int q = nb / kNumXcd, r = nb % kNumXcd;
    const int xcd = b % kNumXcd, local = b / kNumXcd;
    const int blk = (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
    return blk * (int)blockDim.x + (int)threadIdx.x;
}

inline int grid_blocks(long long n, int block = kBlock) { return (int)((n + block - 1) / block); }

}  // namespace bp
