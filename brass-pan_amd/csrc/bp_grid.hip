// bp_grid.hip -- neighbour grid of the particle solver on gfx950:
// prediction + cell hash, stable LSD radix sort of (cell, slot), cell ranges
// and the fused reorder.  Replaces updateGrid (particlesolver.cuh:712-765),
// predictPositions/applyGravity (:49-69), inverseMapping (util.cuh:348-359),
// computeInvScaledMasses (:84-98) and planeStabilize (particleplane.cuh:3-19).
#include "bp_guard.h"
#include "bp_kernels.h"

namespace bp {

// x* = x + v dt; x* += (g dt) dt  (particlesolver.cuh:59-69, 49-57)
__device__ __forceinline__ f3 predict(float4 x4, float4 v4, float dt, f3 gstep) {
    f3 xs = xyz(x4) + xyz(v4) * dt;
    xs += gstep;
    return xs;
}

// Clears the cells set by the previous substep (resetStartEndId, grid.cuh:101-132)
// and hashes the predicted position (updateGridId, grid.cuh:30-47).  The value
// of each key is its slot, supplied implicitly to the first sort pass.
template <class SP>
__global__ __launch_bounds__(kBlock) void k_predict_hash(const float4 *__restrict__ x, const float4 *__restrict__ v,
                                                         int *__restrict__ keys, int n, float dt, f3 gstep, SP sp,
                                                         const int *__restrict__ old_sorted_keys,
                                                         int *__restrict__ ranges_flat) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (old_sorted_keys) {
        const int u = old_sorted_keys[i];
        if (i > 0) {
            const int w = old_sorted_keys[i - 1];
            if (u != w) {
                ranges_flat[2 * u] = -1;
                ranges_flat[2 * w + 1] = -1;
            }
        } else {
            ranges_flat[2 * u] = -1;
        }
        if (i == n - 1) ranges_flat[2 * u + 1] = -1;
    }
    sp.with_scene_ref(i, [&](const SceneConst &S) {
        const f3 xs = predict(x[i], v[i], dt, gstep);
        int g[3];
        S.G.pos(xs, g);
        keys[i] = S.G.addr(g[0], g[1], g[2]);
    });
}

// ---------------------------------------------------------------- radix sort
// Stable LSD radix sort, 8-bit digits, only over the bits the cell count
// needs (keys < numCells, so the order equals a sort over all 32 bits,
// cub::DeviceRadixSort::SortPairs, particlesolver.cuh:721-738).
// A block sorts a tile of kSortWaves * 64 * R keys, R keys per lane.  Below
// kSortBigN keys, R = 4: 4096-key tiles would leave most of the 256 CUs idle
// (283k keys make 70 tiles) while each wave runs 16 dependent rounds.
constexpr int kRadix = 256;
constexpr int kSortWaves = 4;
constexpr int kSortBigN = 1 << 20;
__host__ __device__ constexpr int sort_tile(int rounds) { return kSortWaves * 64 * rounds; }
static int sort_rounds(int n) { return n >= kSortBigN ? 16 : 4; }

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}
// lanes among `valid` holding the same 8-bit digit as this lane
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

// per-block digit histogram, written digit-major: counts[d * nb + block]
template <int R>
__global__ __launch_bounds__(kSortWaves * 64) void k_radix_hist(const int *__restrict__ keys, int n, int shift,
                                                               int *__restrict__ counts, int nb) {
    __shared__ int h[kRadix];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const int base = blockIdx.x * sort_tile(R);
    // each wave reads its own contiguous quarter of the tile in R rounds; the
    // R keys are loaded first, all in flight together (a load per round, each
    // waited for at the start of its round, made R serial round trips)
    const int wave = tid >> 6, lane = tid & 63;
    const int wbase = base + wave * 64 * R;
    int kreg[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int idx = wbase + r * 64 + lane;
        kreg[r] = idx < n ? keys[idx] : 0;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int idx = wbase + r * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = valid ? ((uint32_t)kreg[r] >> shift) & (kRadix - 1) : 0u;
        const uint64_t m = match_digit(d, valid);
        if (valid && (m & lanemask_lt()) == 0ull) atomicAdd(&h[d], (int)__popcll(m));
    }
    __syncthreads();
    counts[tid * nb + blockIdx.x] = h[tid];
}

// one block per digit: exclusive scan of its row of block counts, and the row total
__global__ __launch_bounds__(256) void k_radix_scan_rows(int *__restrict__ counts, int nb, int *__restrict__ totals) {
    __shared__ int s[256];
    const int d = blockIdx.x, tid = threadIdx.x;
    int *row = counts + (size_t)d * nb;
    int carry = 0;
    for (int base = 0; base < nb; base += 256) {
        const int idx = base + tid;
        const int val = idx < nb ? row[idx] : 0;
        s[tid] = val;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int t = tid >= off ? s[tid - off] : 0;
            __syncthreads();
            s[tid] += t;
            __syncthreads();
        }
        if (idx < nb) row[idx] = carry + s[tid] - val;
        carry += s[255];
        __syncthreads();
    }
    if (tid == 0) totals[d] = carry;
}

// stable scatter: rank = digit base + block prefix + earlier waves + earlier lanes
template <int R>
__global__ __launch_bounds__(kSortWaves * 64) void k_radix_scatter(const int *__restrict__ kin,
                                                                  const int *__restrict__ vin,
                                                                  int *__restrict__ kout, int *__restrict__ vout,
                                                                  int n, int shift,
                                                                  const int *__restrict__ prefix,
                                                                  const int *__restrict__ totals, int nb) {
    __shared__ int wcnt[kSortWaves][kRadix];
    __shared__ int dbase[kRadix];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wbase = blockIdx.x * sort_tile(R) + wave * 64 * R;
    // all R keys and values in flight together, before the digit scan and the
    // ranking rounds
    int kreg[R], vreg[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int idx = wbase + r * 64 + lane;
        kreg[r] = idx < n ? kin[idx] : 0;
        vreg[r] = (idx < n && vin) ? vin[idx] : idx;
    }
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) wcnt[w][tid] = 0;
    // exclusive scan of digit totals
    const int tot = totals[tid];
    dbase[tid] = tot;
    __syncthreads();
    for (int off = 1; off < kRadix; off <<= 1) {
        const int t = tid >= off ? dbase[tid - off] : 0;
        __syncthreads();
        dbase[tid] += t;
        __syncthreads();
    }
    const int excl = dbase[tid] - tot + prefix[tid * nb + blockIdx.x];
    __syncthreads();
    dbase[tid] = excl;

    int rreg[R];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int idx = wbase + r * 64 + lane;
        const bool valid = idx < n;
        const int k = kreg[r];
        const uint32_t d = ((uint32_t)k >> shift) & (kRadix - 1);
        const uint64_t m = match_digit(d, valid);
        // the lowest lane of each digit group reserves popc(m) places with one
        // LDS atomic and broadcasts the old count to its group
        const int leader = valid ? (int)__ffsll((unsigned long long)m) - 1 : lane;
        int old = 0;
        if (valid && leader == lane) old = atomicAdd(&wcnt[wave][d], (int)__popcll(m));
        old = __shfl(old, leader);
        kreg[r] = k;
        rreg[r] = valid ? old + (int)__popcll(m & lt) : 0;
    }
    __syncthreads();
    {  // earlier waves of this block, per digit
        int run = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            const int c = wcnt[w][tid];
            wcnt[w][tid] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int idx = wbase + r * 64 + lane;
        if (idx < n) {
            const uint32_t d = ((uint32_t)kreg[r] >> shift) & (kRadix - 1);
            const int pos = dbase[d] + wcnt[wave][d] + rreg[r];
            kout[pos] = kreg[r];
            vout[pos] = vreg[r];
        }
    }
}

// findStartEndId (grid.cuh:70-99) into packed (start, end) pairs
__global__ __launch_bounds__(kBlock) void k_find_ranges(const int *__restrict__ sk, int n, int *__restrict__ rf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int u = sk[i];
    if (i > 0) {
        const int w = sk[i - 1];
        if (u != w) {
            rf[2 * u] = i;
            rf[2 * w + 1] = i;
        }
    } else {
        rf[2 * u] = i;
    }
    if (i == n - 1) rf[2 * u + 1] = n;
}

// reorderParticlesData (grid.cuh:134-164) + inverseMapping (util.cuh:348-359)
// + computeInvScaledMasses (particlesolver.cuh:84-98, on x before
// stabilisation) + planeStabilize x passes (particleplane.cuh:3-19,
// particlesolver.cuh:822-833) + optionally the first iteration's plane
// collisions (particleplane.cuh:21-58).  x* is recomputed from the gathered
// x, v with the same expression k_predict_hash hashed, so it is bit-identical.
template <class SP>
__global__ __launch_bounds__(kBlock) void k_reorder(const int *__restrict__ perm, const float4 *__restrict__ x_in,
                                                    const float4 *__restrict__ v_in, const float *__restrict__ m_in,
                                                    const int *__restrict__ ph_in, const int *__restrict__ o_in,
                                                    const int *__restrict__ sid_in, float4 *__restrict__ xs_out,
                                                    float4 *__restrict__ x_out, float4 *__restrict__ v_out,
                                                    float *__restrict__ m_out, int *__restrict__ ph_out,
                                                    int *__restrict__ o_out, int *__restrict__ sid_out,
                                                    int *__restrict__ map_to_new, int n, float dt, f3 gstep,
                                                    float mass_k, int need_w, int stab_passes, int collide,
                                                    LambdaIn L, SP sp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = perm[i];
    // the source slot's scene (scenes never mix), read wave-uniformly; the
    // plane loops are unrolled over the plane slots so the planes stay in
    // scalar registers
    sp.with_scene_ref(p, [&](const SceneConst &S) {
        const Planes &pl = S.pl;
        const float4 x4 = x_in[p];
        const float4 v4 = v_in[p];
        const float m = m_in[p];
        const int ph = ph_in[p];
        const int o = o_in[p];
        f3 xs = predict(x4, v4, dt, gstep);
        f3 xp = xyz(x4);
        float w = 0.0f;
        if (need_w) {
            const float scale = powf(2.7182818284f, -mass_k * xp.y);
            w = 1.0f / (scale * m);
        }
        for (int s = 0; s < stab_passes; ++s) {
#pragma unroll
            for (int q = 0; q < kMaxPlanes; ++q) {
                if (q >= pl.n) break;
                const f3 on = mk3(pl.o[q][0], pl.o[q][1], pl.o[q][2]);
                const f3 nn = mk3(pl.nrm[q][0], pl.nrm[q][1], pl.nrm[q][2]);
                const f3 o2p = on - xp;
                const float d = dot(o2p, nn) + S.radius;
                if (d <= 0.0f) continue;
                xp += d * nn;
                xs += d * nn;
            }
        }
        if (collide) xs = plane_collide_one(xs, xp, pl, S.radius, S.friction_static, S.friction_dynamic);
        xs_out[i] = f4(xs, __int_as_float(ph));
        store_lambda_in(L, i, xs, m, ph);
        x_out[i] = f4(xp, w);
        v_out[i] = v4;
        m_out[i] = m;
        ph_out[i] = ph;
        o_out[i] = o;
        if (sid_out) sid_out[i] = sid_in[p];
        map_to_new[o] = i;
    });
}

// ------------------------------------------------------------ launchers
template <class SP>
void launch_predict_hash(const float4 *x, const float4 *v, int *keys, int n, float dt, f3 gstep, const SP &sp,
                         const int *old_sorted_keys, int2 *ranges, hipStream_t st) {
    hipLaunchKernelGGL(k_predict_hash<SP>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, x, v, keys, n, dt, gstep, sp,
                       old_sorted_keys, (int *)ranges);
}

size_t radix_counts_len(int n) {
    const int tile = sort_tile(sort_rounds(n));
    const int nb = (n + tile - 1) / tile;
    return (size_t)nb * kRadix;
}

// Sorts (keys_a, implicit slot values) over `bits` bits.  Buffers: keys_a/b,
// vals_a/b of n ints, counts of radix_counts_len(n) ints, totals of 256 ints.
// Returns in *out_keys / *out_vals which buffers hold the result.
void launch_radix_sort(int *keys_a, int *keys_b, int *vals_a, int *vals_b, int *counts, int *totals, int n, int bits,
                       int **out_keys, int **out_vals, hipStream_t st) {
    const bool big = sort_rounds(n) == 16;
    const int tile = sort_tile(sort_rounds(n));
    const int nb = (n + tile - 1) / tile;
    int *kin = keys_a, *kout = keys_b, *vin = nullptr, *vout = vals_b;
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(big ? k_radix_hist<16> : k_radix_hist<4>, dim3(nb), dim3(kSortWaves * 64), 0, st, kin, n,
                           shift, counts, nb);
        hipLaunchKernelGGL(k_radix_scan_rows, dim3(kRadix), dim3(256), 0, st, counts, nb, totals);
        hipLaunchKernelGGL(big ? k_radix_scatter<16> : k_radix_scatter<4>, dim3(nb), dim3(kSortWaves * 64), 0, st, kin,
                           vin, kout, vout, n, shift, counts, totals, nb);
        // next pass reads what this one wrote
        int *nk = kout, *nv = vout;
        kout = kin;
        vout = (vin == nullptr) ? vals_a : vin;
        kin = nk;
        vin = nv;
    }
    *out_keys = kin;
    *out_vals = vin;
}

void launch_find_ranges(const int *sorted_keys, int n, int2 *ranges, hipStream_t st) {
    hipLaunchKernelGGL(k_find_ranges, dim3(grid_blocks(n)), dim3(kBlock), 0, st, sorted_keys, n, (int *)ranges);
}

template <class SP>
void launch_reorder(const int *perm, const float4 *x_in, const float4 *v_in, const float *m_in, const int *ph_in,
                    const int *o_in, const int *sid_in, float4 *xs_out, float4 *x_out, float4 *v_out, float *m_out,
                    int *ph_out, int *o_out, int *sid_out, int *map_to_new, int n, float dt, f3 gstep, float mass_k,
                    int need_w, int stab_passes, int collide, const LambdaIn &L, const SP &sp, hipStream_t st) {
    hipLaunchKernelGGL(k_reorder<SP>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, perm, x_in, v_in, m_in, ph_in, o_in,
                       sid_in, xs_out, x_out, v_out, m_out, ph_out, o_out, sid_out, map_to_new, n, dt, gstep, mass_k,
                       need_w, stab_passes, collide, L, sp);
}

// The renderer boundary without a renderer (mapParticleInfos, main.cu:273-282):
// one {x, y, z, group id} record per particle, the group id's int bits in the
// fourth float, written in original-id order so frames line up across
// substeps.
__global__ __launch_bounds__(kBlock) void k_map_particle_infos(const float4 *__restrict__ x,
                                                               const int *__restrict__ orig,
                                                               const int *__restrict__ group,
                                                               float4 *__restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = orig[i];
    const float4 p = x[i];
    out[o] = make_float4(p.x, p.y, p.z, __int_as_float(group[o]));
}
void launch_map_particle_infos(const float4 *x, const int *orig, const int *group, float4 *out, int n,
                               hipStream_t st) {
    hipLaunchKernelGGL(k_map_particle_infos, dim3(grid_blocks(n)), dim3(kBlock), 0, st, x, orig, group, out, n);
}

// The opt-in non-finite guard (the reference's CHECK_NAN printIfNan,
// kernel/cubrasstower.cuh:15): the lowest slot whose position has a NaN or
// infinite component, by atomicMin into *first, which the host sets to n.
__global__ __launch_bounds__(kBlock) void k_first_nonfinite(const float4 *__restrict__ x, int n, int *first) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = x[i];
    if (guard_flags(p.x, p.y, p.z)) atomicMin(first, i);
}
void launch_first_nonfinite(const float4 *x, int n, int *first, hipStream_t st) {
    hipLaunchKernelGGL(k_first_nonfinite, dim3(grid_blocks(n)), dim3(kBlock), 0, st, x, n, first);
}

// Streaming copy (the bandwidth check of bp_stream_copy_bandwidth): every
// thread moves U float4 per round, grid-stride, all loads before the stores;
// NT marks them non-temporal (streamed past the caches).
template <bool NT, int U>
__global__ __launch_bounds__(kBlock) void k_stream_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                        size_t n) {
    typedef float f4e __attribute__((ext_vector_type(4)));
    const f4e *s4 = reinterpret_cast<const f4e *>(src);
    f4e *d4 = reinterpret_cast<f4e *>(dst);
    const size_t stride = (size_t)gridDim.x * kBlock;
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    for (; i + (U - 1) * stride < n; i += U * stride) {
        f4e r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) r[u] = NT ? __builtin_nontemporal_load(s4 + i + u * stride) : s4[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (NT)
                __builtin_nontemporal_store(r[u], d4 + i + u * stride);
            else
                d4[i + u * stride] = r[u];
        }
    }
    for (; i < n; i += stride) d4[i] = s4[i];
}
void launch_stream_copy(const float4 *src, float4 *dst, size_t n, int blocks, int variant, hipStream_t st) {
    switch (variant) {
    case 0: hipLaunchKernelGGL((k_stream_copy<false, 4>), dim3(blocks), dim3(kBlock), 0, st, src, dst, n); break;
    case 1: hipLaunchKernelGGL((k_stream_copy<true, 4>), dim3(blocks), dim3(kBlock), 0, st, src, dst, n); break;
    case 2: hipLaunchKernelGGL((k_stream_copy<false, 8>), dim3(blocks), dim3(kBlock), 0, st, src, dst, n); break;
    default: hipLaunchKernelGGL((k_stream_copy<true, 8>), dim3(blocks), dim3(kBlock), 0, st, src, dst, n); break;
    }
}

#define BP_INSTANTIATE_GRID(SP)                                                                                 \
    template void launch_predict_hash<SP>(const float4 *, const float4 *, int *, int, float, f3, const SP &,    \
                                          const int *, int2 *, hipStream_t);                                    \
    template void launch_reorder<SP>(const int *, const float4 *, const float4 *, const float *, const int *,    \
                                     const int *, const int *, float4 *, float4 *, float4 *, float *, int *,     \
                                     int *, int *, int *, int, float, f3, float, int, int, int,                  \
                                     const LambdaIn &, const SP &, hipStream_t);
BP_INSTANTIATE_GRID(OneScene)
BP_INSTANTIATE_GRID(SceneTable)

}  // namespace bp
