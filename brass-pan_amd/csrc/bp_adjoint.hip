// bp_adjoint.hip -- hand-written adjoint (reverse-mode derivative) of the
// particle-solver substep on gfx950.
//
// Differentiated: prediction and reorder (with plane stabilisation and the
// inverse scaled mass), aerodynamic wind (wind.cuh:3-50, with a wind vector
// and drag/lift coefficients), plane collisions (particleplane.cuh:21-58), the
// PBF projection -- density/lambda (fluid.cuh:78-159, lambda = -C/D) and the
// position correction with sCorr and the clamp (fluid.cuh:161-228) --,
// distance constraints (distance.cuh:3-41), pins (immovable.cuh:3-14), the
// velocity/position update (particlesolver.cuh:71-81,121-137) and XSPH
// (fluid.cuh:338-385).
//
// Every adjoint pass reads the recorded forward state of its pass and walks
// the same candidate set as the forward pass (the reference's 27 cells around
// the particle's cell at that pass, with the cell table of the substep), so
// the derivative is that of the forward function with its pair sets held
// fixed.  Gather form: each slot accumulates the derivative terms it owns;
// the terms of a pair whose partner's walk does not reach the slot are
// delivered through 2^-40 fixed-point integer atomics (order-independent), so
// the adjoint is deterministic.
#include "bp_kernels.h"

namespace bp {

namespace {

constexpr double kAdjScale = 1099511627776.0;  // 2^40
constexpr int kCorrStride = 6;                  // delivered components per slot

// Range of a delivered term.  |v| <= 2^17 keeps one term below 2^57 in
// 2^-40 fixed point, so a slot's sum of up to 64 terms of one sign cannot
// wrap the int64.  A larger or non-finite term is clamped and counted in the
// guard words after the last slot (corr[kCorrStride * n]: count,
// [kCorrStride * n + 1]: the largest |term| delivered, as float bits);
// bp_backward fails with BP_ERR_NOT_FINITE when the count is not zero.
constexpr float kAdjMaxTerm = 131072.0f;  // 2^17
__device__ __forceinline__ unsigned long long afix(float v, unsigned long long *guard) {
    const float a = fabsf(v);
    if (!(a <= kAdjMaxTerm)) {
        atomicAdd(&guard[0], 1ull);
        v = (v != v) ? 0.0f : copysignf(kAdjMaxTerm, v);
    }
    atomicMax(&guard[1], (unsigned long long)__float_as_uint(a == a ? a : INFINITY));
    return (unsigned long long)(long long)rint((double)v * kAdjScale);
}
__device__ __forceinline__ float aunfix(unsigned long long u) { return (float)((double)(long long)u / kAdjScale); }
__device__ __forceinline__ void deliver3(unsigned long long *corr, int j, int c0, f3 v, unsigned long long *guard) {
    atomicAdd(&corr[kCorrStride * j + c0 + 0], afix(v.x, guard));
    atomicAdd(&corr[kCorrStride * j + c0 + 1], afix(v.y, guard));
    atomicAdd(&corr[kCorrStride * j + c0 + 2], afix(v.z, guard));
}
// adds and clears delivered components [c0, c0 + 3) of slot k
__device__ __forceinline__ f3 take3(unsigned long long *corr, int k, int c0) {
    unsigned long long *p = corr + kCorrStride * k + c0;
    const unsigned long long a = p[0], b = p[1], c = p[2];
    if ((a | b | c) == 0ull) return mk3(0.f, 0.f, 0.f);
    p[0] = 0ull;
    p[1] = 0ull;
    p[2] = 0ull;
    return mk3(aunfix(a), aunfix(b), aunfix(c));
}
__device__ __forceinline__ float take1(unsigned long long *corr, int k, int c0) {
    unsigned long long *p = corr + kCorrStride * k + c0;
    const unsigned long long a = p[0];
    if (a == 0ull) return 0.0f;
    p[0] = 0ull;
    return aunfix(a);
}

// d poly6 / d r^2 = -3 c1 (h^2 - r^2)^2
__device__ __forceinline__ float poly6_d(const SphConst &K, float d2) {
    const float t = K.h2 - d2;
    return -3.0f * K.c1 * t * t;
}

// Pair coefficient c of the position correction (delta = c * diff) and its
// derivatives with respect to r^2 and to lambda_i + lambda_j.
__device__ __forceinline__ void pair_coef_grad(float d2, float ls, const SphConst &K, const FluidParams &F, float &c,
                                               float &dc_dd2, float &dc_dls) {
    float sl = ls, dsl = 0.0f;
    if (!F.akinci) {
        const float t = K.h2 - d2;
        const float q = K.c1 * t * t * t / K.wdq;
        sl += -F.scorr_k * ipowf(q, F.scorr_n);
        if (F.scorr_n > 0)
            dsl = -F.scorr_k * (float)F.scorr_n * ipowf(q, F.scorr_n - 1) * (-3.0f * K.c1 * t * t / K.wdq);
    }
    const float r = fsqrt(d2);
    const float ir = frcp(r);
    const float gc = K.grad_coef(r, ir);
    c = sl * gc;
    if (fabsf(c) * r >= F.delta_clamp) {  // |delta| clamped: c = +-clamp / r
        c = copysignf(F.delta_clamp * ir, c);
        dc_dd2 = -0.5f * c * ir * ir;
        dc_dls = 0.0f;
    } else {
        const float tr = K.h - r;
        const float dgc_dr = -K.c2 * tr * (K.h + r) * ir * ir;
        dc_dd2 = dsl * gc + sl * dgc_dr * 0.5f * ir;
        dc_dls = gc;
    }
}

// one plane of particlePlaneCollisionConstraint (the body of plane_collide_one)
__device__ __forceinline__ f3 plane_fwd(f3 xs, f3 pos, f3 o, f3 n, float r, float fs, float fd) {
    const f3 o2p = o - xs;
    const float d = dot(o2p, n) + r;
    if (d <= 0.0f) return xs;
    const f3 diff = xs - pos;
    const float dn = dot(diff, n);
    const f3 dtan = diff - dn * n;
    const float dtl = length(dtan);
    const f3 resolved = d * n + xs;
    const f3 dX = resolved - pos;
    const f3 tdx = dX - dot(dX, n) * n;
    if (dtl < fs * dn) return resolved - tdx;
    return resolved - tdx * fminf(fd * -dn / dtl, 1.0f);
}
// vector-Jacobian product of plane_fwd: y = adjoint of the output; returns the
// adjoint of xs and adds that of pos to *gpos
__device__ __forceinline__ f3 plane_adj(f3 xs, f3 pos, f3 o, f3 n, float r, float fs, float fd, f3 y, f3 *gpos) {
    const f3 o2p = o - xs;
    const float d = dot(o2p, n) + r;
    if (d <= 0.0f) return y;
    const f3 diff = xs - pos;
    const float dn = dot(diff, n);
    const f3 dtan = diff - dn * n;
    const float dtl = length(dtan);
    const f3 resolved = d * n + xs;
    const f3 dX = resolved - pos;
    const f3 tdx = dX - dot(dX, n) * n;
    f3 gxs = mk3(0.f, 0.f, 0.f), gdiff = mk3(0.f, 0.f, 0.f), gtdx;
    if (dtl < fs * dn) {
        gtdx = -y;
    } else if (dtl == 0.0f && !(fd * -dn > 0.0f)) {
        // no tangential motion: the forward's factor is fminf(NaN or -inf, 1)
        // times a zero tdx; any tangential motion gives fd (-dn) / dtl <= 0,
        // so the derivative is that of the limit (no tangential term)
        gtdx = mk3(0.f, 0.f, 0.f);
    } else {
        const float ratio = fd * -dn / dtl;
        const float mu = fminf(ratio, 1.0f);
        gtdx = -mu * y;
        if (ratio < 1.0f) {  // mu = -fd dn / dtl
            const float gmu = -dot(y, tdx);
            float gdn = -fd * gmu / dtl;
            const float gdtl = fd * dn * gmu / (dtl * dtl);
            const f3 gdtan = (gdtl / dtl) * dtan;
            gdiff += gdtan;
            gdn -= dot(gdtan, n);
            gdiff += gdn * n;
        }
    }
    // tdx = dX - (dX.n) n ; dX = resolved - pos ; resolved = d n + xs ; d = (o - xs).n + r
    const f3 gdX = gtdx - dot(gtdx, n) * n;
    const f3 gres = y + gdX;
    const float gd = dot(gres, n);
    gxs += gres - gd * n;
    *gpos += -gdX;
    gxs += gdiff;
    *gpos += -gdiff;
    return gxs;
}

}  // namespace

// ------------------------------------------------------------ bookkeeping
__global__ __launch_bounds__(kBlock) void ka_inverse(const int *__restrict__ orig, int *__restrict__ map_to_new,
                                                     int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) map_to_new[orig[p]] = p;
}

// g[slot] = scale * seed[orig[slot]] (seed: float3 per original id, may be null)
__global__ __launch_bounds__(kBlock) void ka_seed(const float *__restrict__ seed, const int *__restrict__ orig,
                                                  float4 *__restrict__ g, float scale, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (!seed) {
        g[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const int o = orig[p];
    g[p] = make_float4(scale * seed[3 * o], scale * seed[3 * o + 1], scale * seed[3 * o + 2], 0.f);
}
// out[orig[slot]] = g[slot] / scale (float3 per original id)
__global__ __launch_bounds__(kBlock) void ka_unseed(const float4 *__restrict__ g, const int *__restrict__ orig,
                                                    float *__restrict__ out, float inv_scale, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int o = orig[p];
    const float4 a = g[p];
    out[3 * o] = a.x * inv_scale;
    out[3 * o + 1] = a.y * inv_scale;
    out[3 * o + 2] = a.z * inv_scale;
}

template <class SP>
__global__ __launch_bounds__(kBlock) void ka_cells(const float4 *__restrict__ xs, uint32_t *__restrict__ ccell, int n,
                                                   SP sp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GridDesc &G = sp.at(i).G;
    int g[3];
    G.pos(xyz(xs[i]), g);
    ccell[i] = G.pack(g);
}

// The projection adjoints' per-particle record, one fetch per partner: (lambda,
// m, current cell packed, 0) for the position adjoint; the position adjoint's
// apply step then replaces lambda by the density adjoint for the density
// adjoint.
template <class SP>
__global__ __launch_bounds__(kBlock) void ka_proj_rec(const float4 *__restrict__ xs, const float *__restrict__ m,
                                                      const float *__restrict__ lambda, float4 *__restrict__ rec,
                                                      int n, SP sp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GridDesc &G = sp.at(i).G;
    int g[3];
    G.pos(xyz(xs[i]), g);
    rec[i] = make_float4(lambda[i], m[i], __uint_as_float(G.pack(g)), 0.0f);
}

// ------------------------------------------------------------ XSPH + update
// Forward (fluid i): v3_i = v2_i + c sum_j (v2_j - v2_i) W(|x_i - x_j|^2),
// v2 = (x* - x) / dt.  Own terms of slot i: its equation's derivative with
// respect to v2_i and x*_i, and the terms of every partner j whose equation
// contains i (j's walk reaches i's cell).  The terms of i's equation for a
// partner j whose walk does not reach i are delivered to j.
template <class SP>
__global__ __launch_bounds__(kMaxBlock) void ka_xsph(const float4 *__restrict__ xf, const float4 *__restrict__ x1,
                                                     float inv_dt, const float4 *__restrict__ gv3,
                                                     const uint32_t *__restrict__ ccell, const int *__restrict__ keys,
                                                     const int2 *__restrict__ ranges, float4 *__restrict__ gv2_own,
                                                     float4 *__restrict__ gx_own, float *__restrict__ cacc,
                                                     unsigned long long *__restrict__ corr, int n, SP sp) {
    const int i = xcd_range_slot();
    if (i >= n) return;
    sp.with_scene(i, [&](const SceneConst &S) {
        const GridDesc &G = S.G;
        const SphConst &K = S.K;
        const float c = S.F.xsph_c;
        const float4 a = xf[i];
        const int ph = phase_of(a);
        const f3 yi = xyz(gv3[i]);
        if (ph >= 0) {
            gv2_own[i] = f4(yi, 0.f);
            gx_own[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            return;
        }
        const f3 pi = xyz(a);
        const f3 v2i = (pi - xyz(x1[i])) * inv_dt;
        const uint32_t bi = G.pack_addr(keys[i]);
        int g[3];
        unpack_cell(ccell[i], g);
        f3 gv = mk3(0.f, 0.f, 0.f), gx = mk3(0.f, 0.f, 0.f), vn = mk3(0.f, 0.f, 0.f);
        float wsum = 0.0f;
        for_each_neighbor(
            G, ranges, g, xf, 0,
            [&](int, float4 b) { return (phase_of(b) == ph) & (dist2(pi, xyz(b)) <= K.h2); },
            [&](int j, float4 b) {
                const f3 pj = xyz(b);
                const float d2 = dist2(pi, pj);
                const f3 diff = pi - pj;
                const float W = K.poly6(d2);
                const float Wp2 = 2.0f * poly6_d(K, d2);
                const f3 dv = (pj - xyz(x1[j])) * inv_dt - v2i;
                vn += dv * W;
                wsum += W;
                const float sa = c * dot(yi, dv) * Wp2;
                gx += sa * diff;
                if (G.in_block(ccell[j], bi)) {  // j's equation contains i
                    const f3 yj = xyz(gv3[j]);
                    gv += (c * W) * yj;
                    gx += (c * dot(yj, -dv) * Wp2) * diff;
                } else {  // j never sees this pair: deliver i's terms for j
                    deliver3(corr, j, 0, (c * W) * yi, corr + kCorrStride * n);
                    deliver3(corr, j, 3, -sa * diff, corr + kCorrStride * n);
                }
            },
            AllowNotSelf{i});
        gv2_own[i] = f4(yi - (c * wsum) * yi + gv, 0.f);
        gx_own[i] = f4(gx, 0.f);
        cacc[i] += dot(yi, vn);
    });
}

// Completes the XSPH adjoint and takes it through updateVelocity +
// updatePositions (v2 = (x* - x)/dt; x2 = x* if moved or fluid, else x):
// gxs = adjoint of x* after the iterations, gx1 = adjoint of x.
template <class SP>
__global__ __launch_bounds__(kBlock) void ka_xsph_apply(const float4 *__restrict__ xf, const float4 *__restrict__ x1,
                                                        const float4 *__restrict__ gx2,
                                                        const float4 *__restrict__ gv2_own,
                                                        const float4 *__restrict__ gx_own,
                                                        unsigned long long *__restrict__ corr,
                                                        float4 *__restrict__ gxs, float4 *__restrict__ gx1,
                                                        float *__restrict__ gw, float inv_dt, float eps2, int n,
                                                        SP sp) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 a = xf[k];
    const f3 gv2 = xyz(gv2_own[k]) + take3(corr, k, 0);
    const f3 gxo = xyz(gx_own[k]) + take3(corr, k, 3);
    const f3 out = xyz(a), xp = xyz(x1[k]);
    const bool moved = length2(out - xp) >= eps2 || phase_of(a) < 0;
    const f3 gX = xyz(gx2[k]);
    const f3 z = mk3(0.f, 0.f, 0.f);
    gxs[k] = f4(gv2 * inv_dt + (moved ? gX : z) + gxo, 0.f);
    gx1[k] = f4(-(gv2 * inv_dt) + (moved ? z : gX), 0.f);
    gw[k] = 0.0f;
}

// ------------------------------------------------------------ position (Δp)
// Forward: x*_k += (1/rho0) sum_{pairs (k,j)} m_j c_kj (x*_k - x*_j) over the
// reference's pair set (see k_position).  With y = adjoint / rho0 and
// u = m_j y_k - m_k y_j, each pair adds (u . diff) dc to the adjoints of
// diff and lambda_k + lambda_j.
template <class SP>
__global__ __launch_bounds__(kMaxBlock) void ka_position(const float4 *__restrict__ xs,
                                                         const float4 *__restrict__ prec,
                                                         const int *__restrict__ keys,
                                                         const int2 *__restrict__ ranges,
                                                         const float4 *__restrict__ gpost, float4 *__restrict__ gx_own,
                                                         float *__restrict__ glam_own,
                                                         unsigned long long *__restrict__ corr,
                                                         const uint32_t *__restrict__ trec,
                                                         uint32_t *__restrict__ srec, int n, SP sp) {
    const int k = xcd_range_slot();
    if (k >= n) return;
    sp.with_scene(k, [&](const SceneConst &S) {
        const GridDesc &G = S.G;
        const SphConst &K = S.K;
        const FluidParams &F = S.F;
        const float4 a = xs[k];
        const f3 pk = xyz(a);
        const bool kf = phase_of(a) <= 0;
        const float4 rk = prec[k];  // (lambda, m, cell, -)
        const float lk = rk.x;
        const float mk = rk.y;
        const uint32_t bk = G.pack_addr(keys[k]);
        const float ir0 = F.inv_rest_density;
        const f3 yk = xyz(gpost[k]) * ir0;
        int g[3];
        unpack_cell(__float_as_uint(rk.z), g);
        f3 gx = mk3(0.f, 0.f, 0.f);
        float gl = 0.0f;
        struct Partner {
            float4 b, gp;
            float4 r;  // (lambda, m, cell, -)
        };
        // A fluid particle replays the forward lambda pass's screen of this x*
        // (taped); a solid screens and records for the density adjoint.
        ScreenRec rec;
        rec.in = kf ? trec : nullptr;
        rec.out = kf ? nullptr : srec;
        rec.slot = k;
        rec.n = n;
        for_each_neighbor_nb<kBodyBatch>(
            G, ranges, g, xs, 0,
            // a solid screens for fluid partners only
            [&](int, float4 b) { return (dist2(pk, xyz(b)) <= K.h2) & (kf | (phase_of(b) <= 0)); },
            [&](int j) { return Partner{xs[j], gpost[j], prec[j]}; },
            [&](int j, const Partner &P) {
                const float4 b = P.b;
                // coincident partners are skipped; a solid partner's pair is k's only if j > k
                if (!(dist2(pk, xyz(b)) > 0.0f) | ((phase_of(b) > 0) & !(kf & (j > k)))) return;
                const bool partner_sees_k = G.in_block(__float_as_uint(P.r.z), bk);
                if (j < k && !partner_sees_k) return;
                const f3 diff = pk - xyz(b);
                float c, dc_dd2, dc_dls;
                pair_coef_grad(dist2(pk, xyz(b)), lk + P.r.x, K, F, c, dc_dd2, dc_dls);
                const f3 u = P.r.y * yk - mk * (xyz(P.gp) * ir0);
                const float cbar = dot(u, diff);
                const f3 gdiff = c * u + (2.0f * cbar * dc_dd2) * diff;
                const float gls = cbar * dc_dls;
                gx += gdiff;
                gl += gls;
                if (j > k && !partner_sees_k) {
                    deliver3(corr, j, 0, -gdiff, corr + kCorrStride * n);
                    atomicAdd(&corr[kCorrStride * j + 3], afix(gls, corr + kCorrStride * n));
                }
            },
            [&](int base) { return kf ? ~chunk_bit(k, base) : chunk_below(k, base); }, rec);
        gx_own[k] = f4(gx, 0.f);
        glam_own[k] = gl;
    });
}

// x* adjoint before the correction = after + own + delivered terms; lambda
// adjoint -> density adjoint (lambda = -(rho/rho0 - 1)/D for fluids), which
// replaces lambda in the projection record (prec.x).
template <class SP>
__global__ __launch_bounds__(kBlock) void ka_position_apply(const float4 *__restrict__ xs,
                                                            const float4 *__restrict__ gx_own,
                                                            const float *__restrict__ glam_own,
                                                            unsigned long long *__restrict__ corr,
                                                            float4 *__restrict__ gxs, float4 *__restrict__ prec, int n,
                                                            SP sp) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const FluidParams &F = sp.at(k).F;
    const f3 g = xyz(gxs[k]) + xyz(gx_own[k]) + take3(corr, k, 0);
    const float gl = glam_own[k] + take1(corr, k, 3);
    gxs[k] = f4(g, 0.f);
    const bool fluid = phase_of(xs[k]) <= 0;
    reinterpret_cast<float *>(prec + k)[0] = fluid ? -gl / (F.rest_density * F.lambda_denominator) : 0.0f;
}

// ------------------------------------------------------------ density (λ)
// Forward: rho_i = sum_{j >= i found by i's walk} s_j m_j W(|x_i - x_j|^2)
// (s_j = 1 for fluids, the solid density scale for solids).
template <class SP>
__global__ __launch_bounds__(kMaxBlock) void ka_lambda(const float4 *__restrict__ xs,
                                                       const float4 *__restrict__ prec,
                                                       const int *__restrict__ keys, const int2 *__restrict__ ranges,
                                                       float4 *__restrict__ gx_own,
                                                       unsigned long long *__restrict__ corr,
                                                       const uint32_t *__restrict__ trec,
                                                       const uint32_t *__restrict__ srec, int n, SP sp) {
    const int k = xcd_range_slot();
    if (k >= n) return;
    sp.with_scene(k, [&](const SceneConst &S) {
        const GridDesc &G = S.G;
        const SphConst &K = S.K;
        const float sds = S.F.solid_density_scale;
        const float4 a = xs[k];
        const f3 pk = xyz(a);
        const bool kf = phase_of(a) <= 0;
        const float4 pr = prec[k];  // (density adjoint, m, cell, -)
        const float rk = pr.x;
        const float mk = pr.y;
        const float sk = (phase_of(a) < 0) ? 1.0f : sds;
        const uint32_t bk = G.pack_addr(keys[k]);
        int g[3];
        unpack_cell(__float_as_uint(pr.z), g);
        f3 gx = mk3(0.f, 0.f, 0.f);
        struct Partner {
            float4 b;
            float4 r;  // (density adjoint, m, cell, -)
        };
        ScreenRec rec;  // the records the position adjoint replayed or wrote for this x*
        rec.in = kf ? trec : srec;
        rec.slot = k;
        rec.n = n;
        for_each_neighbor_nb<kBodyBatch>(
            G, ranges, g, xs, 0,
            [&](int, float4 b) { return (dist2(pk, xyz(b)) <= K.h2) & (kf | (phase_of(b) <= 0)); },
            [&](int j) { return Partner{xs[j], prec[j]}; },
            [&](int j, const Partner &P) {
                const float4 b = P.b;
                // coincident partners are skipped; a solid partner's pair is k's only if j > k
                if (!(dist2(pk, xyz(b)) > 0.0f) | ((phase_of(b) > 0) & !(kf & (j > k)))) return;
                const bool partner_sees_k = G.in_block(__float_as_uint(P.r.z), bk);
                const f3 diff = pk - xyz(b);
                const float Wp2 = 2.0f * poly6_d(K, dist2(pk, xyz(b)));
                if (j > k) {  // pair of rho_k
                    const float sj = (phase_of(b) < 0) ? 1.0f : sds;
                    const f3 t = (rk * sj * P.r.y * Wp2) * diff;
                    gx += t;
                    if (!partner_sees_k) deliver3(corr, j, 0, -t, corr + kCorrStride * n);
                } else if (partner_sees_k) {  // pair of rho_j
                    gx += (P.r.x * sk * mk * Wp2) * diff;
                }
            },
            [&](int base) { return kf ? ~chunk_bit(k, base) : chunk_below(k, base); }, rec);
        gx_own[k] = f4(gx, 0.f);
    });
}

__global__ __launch_bounds__(kBlock) void ka_add_own(float4 *__restrict__ gxs, const float4 *__restrict__ gx_own,
                                                     unsigned long long *__restrict__ corr, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    gxs[k] = f4(xyz(gxs[k]) + xyz(gx_own[k]) + take3(corr, k, 0), 0.f);
}

// ------------------------------------------------------------ planes
// Planes are applied one after another (plane_collide_one); the adjoint
// recomputes each plane's input from the recorded x* before the planes.
template <class SP>
__global__ __launch_bounds__(kBlock) void ka_planes(const float4 *__restrict__ xp, const float4 *__restrict__ x1,
                                                    float4 *__restrict__ gxs, float4 *__restrict__ gx1, int n, SP sp) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const SceneConst &S = sp.at(k);
    const Planes &pl = S.pl;
    if (pl.n == 0) return;
    const f3 x0 = xyz(xp[k]), pos = xyz(x1[k]);
    f3 y = xyz(gxs[k]), gpos = xyz(gx1[k]);
    for (int q = pl.n - 1; q >= 0; --q) {
        f3 xin = x0;
        for (int u = 0; u < q; ++u)
            xin = plane_fwd(xin, pos, mk3(pl.o[u][0], pl.o[u][1], pl.o[u][2]),
                            mk3(pl.nrm[u][0], pl.nrm[u][1], pl.nrm[u][2]), S.radius, S.friction_static,
                            S.friction_dynamic);
        y = plane_adj(xin, pos, mk3(pl.o[q][0], pl.o[q][1], pl.o[q][2]),
                      mk3(pl.nrm[q][0], pl.nrm[q][1], pl.nrm[q][2]), S.radius, S.friction_static,
                      S.friction_dynamic, y, &gpos);
    }
    gxs[k] = f4(y, 0.f);
    gx1[k] = f4(gpos, 0.f);
}

// ------------------------------------------------------------ pins, distance
__global__ __launch_bounds__(kBlock) void ka_immovable(float4 *__restrict__ gxs, float4 *__restrict__ gx1,
                                                       const int *__restrict__ ids,
                                                       const int *__restrict__ map_to_new, int ni) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni) return;
    const int slot = map_to_new[ids[t]];
    gx1[slot] = f4(xyz(gx1[slot]) + xyz(gxs[slot]), 0.f);
    gxs[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// distanceConstraints: k = n (d - L) s / (w1 + w2); dx1 = -w1 k, dx2 = +w2 k.
// Per constraint: adjoint of x12 = x1* - x2* and of w1, w2.
__global__ __launch_bounds__(kBlock) void ka_distance(const float4 *__restrict__ xd, const float4 *__restrict__ x1,
                                                      const int2 *__restrict__ pairs,
                                                      const float2 *__restrict__ params,
                                                      const int *__restrict__ map_to_new,
                                                      const float4 *__restrict__ gxs, float4 *__restrict__ cb,
                                                      float2 *__restrict__ wb, int nc) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int2 pr = pairs[c];
    const float2 pa = params[c];
    const int id1 = map_to_new[pr.x], id2 = map_to_new[pr.y];
    const f3 x12 = xyz(xd[id1]) - xyz(xd[id2]);
    const float d = length(x12);
    if (d <= 1e-5f) {
        cb[c] = make_float4(0.f, 0.f, 0.f, 0.f);
        wb[c] = make_float2(0.f, 0.f);
        return;
    }
    const f3 nn = x12 / d;
    const float w1 = x1[id1].w, w2 = x1[id2].w;
    const float sw = w1 + w2;
    const float beta = (d - pa.x) / sw * pa.y;
    const f3 k = nn * beta;
    const f3 y1 = xyz(gxs[id1]), y2 = xyz(gxs[id2]);
    const f3 kbar = -w1 * y1 + w2 * y2;
    const f3 nbar = beta * kbar;
    const float bbar = dot(nn, kbar);
    const float dbar = bbar * pa.y / sw;
    const float swbar = -bbar * beta / sw;
    const f3 g12 = (nbar - dot(nbar, nn) * nn) / d + dbar * nn;
    cb[c] = f4(g12, 1.0f);
    wb[c] = make_float2(-dot(y1, k) + swbar, dot(y2, k) + swbar);
}
__global__ __launch_bounds__(kBlock) void ka_distance_gather(float4 *__restrict__ gxs, float *__restrict__ gw,
                                                             const int *__restrict__ verts,
                                                             const int *__restrict__ offs,
                                                             const int *__restrict__ ents,
                                                             const float4 *__restrict__ cb,
                                                             const float2 *__restrict__ wb,
                                                             const int *__restrict__ map_to_new, int nv) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;
    const int slot = map_to_new[verts[t]];
    f3 g = mk3(0.f, 0.f, 0.f);
    float w = 0.0f;
    gather_entries(
        ents, offs[t], offs[t + 1], [&](int en) { return cb[en >> 1]; },
        [&](int en, float4 c4) {
            if (c4.w == 0.0f) return;
            const float2 w2 = wb[en >> 1];
            if (en & 1) {
                g -= xyz(c4);
                w += w2.y;
            } else {
                g += xyz(c4);
                w += w2.x;
            }
        });
    gxs[slot] = f4(xyz(gxs[slot]) + g, 0.f);
    gw[slot] += w;
}

// ------------------------------------------------------------ wind
// applyWindForce with wind vector W: per face F = drag + lift,
//   drag = -cd |v|^2 A (n.v^) v^,  lift = cl |v|^2 A (n.v^) ((n x v^) x v^),
// v = face-average velocity - W, n the unit normal turned towards v,
// A = |(p2-p1) x (p3-p1)| / 2; each vertex moves by F dt^2 / m.
// Per face: adjoints of the three vertex positions, of the average velocity
// (a third to each vertex), and accumulated d/dW, d/dcd, d/dcl.
template <class SP>
__global__ __launch_bounds__(kBlock) void ka_wind_face(const float4 *__restrict__ xsw, const float4 *__restrict__ v_in,
                                                       const int *__restrict__ perm, const float *__restrict__ m,
                                                       const int4 *__restrict__ faces,
                                                       const int *__restrict__ map_to_new,
                                                       const float4 *__restrict__ gxs, float4 *__restrict__ fbar,
                                                       float4 *__restrict__ facc, int nf, SP sp, float dt) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int4 ids = faces[f];
    const int id1 = map_to_new[ids.x], id2 = map_to_new[ids.y], id3 = map_to_new[ids.z];
    sp.with_scene(id1, [&](const SceneConst &S) {
        const float cd = S.drag_coeff, cl = S.lift_coeff;
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        fbar[4 * f] = z4;
        fbar[4 * f + 1] = z4;
        fbar[4 * f + 2] = z4;
        fbar[4 * f + 3] = z4;
        const f3 p1 = xyz(xsw[id1]), p2 = xyz(xsw[id2]), p3 = xyz(xsw[id3]);
        const f3 a = p2 - p1, b = p3 - p1;
        f3 vel = (xyz(v_in[perm[id1]]) + xyz(v_in[perm[id2]]) + xyz(v_in[perm[id3]])) / 3.0f;
        vel = vel - S.wind;
        const float lv = length(vel);
        if (lv <= 1e-5f) return;
        const f3 vh = vel / lv;
        const f3 fr = cross(a, b);
        const float lfn = length(fr);
        if (lfn <= 1e-5f) return;
        const f3 fu = fr / lfn;
        const float sg = (dot(fu, vel) > 0.0f) ? 1.0f : -1.0f;
        const f3 fn = sg * fu;
        const float area = 0.5f * lfn;
        const float nv = dot(fn, vh);
        const float A = lv * lv * area * nv;
        const f3 wv = cross(fn, vh);
        const f3 uh = cross(wv, vh);
        // adjoint of the force
        const f3 F = dt * dt * (xyz(gxs[id1]) / m[id1] + xyz(gxs[id2]) / m[id2] + xyz(gxs[id3]) / m[id3]);
        const float gcd = -A * dot(F, vh);
        const float gcl = A * dot(F, uh);
        const float gA = dot(F, -cd * vh + cl * uh);
        f3 gvh = (-cd * A) * F;
        const f3 gu = (cl * A) * F;
        const f3 gwv = cross(vh, gu);
        gvh += cross(gu, wv);
        f3 gfn = cross(vh, gwv);
        gvh += cross(gwv, fn);
        const float glv = gA * 2.0f * lv * area * nv;
        const float garea = gA * lv * lv * nv;
        const float gnv = gA * lv * lv * area;
        gfn += gnv * vh;
        gvh += gnv * fn;
        const f3 gfu = sg * gfn;
        const float glfn = 0.5f * garea;
        const f3 gfr = (gfu - dot(gfu, fu) * fu) / lfn + glfn * fu;
        const f3 ga = cross(b, gfr);
        const f3 gb = cross(gfr, a);
        const f3 gvel = (gvh - dot(gvh, vh) * vh) / lv + glv * vh;
        fbar[4 * f] = f4(-ga - gb, 1.0f);
        fbar[4 * f + 1] = f4(ga, 1.0f);
        fbar[4 * f + 2] = f4(gb, 1.0f);
        fbar[4 * f + 3] = f4(gvel / 3.0f, 1.0f);
        const float4 acc0 = facc[2 * f], acc1 = facc[2 * f + 1];
        facc[2 * f] = make_float4(acc0.x - gvel.x, acc0.y - gvel.y, acc0.z - gvel.z, acc0.w + gcd);
        facc[2 * f + 1] = make_float4(acc1.x + gcl, acc1.y, acc1.z, acc1.w);
    });
}
// per vertex (CSR of (face, corner) entries): position and velocity adjoints
__global__ __launch_bounds__(kBlock) void ka_wind_gather(float4 *__restrict__ gxs, float4 *__restrict__ gv1,
                                                         const int *__restrict__ verts, const int *__restrict__ offs,
                                                         const int *__restrict__ ents,
                                                         const float4 *__restrict__ fbar,
                                                         const int *__restrict__ map_to_new, int nv) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;
    const int slot = map_to_new[verts[t]];
    f3 gx = mk3(0.f, 0.f, 0.f), gv = mk3(0.f, 0.f, 0.f);
    gather_entries(
        ents, offs[t], offs[t + 1], [&](int en) { return fbar[4 * (en / 3) + (en - 3 * (en / 3))]; },
        [&](int en, float4 p) {
            gx += xyz(p);
            gv += xyz(fbar[4 * (en / 3) + 3]);
        });
    gxs[slot] = f4(xyz(gxs[slot]) + gx, 0.f);
    gv1[slot] = f4(xyz(gv1[slot]) + gv, 0.f);
}

// ------------------------------------------------------------ reorder
// Forward (new slot p, source o = perm[p]): x0 = x[o], v1 = v[o];
// x* = x0 + v1 dt + g dt^2; w = e^(k y0) / m; plane stabilisation shifts
// x and x* by d n where d = (o - x).n + r > 0.  Writes the adjoints of the
// substep's input state in the previous slot order.
template <class SP>
__global__ __launch_bounds__(kBlock) void ka_reorder(const int *__restrict__ perm, const float4 *__restrict__ x_in,
                                                     const float4 *__restrict__ x1, const float4 *__restrict__ gxs,
                                                     const float4 *__restrict__ gx1, const float4 *__restrict__ gv1,
                                                     const float *__restrict__ gw, float4 *__restrict__ gx_old,
                                                     float4 *__restrict__ gv_old, int n, float dt, float mass_k,
                                                     int stab_passes, SP sp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int o = perm[p];
    const SceneConst &S = sp.at(p);
    const Planes &pl = S.pl;
    f3 xp = xyz(x_in[o]);
    unsigned long long flags = 0ull;
    int step = 0;
    for (int s = 0; s < stab_passes; ++s)
        for (int q = 0; q < pl.n; ++q, ++step) {
            const f3 on = mk3(pl.o[q][0], pl.o[q][1], pl.o[q][2]);
            const f3 nn = mk3(pl.nrm[q][0], pl.nrm[q][1], pl.nrm[q][2]);
            const float d = dot(on - xp, nn) + S.radius;
            if (d <= 0.0f) continue;
            xp += d * nn;
            flags |= 1ull << step;
        }
    const f3 gs = xyz(gxs[p]);
    f3 gx = xyz(gx1[p]);
    for (int s = stab_passes - 1; s >= 0; --s)
        for (int q = pl.n - 1; q >= 0; --q) {
            const int st = s * pl.n + q;
            if (!((flags >> st) & 1ull)) continue;
            const f3 nn = mk3(pl.nrm[q][0], pl.nrm[q][1], pl.nrm[q][2]);
            gx -= dot(nn, gx + gs) * nn;
        }
    gx.y += gw[p] * mass_k * x1[p].w;  // w = e^(k y) / m
    gx += gs;
    const f3 gv = xyz(gv1[p]) + gs * dt;
    gx_old[o] = f4(gx, 0.f);
    gv_old[o] = f4(gv, 0.f);
}

// Per-scene sums of the parameter adjoints: slots [poff[s], poff[s+1]) hold
// scene s's particles at every substep and faces [foff[s], foff[s+1]) its
// faces.  out[6 s + ...] = d/d(wind x, y, z), d/dcd, d/dcl, d/dxsph_c.
__global__ __launch_bounds__(256) void ka_scene_reduce(const float4 *__restrict__ facc, const float *__restrict__ cacc,
                                                       const int *__restrict__ poff, const int *__restrict__ foff,
                                                       float *__restrict__ out) {
    __shared__ double red[6][256];
    const int s = blockIdx.x, t = threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int f = foff[s] + t; f < foff[s + 1]; f += 256) {
        const float4 a = facc[2 * f], b = facc[2 * f + 1];
        acc[0] += a.x;
        acc[1] += a.y;
        acc[2] += a.z;
        acc[3] += a.w;
        acc[4] += b.x;
    }
    for (int p = poff[s] + t; p < poff[s + 1]; p += 256) acc[5] += cacc[p];
    for (int c = 0; c < 6; ++c) red[c][t] = acc[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int c = 0; c < 6; ++c) red[c][t] += red[c][t + w];
        __syncthreads();
    }
    if (t < 6) out[6 * s + t] = (float)red[t][0];
}

// ------------------------------------------------------------ launchers
void launch_adj_inverse(const int *orig, int *map_to_new, int n, hipStream_t st) {
    hipLaunchKernelGGL(ka_inverse, dim3(grid_blocks(n)), dim3(kBlock), 0, st, orig, map_to_new, n);
}
void launch_adj_seed(const float *seed, const int *orig, float4 *g, float scale, int n, hipStream_t st) {
    hipLaunchKernelGGL(ka_seed, dim3(grid_blocks(n)), dim3(kBlock), 0, st, seed, orig, g, scale, n);
}
void launch_adj_unseed(const float4 *g, const int *orig, float *out, float inv_scale, int n, hipStream_t st) {
    hipLaunchKernelGGL(ka_unseed, dim3(grid_blocks(n)), dim3(kBlock), 0, st, g, orig, out, inv_scale, n);
}
template <class SP>
void launch_adj_cells(const float4 *xs, uint32_t *ccell, int n, const SP &sp, hipStream_t st) {
    hipLaunchKernelGGL(ka_cells<SP>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, xs, ccell, n, sp);
}
template <class SP>
void launch_adj_xsph(const float4 *xf, const float4 *x1, const float4 *gx2, const float4 *gv3, const uint32_t *ccell,
                     const int *keys, const int2 *ranges, float4 *gv2_own, float4 *gx_own, float *cacc,
                     unsigned long long *corr, float4 *gxs, float4 *gx1, float *gw, float inv_dt, float eps2, int n,
                     const SP &sp, int block, hipStream_t st) {
    hipLaunchKernelGGL(ka_xsph<SP>, dim3(grid_blocks(n, block)), dim3(block), 0, st, xf, x1, inv_dt, gv3, ccell, keys,
                       ranges, gv2_own, gx_own, cacc, corr, n, sp);
    hipLaunchKernelGGL(ka_xsph_apply<SP>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, xf, x1, gx2, gv2_own, gx_own,
                       corr, gxs, gx1, gw, inv_dt, eps2, n, sp);
}
template <class SP>
void launch_adj_projection(const float4 *xs, const float *m, const float *lambda, float4 *prec, const int *keys,
                           const int2 *ranges, float4 *gxs, float4 *gx_own, float *glam_own,
                           unsigned long long *corr, const uint32_t *trec, uint32_t *srec, int n, const SP &sp,
                           int block, hipStream_t st) {
    hipLaunchKernelGGL(ka_proj_rec<SP>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, xs, m, lambda, prec, n, sp);
    hipLaunchKernelGGL(ka_position<SP>, dim3(grid_blocks(n, block)), dim3(block), 0, st, xs, prec, keys, ranges, gxs,
                       gx_own, glam_own, corr, trec, srec, n, sp);
    hipLaunchKernelGGL(ka_position_apply<SP>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, xs, gx_own, glam_own, corr,
                       gxs, prec, n, sp);
    hipLaunchKernelGGL(ka_lambda<SP>, dim3(grid_blocks(n, block)), dim3(block), 0, st, xs, prec, keys, ranges,
                       gx_own, corr, trec, srec, n, sp);
    hipLaunchKernelGGL(ka_add_own, dim3(grid_blocks(n)), dim3(kBlock), 0, st, gxs, gx_own, corr, n);
}
template <class SP>
void launch_adj_planes(const float4 *xp, const float4 *x1, float4 *gxs, float4 *gx1, int n, const SP &sp,
                       hipStream_t st) {
    hipLaunchKernelGGL(ka_planes<SP>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, xp, x1, gxs, gx1, n, sp);
}
void launch_adj_immovable(float4 *gxs, float4 *gx1, const int *ids, const int *map_to_new, int ni, hipStream_t st) {
    hipLaunchKernelGGL(ka_immovable, dim3(grid_blocks(ni)), dim3(kBlock), 0, st, gxs, gx1, ids, map_to_new, ni);
}
void launch_adj_distance(const float4 *xd, const float4 *x1, const int2 *pairs, const float2 *params,
                         const int *map_to_new, float4 *gxs, float *gw, float4 *cb, float2 *wb, int nc,
                         const int *verts, const int *offs, const int *ents, int nv, hipStream_t st) {
    hipLaunchKernelGGL(ka_distance, dim3(grid_blocks(nc)), dim3(kBlock), 0, st, xd, x1, pairs, params, map_to_new,
                       gxs, cb, wb, nc);
    hipLaunchKernelGGL(ka_distance_gather, dim3(grid_blocks(nv)), dim3(kBlock), 0, st, gxs, gw, verts, offs, ents, cb,
                       wb, map_to_new, nv);
}
template <class SP>
void launch_adj_wind(const float4 *xsw, const float4 *v_in, const int *perm, const float *m, const int4 *faces,
                     const int *map_to_new, float4 *gxs, float4 *gv1, float4 *fbar, float4 *facc, int nf, const SP &sp,
                     float dt, const int *verts, const int *offs, const int *ents, int nv, hipStream_t st) {
    hipLaunchKernelGGL(ka_wind_face<SP>, dim3(grid_blocks(nf)), dim3(kBlock), 0, st, xsw, v_in, perm, m, faces,
                       map_to_new, gxs, fbar, facc, nf, sp, dt);
    hipLaunchKernelGGL(ka_wind_gather, dim3(grid_blocks(nv)), dim3(kBlock), 0, st, gxs, gv1, verts, offs, ents, fbar,
                       map_to_new, nv);
}
template <class SP>
void launch_adj_reorder(const int *perm, const float4 *x_in, const float4 *x1, const float4 *gxs, const float4 *gx1,
                        const float4 *gv1, const float *gw, float4 *gx_old, float4 *gv_old, int n, float dt,
                        float mass_k, int stab_passes, const SP &sp, hipStream_t st) {
    hipLaunchKernelGGL(ka_reorder<SP>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, perm, x_in, x1, gxs, gx1, gv1, gw,
                       gx_old, gv_old, n, dt, mass_k, stab_passes, sp);
}
void launch_adj_scene_reduce(const float4 *facc, const float *cacc, const int *poff, const int *foff, float *out,
                             int nscenes, hipStream_t st) {
    hipLaunchKernelGGL(ka_scene_reduce, dim3(nscenes), dim3(256), 0, st, facc, cacc, poff, foff, out);
}

#define BP_INSTANTIATE_ADJOINT(SP)                                                                              \
    template void launch_adj_cells<SP>(const float4 *, uint32_t *, int, const SP &, hipStream_t);               \
    template void launch_adj_xsph<SP>(const float4 *, const float4 *, const float4 *, const float4 *,           \
                                      const uint32_t *, const int *, const int2 *, float4 *, float4 *, float *,  \
                                      unsigned long long *, float4 *, float4 *, float *, float, float, int,      \
                                      const SP &, int, hipStream_t);                                             \
    template void launch_adj_projection<SP>(const float4 *, const float *, const float *, float4 *, const int *, \
                                            const int2 *, float4 *, float4 *, float *, unsigned long long *,     \
                                            const uint32_t *, uint32_t *, int, const SP &, int, hipStream_t);    \
    template void launch_adj_planes<SP>(const float4 *, const float4 *, float4 *, float4 *, int, const SP &,    \
                                        hipStream_t);                                                           \
    template void launch_adj_wind<SP>(const float4 *, const float4 *, const int *, const float *, const int4 *, \
                                      const int *, float4 *, float4 *, float4 *, float4 *, int, const SP &,      \
                                      float, const int *, const int *, const int *, int, hipStream_t);           \
    template void launch_adj_reorder<SP>(const int *, const float4 *, const float4 *, const float4 *,           \
                                         const float4 *, const float4 *, const float *, float4 *, float4 *, int, \
                                         float, float, int, const SP &, hipStream_t);
BP_INSTANTIATE_ADJOINT(OneScene)
BP_INSTANTIATE_ADJOINT(SceneTable)

}  // namespace bp
