// bp_fluid.hip -- position-based fluid passes on gfx950: density/lambda,
// position correction, vorticity, Akinci normal + tension and XSPH.
// Gather form: each thread owns one sorted slot and writes only its own
// outputs, so there are no floating-point atomics and results are
// deterministic.  Neighbour candidates are the reference's 27 cells around the
// particle's current cell (fluid.cuh:104-115), walked as 9 column runs of
// consecutive slots.
#include "bp_kernels.h"

namespace bp {

// The lambda pass's inputs (LambdaIn) from x* and the masses; with m null,
// only the position planes.
__global__ __launch_bounds__(kBlock) void k_lambda_in(const float4 *__restrict__ xs, const float *__restrict__ m,
                                                      LambdaIn L, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = xs[i];
    if (!m) L.xm = nullptr;
    store_lambda_in(L, i, xyz(a), m ? m[i] : 0.0f, phase_of(a));
}

// fluidLambda (fluid.cuh:78-159).  Also writes the record the position pass
// reads for this particle as self and as partner, (lambda, m, current cell
// packed, 0), so that pass fetches two words-of-four per partner (x* and this)
// instead of four separate arrays; it searches from the same x* and cell.
// Cap: 0 = every lane reads its candidates from global memory; else the
// staged walk with a region of Cap records per wave and column offset,
// screening SB candidates per batch (4 or 8: fewer registers, more waves).
template <class SP, int Cap, int SB>
__global__ __launch_bounds__(kMaxBlock) kPassWaves(SP, Cap ? 5 : 6) void k_lambda(const float4 *__restrict__ xs, const float *__restrict__ m,
                                                   const int2 *__restrict__ ranges, float *__restrict__ lambda,
                                                   float *__restrict__ density, float4 *__restrict__ lam_rec,
                                                   uint32_t *__restrict__ srec, LambdaIn L, int n, SP sp) {
    const int i = pass_slot();
    if (i >= n) return;
    sp.with_scene(i, [&](const SceneConst &S) {
        const GridDesc &G = S.G;
        const SphConst &K = S.K;
        const FluidParams &F = S.F;
        const float4 a = xs[i];
        const f3 pi = xyz(a);
        int g[3];
        G.pos(pi, g);
        const float cell_bits = __uint_as_float(G.pack(g));
        if (phase_of(a) > 0) {  // solids: lambda 0 (the reference never writes theirs)
            lambda[i] = 0.0f;
            density[i] = 0.0f;
            lam_rec[i] = make_float4(0.0f, m[i], cell_bits, 0.0f);
            return;
        }
        float dens = 0.0f, sumg2 = 0.0f;
        f3 gi = mk3(0.f, 0.f, 0.f);
        const float sds = F.solid_density_scale;
        // Q1: only j >= i (fluid.cuh:117).  With a screen record the walk covers
        // every j (the position pass replays it) and j < i are masked off.
        const float inv_rho0 = F.inv_rest_density;
        ScreenRec rec;
        rec.out = srec;
        rec.slot = i;
        rec.n = n;
        rec.planes = L.planes;  // the screen tests distance only
        rec.pstride = L.pstride;
        auto pick = [&](int, float4 b) { return dist2(pi, xyz(b)) <= K.h2; };
        auto body = [&](int, const float4 &b) {
                const float d2 = dist2(pi, xyz(b));
                const float mj = fabsf(b.w);
                const float ld = mj * K.poly6(d2);
                dens += signbit(b.w) ? sds * ld : ld;
                if (F.lambda_mode == 1 && d2 > 0.0f) {
                    // grad C_j = -m_j gradW / rho0 (fluid.cuh:132-137)
                    const float r = fsqrt(d2);
                    const float gc = -mj * K.grad_coef(r, frcp(r)) * inv_rho0;
                    const f3 gr = gc * (pi - xyz(b));
                    sumg2 += gc * gc * d2;
                    gi -= gr;
                }
        };
        auto allow = [&](int base) { return ~chunk_below(i, base); };
        // one record per partner, so four partners per fetch fit the register
        // budget; the partner record is (x*, m), m negated for a non-fluid partner
        if constexpr (Cap != 0) {
            const float4 *const srcs[1] = {L.xm};
            for_each_neighbor_staged<4, 1, Cap, SB>(
                G, ranges, g, srcs, srec ? 0 : i, pick, [&](int j, const StagedRec<1> &R) { body(j, R.r[0]); },
                allow, rec);
        } else {
            for_each_neighbor_nb<4>(
                G, ranges, g, xs, srec ? 0 : i, pick, [&](int j) { return ld32(L.xm, j); }, body, allow, rec);
        }
        if (F.lambda_mode == 1) sumg2 += dot(gi, gi);
        float C = dens / F.rest_density - 1.0f;
        if (F.akinci) C = fmaxf(C, 0.0f);
        const float lam = (F.lambda_mode == 1) ? -C / (sumg2 + F.lambda_relaxation) : -C / F.lambda_denominator;
        lambda[i] = lam;
        density[i] = dens;
        lam_rec[i] = make_float4(lam, m[i], cell_bits, 0.0f);
    });
}

// fluidPosition pair term (fluid.cuh:199-221) of pair (k,j) seen from k, as a
// multiple of diff = x_k - x_j: delta = c diff with c = lambda_sum * gradW / r,
// so |delta| = |c| r and the clamp |delta| <= 5 becomes |c| <= 5 / r.
__device__ __forceinline__ float pair_coef(float d2, float lam_sum, const SphConst &K, const FluidParams &F) {
    float sl = lam_sum;
    if (!F.akinci) sl += -F.scorr_k * ipowf(K.over_wdq(K.poly6(d2)), F.scorr_n);
    const float r = fsqrt(d2);
    const float ir = frcp(r);
    float c = sl * K.grad_coef(r, ir);
    if (fabsf(c) * r >= F.delta_clamp) c = copysignf(F.delta_clamp * ir, c);
    return c;
}

// Fixed-point scale of the pair corrections delivered by atomics: integer
// addition makes their sum independent of arrival order.
constexpr double kCorrScale = 1099511627776.0;  // 2^40
__device__ __forceinline__ unsigned long long to_fixed(float v) {
    return (unsigned long long)(long long)rint((double)v * kCorrScale);
}

// fluidPosition (fluid.cuh:161-228) in gather form, with the reference's pair
// set.  The pair (i<j) exists iff particle i is not a solid (fluid.cuh:176)
// and j's cell (from the grid build) lies in the 27 cells around i's current
// cell.  Slot k sums, as m_j delta(k,j) with delta(k,j) = -delta(j,k):
//   partners j > k found by its own search (k not solid), and
//   partners j < k (not solid) found by its own search whose search contains
//   k's cell.
// A pair k owns whose partner j's search does not reach k's cell is delivered
// to j through corr[j] (fixed point).  The new x* is
//   out = (x* + own sum / rho0) + delivered / rho0,
// the second term only where something was delivered.  With Direct the pass
// writes x* + own sum / rho0 to `out` itself and sets each receiver's byte in
// dflag (a plain store: every writer stores 1), so k_position_fixup reads one
// byte per slot and the corrections only where they exist; otherwise it
// writes the own sum to `out` and k_position_apply adds both parts and runs
// the fused epilogue.
// NB: partners fetched together per round trip (2 or 4; the pass waits on
// these loads, and more in flight per wave costs registers).
template <class SP, bool Flat, int NB, bool Direct>
__global__ __launch_bounds__(kMaxBlock) kPassWaves(SP, Flat ? 5 : 6) void k_position(const float4 *__restrict__ xs,
                                                     const float4 *__restrict__ lam_rec,
                                                     const int *__restrict__ build_keys,
                                                     const int2 *__restrict__ ranges,
                                                     const uint32_t *__restrict__ srec,
                                                     const uint32_t *__restrict__ fluid_stamp,
                                                     const uint32_t *__restrict__ stamp, float4 *__restrict__ out,
                                                     unsigned long long *__restrict__ corr,
                                                     uint8_t *__restrict__ dflag, int n,
                                                     SP sp) {
    const int k = pass_slot();
    if (k >= n) return;
    sp.with_scene(k, [&](const SceneConst &S) {
        const GridDesc &G = S.G;
        const SphConst &K = S.K;
        const FluidParams &F = S.F;
        const float4 a = xs[k];
        const f3 pk = xyz(a);
        const bool kf = phase_of(a) <= 0;
        const float4 rk = lam_rec[k];  // (lambda, m, cell, -) from the lambda pass
        const float lk = rk.x;
        const float mk = rk.y;
        const uint32_t ck = __float_as_uint(rk.z);
        const uint32_t bk = G.pack_addr(build_keys[k]);
        int g[3];
        unpack_cell(ck, g);
        f3 sum = mk3(0.f, 0.f, 0.f);
        auto put = [&]() {
            if constexpr (Direct)
                out[k] = f4(pk + sum * F.inv_rest_density, a.w);
            else
                out[k] = f4(sum, 0.0f);
        };
        // A solid's pairs all have a fluid partner: with no fluid particle in its
        // 27 cells the sum stays zero.
        if (!kf && fluid_stamp && !stamp_near(G, g, fluid_stamp, *stamp)) {
            put();
            return;
        }
        struct Partner {
            float4 b;
            float4 r;  // (lambda, m, cell, -)
        };
        ScreenRec rec;  // the lambda pass screened fluid particles at this x*
        rec.in = kf ? srec : nullptr;
        rec.slot = k;
        rec.n = n;
        walk_form<Flat, NB, kReplayScreen>(
            G, ranges, g, xs, 0,
            // (fluid particles replay the lambda pass's masks; a solid screens for
            // fluid partners only)
            [&](int, float4 b) { return (dist2(pk, xyz(b)) <= K.h2) & (kf | (phase_of(b) <= 0)); },
            [&](int j) { return Partner{ld32(xs, j), ld32(lam_rec, j)}; },
            [&](int j, const Partner &P) {
                const float4 b = P.b;
                // coincident partners are skipped; a solid partner's pair is k's only if j > k
                if (!(dist2(pk, xyz(b)) > 0.0f) | ((phase_of(b) > 0) & !(kf & (j > k)))) return;
                const bool partner_sees_k = G.in_block(__float_as_uint(P.r.z), bk);
                if (j < k && !partner_sees_k) return;
                const f3 diff = pk - xyz(b);
                const float c = pair_coef(dist2(pk, xyz(b)), lk + P.r.x, K, F);
                const float cj = c * P.r.y;
                sum.x = fmaf(cj, diff.x, sum.x);
                sum.y = fmaf(cj, diff.y, sum.y);
                sum.z = fmaf(cj, diff.z, sum.z);
                if (j > k && !partner_sees_k) {  // j will not find this pair: deliver -m_k delta(k,j)
                    const float ck2 = -c * mk;
                    atomicAdd(&corr[3 * j + 0], to_fixed(ck2 * diff.x));
                    atomicAdd(&corr[3 * j + 1], to_fixed(ck2 * diff.y));
                    atomicAdd(&corr[3 * j + 2], to_fixed(ck2 * diff.z));
                    if constexpr (Direct) dflag[j] = 1;
                }
            },
            // partners j != k if k is fluid, only j < k if k is a solid
            [&](int base) { return kf ? ~chunk_bit(k, base) : chunk_below(k, base); }, rec);
        put();
    });
}

// The delivered pairs of a direct position pass: x* += delivered / rho0 for
// each slot whose dflag byte is set, clearing corr and dflag for the next pass.
template <class SP>
__global__ __launch_bounds__(kBlock) void k_position_fixup(float4 *__restrict__ xs_out,
                                                           unsigned long long *__restrict__ corr,
                                                           uint8_t *__restrict__ dflag, int n, SP sp) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || dflag[k] == 0) return;
    dflag[k] = 0;
    const unsigned long long cx = corr[3 * k], cy = corr[3 * k + 1], cz = corr[3 * k + 2];
    if ((cx | cy | cz) == 0ull) return;
    corr[3 * k] = 0ull;
    corr[3 * k + 1] = 0ull;
    corr[3 * k + 2] = 0ull;
    sp.with_scene_ref(k, [&](const SceneConst &S) {
        const float inv = (float)(1.0 / kCorrScale);
        const f3 dv = mk3((float)(long long)cx * inv, (float)(long long)cy * inv, (float)(long long)cz * inv);
        const float4 o = xs_out[k];
        xs_out[k] = f4(xyz(o) + dv * S.F.inv_rest_density, o.w);
    });
}

// x* += own sum / rho0, then += delivered pairs / rho0 where any were
// delivered (particlesolver.cuh:919-932), and the fused epilogue; clears corr
// for the next pass.
template <class SP>
__global__ __launch_bounds__(kBlock) void k_position_apply(const float4 *__restrict__ xs,
                                                           const float4 *__restrict__ sum_in,
                                                           unsigned long long *__restrict__ corr,
                                                           float4 *__restrict__ xs_out, float4 *__restrict__ x_io,
                                                           float4 *__restrict__ v_out, int n, SP sp, Epilogue E) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    sp.with_scene_ref(k, [&](const SceneConst &S) {
        const FluidParams &F = S.F;
        const float4 a = xs[k];
        f3 out = xyz(a) + xyz(sum_in[k]) * F.inv_rest_density;
        const unsigned long long cx = corr[3 * k], cy = corr[3 * k + 1], cz = corr[3 * k + 2];
        if ((cx | cy | cz) != 0ull) {
            const float inv = (float)(1.0 / kCorrScale);
            const f3 dv = mk3((float)(long long)cx * inv, (float)(long long)cy * inv, (float)(long long)cz * inv);
            out = out + dv * F.inv_rest_density;
            corr[3 * k] = 0ull;
            corr[3 * k + 1] = 0ull;
            corr[3 * k + 2] = 0ull;
        }
        if (E.planes_next || E.finish) {
            float4 x4 = x_io[k];
            const f3 xp = xyz(x4);
            if (E.planes_next)
                out = plane_collide_one(out, xp, S.pl, S.radius, S.friction_static, S.friction_dynamic);
            if (E.finish) {  // updateVelocity + updatePositions (particlesolver.cuh:71-81,121-137)
                v_out[k] = f4((out - xp) * E.inv_dt, (E.vw && phase_of(a) < 0) ? E.vw[k] : 0.0f);
                const float d2 = length2(out - xp);
                if (d2 >= E.sleep_eps2 || phase_of(a) < 0) x_io[k] = f4(out, x4.w);
            }
        }
        xs_out[k] = f4(out, a.w);
        if (E.lin.planes) store_lambda_in(E.lin, k, out, E.m[k], phase_of(a));
    });
}

// fluidOmega (fluid.cuh:231-279) and fluidNormal (fluid.cuh:387-434) in one
// sweep (independent sums over the same neighbours).  omega.w = |omega|.
// Mode: bit 0 = omega (vorticity confinement follows), bit 1 = normals
// (Akinci tension follows); one kernel per combination, so a pass carries the
// registers of the terms it computes only.  Cap != 0: candidates and partner
// records from the wave's LDS region of Cap records per column offset
// (for_each_neighbor_staged, 4 waves per SIMD: it needs 111-121 VGPRs);
// Cap = 0: per lane from global memory.  SB: the staged screen batch (4 or 8).
template <class SP, int Mode, int Cap, int SB>
__global__ __launch_bounds__(kMaxBlock) kPassWaves(SP, Cap ? 4 : 6) void k_post_a(const float4 *__restrict__ xs, const float4 *__restrict__ v,
                                                   const float *__restrict__ density,
                                                   const int2 *__restrict__ ranges, float4 *__restrict__ omega,
                                                   float4 *__restrict__ xom,
                                                   float4 *__restrict__ normal, uint32_t *__restrict__ ccell,
                                                   uint32_t *__restrict__ srec, const float *__restrict__ planes,
                                                   int pstride, int n, SP sp) {
    constexpr bool do_omega = (Mode & 1) != 0, do_normal = (Mode & 2) != 0;
    const int i = pass_slot();
    if (i >= n) return;
    sp.with_scene(i, [&](const SceneConst &S) {
        const GridDesc &G = S.G;
        const SphConst &K = S.K;
        const float4 a = xs[i];
        const int ph = phase_of(a);
        if (ph >= 0) return;
        const f3 pi = xyz(a);
        int g[3];
        G.pos(pi, g);
        ccell[i] = G.pack(g);
        // v.w holds the density (written with v for this pass), so a
        // partner's velocity and density are one fetch
        const f3 vi = xyz(v[i]);
        f3 om = mk3(0.f, 0.f, 0.f), nn = mk3(0.f, 0.f, 0.f);
        ScreenRec rec;  // replayed by the vorticity/tension and XSPH passes
        rec.out = srec;
        rec.slot = i;
        rec.n = n;
        // with position planes every particle has phase ph: the screen tests distance only
        rec.planes = planes;
        rec.pstride = pstride;
        const bool any_phase = planes != nullptr;
        // partner records: x*, (v, density)
        auto pick = [&](int, float4 b) { return (any_phase | (phase_of(b) == ph)) & (dist2(pi, xyz(b)) <= K.h2); };
        auto body = [&](int, const StagedRec<2> &P) {
            const float4 b = P.r[0];
            if (!(dist2(pi, xyz(b)) > 0.0f)) return;  // coincident partners are skipped
            const f3 diff = pi - xyz(b);
            const float r = fsqrt(dist2(pi, xyz(b)));
            const f3 gr = K.grad_coef(r, frcp(r)) * diff;
            if (do_omega) om += cross(xyz(P.r[1]) - vi, gr);
            if (do_normal) nn += frcp(P.r[1].w) * gr;
        };
        if constexpr (Cap != 0) {
            const float4 *const srcs[2] = {xs, v};
            for_each_neighbor_staged<kBodyBatch, 2, Cap, SB>(G, ranges, g, srcs, 0, pick, body, AllowNotSelf{i}, rec);
        } else {
            for_each_neighbor_nb<kBodyBatch>(
                G, ranges, g, xs, 0, pick, [&](int j) { return StagedRec<2>{{ld32(xs, j), ld32(v, j)}}; }, body,
                AllowNotSelf{i}, rec);
        }
        // (x, |omega|) and (normal, density) are the two records the
        // vorticity/tension pass fetches per partner
        const float mag = do_omega ? fsqrt(dot(om, om)) : 0.0f;
        if (do_omega) omega[i] = f4(om, mag);
        xom[i] = f4(pi, mag);
        if (do_normal) normal[i] = f4(K.h * nn, density[i]);
    });
}

// fluidVorticity (fluid.cuh:281-336, in place) followed by fluidAkinciTension
// (fluid.cuh:436-489) for the same particle: both only change the particle's
// own velocity, so the fused pass writes v in place.
// Mode: bit 0 = vorticity confinement, bit 1 = Akinci tension (one kernel per
// combination, as for k_post_a).
template <class SP, bool Flat, int Mode>
__global__ __launch_bounds__(kMaxBlock) kPassWaves(SP, 1) void k_post_b(const float4 *__restrict__ xs, float4 *__restrict__ v,
                                                   const float4 *__restrict__ omega,
                                                   const float4 *__restrict__ xom,
                                                   const float4 *__restrict__ normal,
                                                   const float *__restrict__ density,
                                                   const int2 *__restrict__ ranges,
                                                   const uint32_t *__restrict__ ccell,
                                                   const uint32_t *__restrict__ srec, int n, SP sp, float dt) {
    constexpr bool do_vort = (Mode & 1) != 0, do_akinci = (Mode & 2) != 0;
    const int i = pass_slot();
    if (i >= n) return;
    sp.with_scene(i, [&](const SceneConst &S) {
        const GridDesc &G = S.G;
        const SphConst &K = S.K;
        const FluidParams &F = S.F;
        const float4 a = xs[i];
        const int ph = phase_of(a);
        if (ph >= 0) return;
        const f3 pi = xyz(a);
        int g[3];
        unpack_cell(ccell[i], g);
        f3 eta = mk3(0.f, 0.f, 0.f), f = mk3(0.f, 0.f, 0.f);
        const f3 ni = do_akinci ? xyz(normal[i]) : mk3(0.f, 0.f, 0.f);
        const float rhoi = do_akinci ? density[i] : 0.0f;
        const float gam = F.tension_gamma;
        ScreenRec rec;  // the omega/normal pass screened the same positions
        rec.in = srec;
        rec.slot = i;
        rec.n = n;
        struct Partner {
            float4 xo;   // (x, |omega|) of the partner
            float4 nrm;  // nrm.w: the partner's density
        };
        // one 16-byte fetch of the second record whatever is used of it (a
        // conditional fetch is split into per-component loads)
        const float4 *__restrict__ second = do_akinci ? normal : xom;
        // four partners per round trip: the pass mostly waits on these loads
        // (batch 0.79 -> 0.72 ms, 4M block 0.95 -> 0.90 ms against two)
        walk_form<Flat, 4, kReplayScreen>(
            G, ranges, g, xs, 0,
            [&](int, float4 b) { return (phase_of(b) == ph) & (dist2(pi, xyz(b)) <= K.h2); },
            [&](int j) { return Partner{ld32(xom, j), ld32(second, j)}; },
            [&](int, const Partner &P) {
                const float4 b = P.xo;
                if (!(dist2(pi, xyz(b)) > 0.0f)) return;  // coincident partners are skipped
                const f3 diff = pi - xyz(b);
                const float r = fsqrt(dist2(pi, xyz(b)));
                const float ir = frcp(r);
                if (do_vort) eta += (b.w * K.grad_coef(r, ir)) * diff;
                if (do_akinci) {
                    // k_ij (-gamma C(r) diff / r - gamma (n_i - n_j)) (fluid.cuh:476-483)
                    const float kij = 2.0f * F.rest_density * frcp(rhoi + P.nrm.w);
                    const f3 fcoh = (-gam * K.akinci(r) * ir) * diff;
                    const f3 fcur = -gam * (ni - xyz(P.nrm));
                    f += kij * (fcoh + fcur);
                }
            },
            AllowNotSelf{i}, rec);
        f3 vi = xyz(v[i]);
        if (do_vort && length2(eta) > 1e-3f) {
            const f3 nrm = eta * (1.0f / sqrtf(dot(eta, eta)));
            vi += F.vorticity_eps * cross(nrm, xyz(omega[i])) * dt;
        }
        if (do_akinci) vi = vi + f * dt;
        v[i] = f4(vi, 0.0f);
    });
}

// fluidXSph (fluid.cuh:338-385).  Non-fluid particles keep their velocity.
template <class SP>
__global__ __launch_bounds__(kMaxBlock) kPassWaves(SP, 6) void k_xsph(const float4 *__restrict__ xs, const float4 *__restrict__ v,
                                                 const int2 *__restrict__ ranges,
                                                 const uint32_t *__restrict__ ccell,
                                                 const uint32_t *__restrict__ srec, float4 *__restrict__ v_out,
                                                 int n, SP sp, int use_ccell) {
    const int i = pass_slot();
    if (i >= n) return;
    sp.with_scene(i, [&](const SceneConst &S) {
        const GridDesc &G = S.G;
        const SphConst &K = S.K;
        const float c = S.F.xsph_c;
        const float4 a = xs[i];
        const int ph = phase_of(a);
        const float4 v4 = v[i];
        if (ph >= 0) {
            v_out[i] = v4;
            return;
        }
        const f3 pi = xyz(a);
        int g[3];
        if (use_ccell)
            unpack_cell(ccell[i], g);
        else
            G.pos(pi, g);
        const f3 vi = xyz(v4);
        f3 vn = mk3(0.f, 0.f, 0.f);
        ScreenRec rec;  // after the omega/normal pass: its screen of the same positions
        rec.in = use_ccell ? srec : nullptr;
        rec.slot = i;
        rec.n = n;
        struct Partner {
            float4 b, v;
        };
        for_each_neighbor_nb<kBodyBatch, kReplayScreen>(
            G, ranges, g, xs, 0,
            [&](int, float4 b) { return (phase_of(b) == ph) & (dist2(pi, xyz(b)) <= K.h2); },
            [&](int j) { return Partner{ld32(xs, j), ld32(v, j)}; },
            [&](int, const Partner &P) { vn += (xyz(P.v) - vi) * K.poly6(dist2(pi, xyz(P.b))); }, AllowNotSelf{i},
            rec);
        v_out[i] = f4(vi + c * vn, 0.0f);
    });
}

// ------------------------------------------------------------ launchers
template <class SP>
void launch_lambda(const float4 *xs, const float *m, const int2 *ranges, float *lambda, float *density,
                   float4 *lam_rec, uint32_t *srec, const LambdaIn &L, int lin_ready, int n, const SP &sp,
                   int block, int staged, int stage_sb, hipStream_t st) {
    if (!lin_ready) hipLaunchKernelGGL(k_lambda_in, dim3(grid_blocks(n)), dim3(kBlock), 0, st, xs, m, L, n);
    const dim3 grid(grid_blocks(n, block)), blk(block);
    const size_t lds = staged ? staged_lds_bytes(block, 1, staged) : 0;
#define BP_LAMBDA(C, B)                                                                                          \
    hipLaunchKernelGGL((k_lambda<SP, C, B>), grid, blk, lds, st, xs, m, ranges, lambda, density, lam_rec, srec, L, \
                       n, sp)
    switch (staged | (stage_sb == 4 ? 1 : 0)) {
    case 128: BP_LAMBDA(128, 8); break;
    case 256: BP_LAMBDA(256, 8); break;
    case 512: BP_LAMBDA(512, 8); break;
    case 129: BP_LAMBDA(128, 4); break;
    case 257: BP_LAMBDA(256, 4); break;
    case 513: BP_LAMBDA(512, 4); break;
    default: BP_LAMBDA(0, 8); break;
    }
#undef BP_LAMBDA
}
template <class SP>
void launch_position(const float4 *xs, const float4 *lam_rec, const int *build_keys, const int2 *ranges,
                     const uint32_t *srec, const uint32_t *fluid_stamp, const uint32_t *stamp, float4 *sum_tmp,
                     unsigned long long *corr, uint8_t *dflag, float4 *xs_out, float4 *x_io,
                     float4 *v_out, int n, const SP &sp, const Epilogue &E, int block, int flat, int partners,
                     hipStream_t st) {
    const dim3 grid(grid_blocks(n, block)), blk(block);
    const size_t lds = flat ? flat_lds_bytes(block) : 0;
    // no epilogue: the pass writes x* itself and a fix-up adds the deliveries
    const bool direct = !E.planes_next && !E.finish && !E.lin.planes;
    float4 *const out = direct ? xs_out : sum_tmp;
#define BP_POSITION(F, NB, D)                                                                                   \
    hipLaunchKernelGGL((k_position<SP, F, NB, D>), grid, blk, lds, st, xs, lam_rec, build_keys, ranges, srec, \
                       fluid_stamp, stamp, out, corr, dflag, n, sp)
    switch ((flat ? 1 : 0) | (partners == 4 ? 2 : 0) | (direct ? 4 : 0)) {
    case 0: BP_POSITION(false, 2, false); break;
    case 1: BP_POSITION(true, 2, false); break;
    case 2: BP_POSITION(false, 4, false); break;
    case 3: BP_POSITION(true, 4, false); break;
    case 4: BP_POSITION(false, 2, true); break;
    case 5: BP_POSITION(true, 2, true); break;
    case 6: BP_POSITION(false, 4, true); break;
    default: BP_POSITION(true, 4, true); break;
    }
#undef BP_POSITION
    if (direct) {
        hipLaunchKernelGGL(k_position_fixup<SP>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, xs_out, corr, dflag, n,
                           sp);
    } else {
        hipLaunchKernelGGL(k_position_apply<SP>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, xs, sum_tmp, corr,
                           xs_out, x_io, v_out, n, sp, E);
    }
}
template <class SP>
void launch_post_a(const float4 *xs, const float4 *v, const float *density, const int2 *ranges, float4 *omega,
                   float4 *xom, float4 *normal, uint32_t *ccell, uint32_t *srec, const LambdaIn *one_phase,
                   int planes_ready, int n, const SP &sp, int do_omega, int do_normal, int block, int staged,
                   int stage_sb, hipStream_t st) {
    if (one_phase && !planes_ready)
        hipLaunchKernelGGL(k_lambda_in, dim3(grid_blocks(n)), dim3(kBlock), 0, st, xs, nullptr, *one_phase, n);
    const float *planes = one_phase ? one_phase->planes : nullptr;
    const int pstride = one_phase ? one_phase->pstride : 0;
    const dim3 grid(grid_blocks(n, block)), blk(block);
    const size_t lds = staged ? staged_lds_bytes(block, 2, staged) : 0;
#define BP_POST_A(M, C, B)                                                                                       \
    hipLaunchKernelGGL((k_post_a<SP, M, C, B>), grid, blk, lds, st, xs, v, density, ranges, omega, xom, normal,    \
                       ccell, srec, planes, pstride, n, sp)
#define BP_POST_A_CAP(M)                                \
    switch (staged | (stage_sb == 4 ? 1 : 0)) {         \
    case 128: BP_POST_A(M, 128, 8); break;              \
    case 256: BP_POST_A(M, 256, 8); break;              \
    case 512: BP_POST_A(M, 512, 8); break;              \
    case 129: BP_POST_A(M, 128, 4); break;              \
    case 257: BP_POST_A(M, 256, 4); break;              \
    case 513: BP_POST_A(M, 512, 4); break;              \
    default: BP_POST_A(M, 0, 8); break;                 \
    }
    switch ((do_omega ? 1 : 0) | (do_normal ? 2 : 0)) {
    case 1: BP_POST_A_CAP(1); break;
    case 2: BP_POST_A_CAP(2); break;
    case 3: BP_POST_A_CAP(3); break;
    default: break;
    }
#undef BP_POST_A_CAP
#undef BP_POST_A
}
template <class SP>
void launch_post_b(const float4 *xs, float4 *v, const float4 *omega, const float4 *xom, const float4 *normal,
                   const float *density, const int2 *ranges, const uint32_t *ccell, const uint32_t *srec, int n,
                   const SP &sp, int do_vort, int do_akinci, float dt, int block, int flat, hipStream_t st) {
    const dim3 grid(grid_blocks(n, block)), blk(block);
    const size_t lds = flat ? flat_lds_bytes(block) : 0;
#define BP_POST_B(F, M)                                                                                           \
    hipLaunchKernelGGL((k_post_b<SP, F, M>), grid, blk, lds, st, xs, v, omega, xom, normal, density, ranges, ccell, \
                       srec, n, sp, dt)
    switch ((do_vort ? 1 : 0) | (do_akinci ? 2 : 0) | (flat ? 4 : 0)) {
    case 1: BP_POST_B(false, 1); break;
    case 2: BP_POST_B(false, 2); break;
    case 3: BP_POST_B(false, 3); break;
    case 5: BP_POST_B(true, 1); break;
    case 6: BP_POST_B(true, 2); break;
    case 7: BP_POST_B(true, 3); break;
    default: break;
    }
#undef BP_POST_B
}
template <class SP>
void launch_xsph(const float4 *xs, const float4 *v, const int2 *ranges, const uint32_t *ccell, const uint32_t *srec,
                 float4 *v_out, int n, const SP &sp, int use_ccell, int block, hipStream_t st) {
    hipLaunchKernelGGL(k_xsph<SP>, dim3(grid_blocks(n, block)), dim3(block), 0, st, xs, v, ranges, ccell, srec,
                       v_out, n, sp, use_ccell);
}

#define BP_INSTANTIATE_FLUID(SP)                                                                                \
    template void launch_lambda<SP>(const float4 *, const float *, const int2 *, float *, float *, float4 *,    \
                                    uint32_t *, const LambdaIn &, int, int, const SP &, int, int, int,          \
                                    hipStream_t);                                                               \
    template void launch_position<SP>(const float4 *, const float4 *, const int *, const int2 *,                \
                                      const uint32_t *, const uint32_t *, const uint32_t *, float4 *,           \
                                      unsigned long long *, uint8_t *, float4 *, float4 *, float4 *,            \
                                      int, const SP &, const Epilogue &, int, int, int, hipStream_t);           \
    template void launch_post_a<SP>(const float4 *, const float4 *, const float *, const int2 *, float4 *,      \
                                    float4 *, float4 *, uint32_t *, uint32_t *, const LambdaIn *, int, int,     \
                                    const SP &, int, int, int, int, int, hipStream_t);                          \
    template void launch_post_b<SP>(const float4 *, float4 *, const float4 *, const float4 *, const float4 *,   \
                                    const float *, const int2 *, const uint32_t *, const uint32_t *, int,       \
                                    const SP &, int, int, float, int, int, hipStream_t);                        \
    template void launch_xsph<SP>(const float4 *, const float4 *, const int2 *, const uint32_t *,               \
                                  const uint32_t *, float4 *, int, const SP &, int, int, hipStream_t);
BP_INSTANTIATE_FLUID(OneScene)
BP_INSTANTIATE_FLUID(SceneTable)

}  // namespace bp
