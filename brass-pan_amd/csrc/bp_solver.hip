// bp_solver.hip -- C ABI (include/brasspan.h) and substep orchestration.
//
// Replaces struct ParticleSolver (brasstower/particlesolver.cuh:146-1148).
// One handle owns all device memory of one scene and one HIP stream; the
// substep pipeline is ParticleSolver::update (:772-1072) with the projection
// and wind blocks active, fused into fewer passes:
//   predict+hash(+reset) | radix sort | ranges | reorder(+inverse map, inverse
//   scaled mass, stabilisation, first plane collisions) | [wind] |
//   iterations { [planes] [particle-particle] lambda  position(+next planes |
//   +velocity/position update) [distance] [bending] [shape matching] [pins] } |
//   [finish] | [omega+normal] [vorticity+tension] [xsph]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bp_guard.h"
#include "bp_solver_impl.h"

using namespace bp;
using namespace bp_host;

namespace bp_host {

thread_local std::string g_last_error;

bp_status fail(bp_status st, const std::string &msg) {
    g_last_error = msg;
    return st;
}

}  // namespace bp_host

namespace {

const char *kFamilyNames[F_COUNT] = {"predict_hash", "radix_sort", "cell_ranges", "reorder", "wind",
                                     "plane_collide", "particle_particle", "fluid_lambda", "fluid_position",
                                     "distance", "bending", "shape_matching", "immovable", "finish",
                                     "omega_normal", "vorticity_tension", "xsph", "pick"};

}  // namespace

namespace {

bp_status build_csr(bp_solver *s, Csr &c, int nent_per, const int *ids, int count, int N, int div) {
    // entries (item * nent_per + corner) / div, grouped by particle id, in item order
    std::vector<int> cnt(N, 0);
    for (int i = 0; i < count * nent_per; ++i) cnt[ids[i]]++;
    std::vector<int> verts, offs;
    std::vector<int> pos(N, -1);
    for (int p = 0; p < N; ++p)
        if (cnt[p] > 0) {
            pos[p] = (int)verts.size();
            verts.push_back(p);
        }
    offs.assign(verts.size() + 1, 0);
    for (size_t t = 0; t < verts.size(); ++t) offs[t + 1] = offs[t] + cnt[verts[t]];
    std::vector<int> fill(offs.begin(), offs.end() - 1), ents((size_t)count * nent_per);
    for (int i = 0; i < count * nent_per; ++i) ents[fill[pos[ids[i]]]++] = i / div;
    c.nv = (int)verts.size();
    bp_status st;
    if ((st = s->upload(&c.verts, verts.data(), verts.size())) != BP_OK) return st;
    if ((st = s->upload(&c.offs, offs.data(), offs.size())) != BP_OK) return st;
    if ((st = s->upload(&c.ents, ents.data(), ents.size())) != BP_OK) return st;
    return BP_OK;
}

void set_kernel_const(SphConst &K, float h, float scorr_dq) {
    // SetKernelRadius (fluid.cuh:11-27), host powf
    K.c1 = 315.f / 64.f / 3.141592f / powf(h, 9.0f);
    K.c2 = -45.f / 3.141592f / powf(h, 6.f);
    K.c3 = 32.0f / 3.141592f / powf(h, 9.f);
    K.c4 = powf(h, 6.0f) / 64.0f;
    K.h2 = h * h;
    K.halfh = h * 0.5f;
    K.h = h;
    // poly6Kernel(powf(0.03f * KernelRadius, 2.f)) (fluid.cuh:208)
    const float r2 = powf(scorr_dq * h, 2.0f);
    const float t = K.h2 - r2;
    K.wdq = K.c1 * t * t * t;
    K.inv_wdq = (K.wdq >= 0x1p-20f && K.wdq <= 0x1p20f) ? 1.0f / K.wdq : 0.0f;
}

// per-scene constants from the scene description and its parameters
void set_scene_const(SceneConst &c, const bp_params &P, float h) {
    set_kernel_const(c.K, h, P.scorr_dq);
    c.F.rest_density = P.fluid_rest_density;
    c.F.inv_rest_density = 1.0f / P.fluid_rest_density;
    c.F.solid_density_scale = P.solid_density_scale;
    c.F.lambda_mode = P.lambda_mode;
    c.F.lambda_denominator = P.lambda_denominator;
    c.F.lambda_relaxation = P.lambda_relaxation;
    c.F.akinci = P.akinci;
    c.F.scorr_k = P.scorr_k;
    c.F.scorr_n = P.scorr_n;
    c.F.delta_clamp = P.delta_clamp;
    c.F.vorticity_eps = P.vorticity_eps;
    c.F.tension_gamma = P.tension_gamma;
    c.F.xsph_c = P.xsph_c;
    c.friction_static = P.friction_static;
    c.friction_dynamic = P.friction_dynamic;
    c.wind = mk3(P.wind[0], P.wind[1], P.wind[2]);
    c.drag_coeff = P.drag_coeff;
    c.lift_coeff = P.lift_coeff;
}

bp_status check_params(const bp_params *p) {
    if (!p) return fail(BP_ERR_INVALID_ARG, "params is NULL");
    if (p->iterations < 0 || p->stabilization_passes < 0)
        return fail(BP_ERR_INVALID_ARG, "iterations and stabilization_passes must be >= 0");
    if (!(p->fluid_rest_density > 0.0f)) return fail(BP_ERR_INVALID_ARG, "fluid_rest_density must be > 0");
    if (p->lambda_mode != 0 && p->lambda_mode != 1) return fail(BP_ERR_INVALID_ARG, "lambda_mode must be 0 or 1");
    if (p->shape_matching_iterations < 0) return fail(BP_ERR_INVALID_ARG, "shape_matching_iterations < 0");
    if (p->scorr_n < 0 || p->scorr_n > 64) return fail(BP_ERR_INVALID_ARG, "scorr_n must be in [0, 64]");
    return BP_OK;
}

// Parameters that shape the pass sequence must agree across the scenes of a batch.
bool same_schedule(const bp_params &a, const bp_params &b) {
    return a.iterations == b.iterations && a.stabilization_passes == b.stabilization_passes &&
           a.gravity[0] == b.gravity[0] && a.gravity[1] == b.gravity[1] && a.gravity[2] == b.gravity[2] &&
           a.akinci == b.akinci && (a.vorticity_eps != 0.0f) == (b.vorticity_eps != 0.0f) &&
           (a.xsph_c != 0.0f) == (b.xsph_c != 0.0f) && a.mass_scaling_k == b.mass_scaling_k &&
           a.sleep_epsilon == b.sleep_epsilon && a.wind_enabled == b.wind_enabled &&
           a.bend_stiffness == b.bend_stiffness && a.shape_matching_iterations == b.shape_matching_iterations &&
           a.nan_guard == b.nan_guard;
}

}  // namespace

namespace bp_host {

namespace {
// device-to-device snapshot of `count` elements on the solver's stream
template <class T> hipError_t snap(T *dst, const T *src, size_t count, hipStream_t st) {
    if (!dst) return hipSuccess;
    return hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToDevice, st);
}
}  // namespace

// one substep of ParticleSolver::update (particlesolver.cuh:787-1063); `sp()`
// returns the scene parameter source for the current slot order
template <class SPF>
bp_status substep_t(bp_solver *s, float subdt, float inv_dt, f3 gstep, bool pick, int picked, f3 ppos, SPF sp,
                    bool no_fuse, const SubstepHooks *hk) {
    const int N = s->N;
    const bp_params &P = s->P;
    hipStream_t st = s->st;
    const f3 zero = mk3(0.f, 0.f, 0.f);
    const bool wind_on = P.wind_enabled && s->nf > 0;
    const bool tail_groups = s->nd > 0 || s->nb > 0 || s->nr > 0 || s->ni > 0;
    const bool need_w = s->has_solid || s->nd > 0 || s->nb > 0;
    const int iters = P.iterations;
    const SubstepHooks H = hk ? *hk : SubstepHooks{};

    if (pick) TIMED(F_PICK, launch_pick(s->x.cur(), s->v.cur(), s->map_to_new, picked, ppos, 0, zero, st));
    HIP_TRY(snap(H.x_in, s->x.cur(), N, st));
    HIP_TRY(snap(H.v_in, s->v.cur(), N, st));

    // predict + hash, clearing the ranges of the previous substep
    int kin_idx = (s->sk_idx == 0) ? 1 : 0;
    int *kin = s->keys[kin_idx], *kother = s->keys[1 - kin_idx];
    TIMED(F_PREDICT, launch_predict_hash(s->x.cur(), s->v.cur(), kin, N, subdt, gstep, sp(),
                                         s->sk_idx >= 0 ? s->sorted_keys : nullptr, s->ranges, st));
    int *sk = nullptr, *pm = nullptr;
    TIMED(F_SORT, launch_radix_sort(kin, kother, s->vals[0], s->vals[1], s->counts, s->totals, N, s->sort_bits, &sk,
                                    &pm, st));
    s->sorted_keys = sk;
    s->perm = pm;
    s->sk_idx = (sk == s->keys[0]) ? 0 : 1;
    TIMED(F_RANGES, launch_find_ranges(sk, N, s->ranges, st));
    HIP_TRY(snap(H.keys, sk, N, st));
    HIP_TRY(snap(H.perm, pm, N, st));

    const bool fuse_first = !no_fuse && s->any_planes && !wind_on && !pick && iters > 0;
    // The lambda pass reads LambdaIn besides x*.  The pass that writes x* right
    // before it also writes LambdaIn when no other pass moves x* in between
    // (fluid-only scenes); otherwise launch_lambda derives it from x* first.
    const bool lone_fluid = s->lin.planes && !tail_groups && !s->has_solid && !pick;
    const bool planes_first = lone_fluid && !wind_on && (fuse_first || !s->any_planes);
    TIMED(F_REORDER, launch_reorder(pm, s->x.cur(), s->v.cur(), s->m.cur(), s->ph.cur(), s->orig.cur(), s->sid.cur(),
                                    s->xs.cur(), s->x.alt(), s->v.alt(), s->m.alt(), s->ph.alt(), s->orig.alt(),
                                    s->sid.alt(), s->map_to_new, N, subdt, gstep, P.mass_scaling_k, need_w ? 1 : 0,
                                    P.stabilization_passes, fuse_first ? 1 : 0,
                                    planes_first ? s->lin : LambdaIn{}, sp(), st));
    s->x.flip();
    s->v.flip();
    s->m.flip();
    s->ph.flip();
    s->orig.flip();
    if (s->sid.b[0]) s->sid.flip();
    if (s->cell_stamp)
        TIMED(F_PP, launch_mark_cells(sk, s->ph.cur(), N, s->stamp, s->cell_stamp, s->fluid_stamp, sp(), st));
    HIP_TRY(snap(H.x1, s->x.cur(), N, st));
    HIP_TRY(snap(H.xsw, s->xs.cur(), N, st));
    HIP_TRY(snap(H.m, s->m.cur(), N, st));
    HIP_TRY(snap(H.orig, s->orig.cur(), N, st));
    if (s->sid.b[0]) HIP_TRY(snap(H.sid, s->sid.cur(), N, st));

    if (wind_on)
        TIMED(F_WIND, launch_wind(s->xs.cur(), s->v.cur(), s->m.cur(), s->faces, s->map_to_new, s->fbuf, s->nf, sp(),
                                  subdt, s->fcsr.verts, s->fcsr.offs, s->fcsr.ents, s->fcsr.nv, st));

    // the omega/normal pass reads a fluid partner's density from v.w
    const float *vw = (s->has_fluid && (P.vorticity_eps != 0.0f || P.akinci)) ? s->density : nullptr;
    bool planes_done = fuse_first;
    bool planes_ready = planes_first;
    bool finished = false;
    for (int it = 0; it < iters; ++it) {
        if (pick) TIMED(F_PICK, launch_pick(s->xs.cur(), s->v.cur(), s->map_to_new, picked, ppos, 0, zero, st));
        if (H.xp) HIP_TRY(snap(H.xp[it], s->xs.cur(), N, st));
        if (s->any_planes && !planes_done) TIMED(F_PLANES, launch_plane_collide(s->xs.cur(), s->x.cur(), N, sp(), st));
        planes_done = false;
        if (s->has_solid) {
            // the last pass to move x* before the lambda pass: it also writes LambdaIn
            TIMED(F_PP, launch_particle_particle(s->xs.cur(), s->x.cur(), s->ranges, s->xs.alt(), N, sp(), s->blk_pp,
                                                 H.contacts, s->cell_stamp, s->stamp, s->lin, s->m.cur(), st));
            s->xs.flip();
            planes_ready = s->lin.planes != nullptr;
        }
        if (H.xl) HIP_TRY(snap(H.xl[it], s->xs.cur(), N, st));
        if (s->has_fluid) {
            TIMED(F_LAMBDA, launch_lambda(s->xs.cur(), s->m.cur(), s->ranges, s->lambda, s->density, s->lam_rec,
                                          H.rec ? H.rec[it] : s->srec, s->lin, planes_ready ? 1 : 0, N, sp(),
                                          s->blk_lambda, s->staged_lambda, s->stage_sb_lambda, st));
            if (H.lam) HIP_TRY(snap(H.lam[it], s->lambda, N, st));
            Epilogue E{};
            E.planes_next = (!no_fuse && s->any_planes && it + 1 < iters && !tail_groups && !pick) ? 1 : 0;
            E.finish = (!no_fuse && it + 1 == iters && !tail_groups) ? 1 : 0;
            E.inv_dt = inv_dt;
            E.sleep_eps2 = P.sleep_epsilon * P.sleep_epsilon;
            E.vw = vw;
            // (after the last iteration: the final x*, which the omega/normal pass screens)
            planes_ready = lone_fluid && (it + 1 < iters ? (E.planes_next || !s->any_planes) : E.finish != 0);
            if (planes_ready) {
                E.lin = s->lin;
                E.m = s->m.cur();
            }
            TIMED(F_POSITION, launch_position(s->xs.cur(), s->lam_rec, sk, s->ranges, H.rec ? H.rec[it] : s->srec,
                                              s->fluid_stamp, s->stamp, s->sum_tmp, s->corr, s->dflag,
                                              s->xs.alt(), s->x.cur(), s->v.cur(), N, sp(), E, s->blk_position,
                                              s->flat_position, s->pos_partners, st));
            s->xs.flip();
            planes_done = E.planes_next;
            finished = E.finish;
        }
        if (H.xd) HIP_TRY(snap(H.xd[it], s->xs.cur(), N, st));
        if (s->nd > 0)
            TIMED(F_DISTANCE, launch_distance(s->xs.cur(), s->x.cur(), s->dpairs, s->dparams, s->map_to_new, s->kbuf,
                                              s->nd, s->dcsr.verts, s->dcsr.offs, s->dcsr.ents, s->dcsr.nv, st));
        if (s->nb > 0)
            TIMED(F_BENDING, launch_bending(s->xs.cur(), s->x.cur(), s->bends, s->map_to_new, s->bbuf, s->nb,
                                            P.bend_stiffness, s->bcsr.verts, s->bcsr.offs, s->bcsr.ents, s->bcsr.nv,
                                            st));
        // pins apply after shape matching; in the same launch when they touch no rigid body
        const int fused_pins = s->pins_apart_from_rigid ? s->ni : 0;
        if (s->nr > 0)
            TIMED(F_SHAPE, launch_shape_matching(s->xs.cur(), s->map_to_new, s->rrest, s->rranges, s->rrot, s->rcm,
                                                 s->nr, P.shape_matching_iterations, st, s->x.cur(), s->immov,
                                                 fused_pins));
        if (s->ni > 0 && !(s->nr > 0 && fused_pins > 0))
            TIMED(F_IMMOVABLE, launch_immovable(s->xs.cur(), s->x.cur(), s->immov, s->map_to_new, s->ni, st));
    }
    HIP_TRY(snap(H.xf, s->xs.cur(), N, st));
    if (!finished)
        TIMED(F_FINISH, launch_finish(s->xs.cur(), s->x.cur(), s->v.cur(), vw, N, inv_dt,
                                      P.sleep_epsilon * P.sleep_epsilon, st));

    if (s->has_fluid) {
        const int do_vort = P.vorticity_eps != 0.0f ? 1 : 0;
        const int do_ak = P.akinci ? 1 : 0;
        if (do_vort || do_ak) {
            TIMED(F_POST_A, launch_post_a(s->xs.cur(), s->v.cur(), s->density, s->ranges, s->omega, s->xom, s->normal,
                                          s->ccell, s->srec, s->one_phase ? &s->lin : nullptr,
                                          finished && planes_ready ? 1 : 0, N, sp(), do_vort, do_ak, s->blk_post_a,
                                          s->staged_post_a, s->stage_sb_post_a, st));
            TIMED(F_POST_B, launch_post_b(s->xs.cur(), s->v.cur(), s->omega, s->xom, s->normal, s->density, s->ranges,
                                          s->ccell, s->srec, N, sp(), do_vort, do_ak, subdt, s->blk_post_b,
                                          s->flat_post_b, st));
        }
        if (P.xsph_c != 0.0f) {
            TIMED(F_XSPH, launch_xsph(s->xs.cur(), s->v.cur(), s->ranges, s->ccell, s->srec, s->v.alt(), N, sp(),
                                      (do_vort || do_ak) ? 1 : 0, s->blk_xsph, st));
            s->v.flip();
        }
    }
    HIP_TRY(hipGetLastError());
    return BP_OK;
}

bp_status substep(bp_solver *s, float subdt, float inv_dt, f3 gstep, bool pick, int picked, f3 ppos, bool no_fuse,
                  const SubstepHooks *hk) {
    if (s->nscenes == 1) {
        const OneScene one{s->hsc[0]};
        return substep_t(s, subdt, inv_dt, gstep, pick, picked, ppos, [&]() { return one; }, no_fuse, hk);
    }
    return substep_t(s, subdt, inv_dt, gstep, pick, picked, ppos, [&]() { return SceneTable{s->dsc, s->sid.cur()}; },
                     no_fuse, hk);
}

f3 gravity_step(const bp_params &P, float subdt) {
    // make_float3(0,-9.8,0) * dt * dt (particlesolver.cuh:56), per component
    return mk3(P.gravity[0] * subdt * subdt, P.gravity[1] * subdt * subdt, P.gravity[2] * subdt * subdt);
}

}  // namespace bp_host

namespace {

// Runs `count` plain substeps; with a graph enabled, pairs of substeps replay a
// captured graph (the ping-pong buffers return to their start after two).
bp_status run_substeps(bp_solver *s, int count, float subdt) {
    const float inv_dt = 1.0f / subdt;  // particlesolver.cuh:1000
    const f3 gstep = gravity_step(s->P, subdt);
    int k = 0;
    if (s->use_graph && !s->timing) {
        while (k < count && s->sk_idx < 0) {  // the first substep has no ranges to clear
            bp_status r = substep(s, subdt, inv_dt, gstep, false, -1, mk3(0.f, 0.f, 0.f), !s->fuse);
            if (r != BP_OK) return r;
            ++k;
        }
        while (count - k >= 2) {
            if (!s->gexec || s->graph_dt != subdt || s->graph_sig != s->buffer_sig()) {
                s->drop_graph();
                const int sig0 = s->buffer_sig();
                HIP_TRY(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
                bp_status r1 = substep(s, subdt, inv_dt, gstep, false, -1, mk3(0.f, 0.f, 0.f), !s->fuse);
                bp_status r2 = (r1 == BP_OK) ? substep(s, subdt, inv_dt, gstep, false, -1, mk3(0.f, 0.f, 0.f), !s->fuse) : r1;
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(s->st, &g);
                if (r2 != BP_OK) return r2;
                if (e != hipSuccess) return fail(BP_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
                s->graph = g;
                HIP_TRY(hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
                s->graph_dt = subdt;
                s->graph_sig = sig0;
                if (s->buffer_sig() != sig0) return fail(BP_ERR_HIP, "graph capture: buffers not periodic over two substeps");
            }
            // capture recorded the launches without running them
            HIP_TRY(hipGraphLaunch(s->gexec, s->st));
            k += 2;
        }
    }
    for (; k < count; ++k) {
        bp_status r = substep(s, subdt, inv_dt, gstep, false, -1, mk3(0.f, 0.f, 0.f), !s->fuse);
        if (r != BP_OK) return r;
    }
    return BP_OK;
}

bp_status collect_timing(bp_solver *s) {
    if (s->recs.empty()) return BP_OK;
    HIP_TRY(hipStreamSynchronize(s->st));
    for (const auto &r : s->recs) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
        s->fam_ms[r.fam] += ms;
        s->fam_n[r.fam] += 1;
    }
    s->recs.clear();
    s->ev_next = 0;
    return BP_OK;
}

bp_status validate_scene(const bp_scene_desc *sc) {
    const int N = sc->num_particles;
    if (N < 0) return fail(BP_ERR_INVALID_ARG, "num_particles < 0");
    if (N > 0 && (!sc->positions || !sc->masses || !sc->phases))
        return fail(BP_ERR_INVALID_ARG, "positions, masses and phases are required");
    // the lambda pass's partner record carries the mass with the partner's
    // fluid flag in its sign, so a negative mass would not round-trip
    for (int i = 0; i < N; ++i)
        if (!(sc->masses[i] >= 0.0f)) return fail(BP_ERR_INVALID_ARG, "masses must be >= 0");
    for (int a = 0; a < 3; ++a) {
        if (!is_pow2(sc->grid_dims[a]) || sc->grid_dims[a] < 4 || sc->grid_dims[a] > 1024)
            return fail(BP_ERR_INVALID_ARG, "grid_dims must be powers of two in [4, 1024]");
        if (!(sc->cell_size[a] > 0.0f)) return fail(BP_ERR_INVALID_ARG, "cell_size must be > 0");
    }
    if (!(sc->kernel_radius > 0.0f)) return fail(BP_ERR_INVALID_ARG, "kernel_radius must be > 0");
    if (sc->num_planes < 0 || sc->num_planes > kMaxPlanes)
        return fail(BP_ERR_INVALID_ARG, "num_planes must be in [0, 16]");
    if (sc->num_planes > 0 && (!sc->plane_origins || !sc->plane_normals))
        return fail(BP_ERR_INVALID_ARG, "plane arrays are NULL");
    auto ids_ok = [&](const int *ids, long long count) {
        if (count > 0 && !ids) return false;
        for (long long i = 0; i < count; ++i)
            if (ids[i] < 0 || ids[i] >= N) return false;
        return true;
    };
    if (sc->num_distance < 0 || !ids_ok(sc->distance_pairs, 2LL * sc->num_distance) ||
        (sc->num_distance > 0 && !sc->distance_params))
        return fail(BP_ERR_INVALID_ARG, "invalid distance constraints");
    if (sc->num_bendings < 0 || !ids_ok(sc->bendings, 4LL * sc->num_bendings))
        return fail(BP_ERR_INVALID_ARG, "invalid bending constraints");
    if (sc->num_faces < 0 || !ids_ok(sc->faces, 3LL * sc->num_faces))
        return fail(BP_ERR_INVALID_ARG, "invalid wind faces");
    if (sc->num_immovables < 0 || !ids_ok(sc->immovables, sc->num_immovables))
        return fail(BP_ERR_INVALID_ARG, "invalid immovable ids");
    if (sc->num_rigid < 0) return fail(BP_ERR_INVALID_ARG, "num_rigid < 0");
    if (sc->num_rigid > 0 && !sc->rigid_ranges) return fail(BP_ERR_INVALID_ARG, "rigid_ranges is NULL");
    std::vector<std::pair<int, int>> bodies;
    for (int b = 0; b < sc->num_rigid; ++b) {
        const int lo = sc->rigid_ranges[2 * b], hi = sc->rigid_ranges[2 * b + 1];
        if (lo < 0 || hi <= lo || hi > N || hi > sc->rigid_rest_len || !sc->rigid_rest_positions)
            return fail(BP_ERR_INVALID_ARG, "invalid rigid range or rest positions");
        if (hi - lo > kMaxRigidParticles)
            return fail(BP_ERR_CAPACITY, "rigid body has more than 64 particles "
                                         "(NUM_MAX_PARTICLE_PER_RIGID_BODY)");
        bodies.emplace_back(lo, hi);
    }
    // one shape-matching workgroup per body writes its particles: two bodies
    // sharing a particle would race on it
    std::sort(bodies.begin(), bodies.end());
    for (size_t b = 1; b < bodies.size(); ++b)
        if (bodies[b].first < bodies[b - 1].second)
            return fail(BP_ERR_INVALID_ARG, "rigid ranges overlap");
    return BP_OK;
}

// Creates a solver over `ns` scenes laid end to end (ns == 1: the plain solver).
bp_status create_impl(const bp_scene_desc *scs, const bp_params *params, int ns, void *hip_stream,
                      bp_solver **out) {
    if (!out) return fail(BP_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!scs || !params || ns <= 0) return fail(BP_ERR_INVALID_ARG, "scenes / params missing");
    long long total = 0, cells_total = 0;
    for (int k = 0; k < ns; ++k) {
        bp_status st = check_params(&params[k]);
        if (st != BP_OK) return st;
        if (!same_schedule(params[0], params[k]))
            return fail(BP_ERR_INVALID_ARG, "scenes of a batch must share iterations, pass enables, gravity and "
                                            "solid/cloth constants");
        st = validate_scene(&scs[k]);
        if (st != BP_OK) return st;
        total += scs[k].num_particles;
        cells_total += (long long)scs[k].grid_dims[0] * scs[k].grid_dims[1] * scs[k].grid_dims[2];
    }
    if (total > kMaxSlots || cells_total > (1LL << 30))
        return fail(BP_ERR_CAPACITY, "at most 2^27 particles and 2^30 cells per solver");

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(BP_ERR_NO_DEVICE, "no HIP device visible");
    bp_solver *s = new bp_solver();
#define TRY(x)                    \
    do {                          \
        bp_status r_ = (x);       \
        if (r_ != BP_OK) {        \
            bp_solver_destroy(s); \
            return r_;            \
        }                         \
    } while (0)
#define HTRY(x) TRY(([&]() -> bp_status { HIP_TRY(x); return BP_OK; })())
    HTRY(hipGetDevice(&s->device));
    {
        hipDeviceProp_t prop;
        HTRY(hipGetDeviceProperties(&prop, s->device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            bp_solver_destroy(s);
            return fail(BP_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", this build targets gfx950");
        }
    }
    if (hip_stream) {
        s->st = (hipStream_t)hip_stream;
    } else {  // own stream (graph capture needs one); default priority
        HTRY(hipStreamCreate(&s->st));
        s->own_stream = true;
    }
    s->P = params[0];
    s->N = (int)total;
    s->nscenes = ns;
    s->num_cells = (int)cells_total;
    s->sort_bits = ilog2((unsigned)s->num_cells);
    const int N = s->N;

    // concatenate: particles, constraints (ids offset), per-scene constants
    std::vector<float4> hx(N), hv(N);
    std::vector<int> iota(N), hsid(N);
    std::vector<float> hm(N);
    std::vector<int> hph(N), hgrp(N);
    std::vector<int> dp, bd, fc, rr, im;
    std::vector<float> dpar;
    std::vector<float4> rest;
    s->poff.assign(ns + 1, 0);
    uint32_t cell_base = 0;
    s->hsc.resize(ns);
    for (int k = 0; k < ns; ++k) {
        const bp_scene_desc &sc = scs[k];
        const int off = s->poff[k];
        s->poff[k + 1] = off + sc.num_particles;
        for (int i = 0; i < sc.num_particles; ++i) {
            const int g = off + i;
            hx[g] = make_float4(sc.positions[3 * i], sc.positions[3 * i + 1], sc.positions[3 * i + 2], 0.f);
            hv[g] = sc.velocities ? make_float4(sc.velocities[3 * i], sc.velocities[3 * i + 1],
                                                sc.velocities[3 * i + 2], 0.f)
                                  : make_float4(0.f, 0.f, 0.f, 0.f);
            hm[g] = sc.masses[i];
            hph[g] = sc.phases[i];
            hgrp[g] = sc.group_ids ? sc.group_ids[i] : 0;
            iota[g] = g;
            hsid[g] = k;
            if (sc.phases[i] > 0) s->has_solid = true; else s->has_fluid = true;
        }
        if (k == 0) s->one_phase = sc.num_particles > 0;
        for (int i = 0; i < sc.num_particles; ++i) s->one_phase = s->one_phase && sc.phases[i] == hph[0] && hph[0] < 0;
        for (int i = 0; i < 2 * sc.num_distance; ++i) dp.push_back(sc.distance_pairs[i] + off);
        for (int i = 0; i < 2 * sc.num_distance; ++i) dpar.push_back(sc.distance_params[i]);
        for (int i = 0; i < 4 * sc.num_bendings; ++i) bd.push_back(sc.bendings[i] + off);
        for (int i = 0; i < 3 * sc.num_faces; ++i) fc.push_back(sc.faces[i] + off);
        for (int i = 0; i < sc.num_immovables; ++i) im.push_back(sc.immovables[i] + off);
        for (int b = 0; b < sc.num_rigid; ++b) {
            rr.push_back(sc.rigid_ranges[2 * b] + off);
            rr.push_back(sc.rigid_ranges[2 * b + 1] + off);
        }
        if (sc.num_rigid > 0) {
            if (rest.size() < (size_t)(off + sc.rigid_rest_len)) rest.resize(off + sc.rigid_rest_len);
            for (int i = 0; i < sc.rigid_rest_len; ++i)
                rest[off + i] = make_float4(sc.rigid_rest_positions[3 * i], sc.rigid_rest_positions[3 * i + 1],
                                            sc.rigid_rest_positions[3 * i + 2], 0.f);
        }
        SceneConst &c = s->hsc[k];
        c = SceneConst{};
        for (int a = 0; a < 3; ++a) {
            c.G.cs[a] = sc.cell_size[a];
            c.G.dims[a] = sc.grid_dims[a];
            c.G.mask[a] = (uint32_t)sc.grid_dims[a] - 1u;
        }
        c.G.stride_x = (uint32_t)(sc.grid_dims[1] * sc.grid_dims[2]);
        c.G.shift_y = (uint32_t)ilog2((unsigned)sc.grid_dims[2]);
        c.G.shift_x = (uint32_t)ilog2((unsigned)(sc.grid_dims[1] * sc.grid_dims[2]));
        c.G.base = cell_base;
        cell_base += (uint32_t)(sc.grid_dims[0] * sc.grid_dims[1] * sc.grid_dims[2]);
        set_scene_const(c, params[k], sc.kernel_radius);
        c.radius = sc.particle_radius;
        c.pl.n = sc.num_planes;
        for (int p = 0; p < sc.num_planes; ++p)
            for (int a = 0; a < 3; ++a) {
                c.pl.o[p][a] = sc.plane_origins[3 * p + a];
                c.pl.nrm[p][a] = sc.plane_normals[3 * p + a];
            }
        if (sc.num_planes > 0) s->any_planes = true;
    }
    s->mass_by_orig = hm;
    s->phase_by_orig = hph;
    s->scene_of_orig = hsid;

    for (int k = 0; k < 2; ++k) {
        TRY(s->alloc(&s->xs.b[k], (size_t)N + kWalkPad));  // neighbour walks read past the last slot
        TRY(s->alloc(&s->x.b[k], N));
        TRY(s->alloc(&s->v.b[k], N));
        TRY(s->alloc(&s->m.b[k], N));
        TRY(s->alloc(&s->ph.b[k], N));
        TRY(s->alloc(&s->orig.b[k], N));
        if (ns > 1) TRY(s->alloc(&s->sid.b[k], N));
        TRY(s->alloc(&s->keys[k], N));
        TRY(s->alloc(&s->vals[k], N));
    }
    TRY(s->alloc(&s->map_to_new, N));
    TRY(s->alloc(&s->guard_slot, 1));
    TRY(s->alloc(&s->counts, radix_counts_len(N)));
    TRY(s->alloc(&s->totals, 256));
    TRY(s->alloc(&s->ranges, s->num_cells));
    TRY(s->alloc(&s->lambda, N));
    TRY(s->alloc(&s->density, N));
    TRY(s->alloc(&s->ccell, N));
    TRY(s->alloc(&s->group, N));
    if (s->has_fluid) TRY(s->alloc(&s->srec, (size_t)kRecChunks * N));
    if (s->has_fluid) TRY(s->alloc(&s->lam_rec, N));
    if (s->has_fluid) {
        s->lin.pstride = (N + kWalkPad + 63) / 64 * 64;  // kWalkPad slots of slack, 256-byte aligned planes
        TRY(s->alloc(&s->lin.planes, 3 * (size_t)s->lin.pstride));
        TRY(s->alloc(&s->lin.xm, N));
    }
    TRY(s->alloc(&s->sum_tmp, N));
    TRY(s->alloc(&s->corr, 3 * (size_t)N));
    TRY(s->alloc(&s->dflag, (size_t)N));
    TRY(s->alloc(&s->omega, N));
    TRY(s->alloc(&s->xom, N));
    TRY(s->alloc(&s->normal, N));
    if (s->has_solid && s->has_fluid) {
        TRY(s->alloc(&s->cell_stamp, s->num_cells));
        TRY(s->alloc(&s->fluid_stamp, s->num_cells));
        TRY(s->alloc(&s->stamp, 1));
        HTRY(hipMemsetAsync(s->cell_stamp, 0, (size_t)s->num_cells * sizeof(uint32_t), s->st));
        HTRY(hipMemsetAsync(s->fluid_stamp, 0, (size_t)s->num_cells * sizeof(uint32_t), s->st));
        HTRY(hipMemsetAsync(s->stamp, 0, sizeof(uint32_t), s->st));
    }
    if (ns > 1) TRY(s->upload(&s->dsc, s->hsc.data(), s->hsc.size()));
    if (N > 0) {
        HTRY(hipMemcpyAsync(s->x.cur(), hx.data(), N * sizeof(float4), hipMemcpyHostToDevice, s->st));
        HTRY(hipMemcpyAsync(s->v.cur(), hv.data(), N * sizeof(float4), hipMemcpyHostToDevice, s->st));
        HTRY(hipMemcpyAsync(s->m.cur(), hm.data(), N * sizeof(float), hipMemcpyHostToDevice, s->st));
        HTRY(hipMemcpyAsync(s->ph.cur(), hph.data(), N * sizeof(int), hipMemcpyHostToDevice, s->st));
        HTRY(hipMemcpyAsync(s->group, hgrp.data(), N * sizeof(int), hipMemcpyHostToDevice, s->st));
        HTRY(hipMemcpyAsync(s->orig.cur(), iota.data(), N * sizeof(int), hipMemcpyHostToDevice, s->st));
        HTRY(hipMemcpyAsync(s->map_to_new, iota.data(), N * sizeof(int), hipMemcpyHostToDevice, s->st));
        if (ns > 1) HTRY(hipMemcpyAsync(s->sid.cur(), hsid.data(), N * sizeof(int), hipMemcpyHostToDevice, s->st));
        HTRY(hipMemsetAsync(s->lambda, 0, N * sizeof(float), s->st));
        HTRY(hipMemsetAsync(s->density, 0, N * sizeof(float), s->st));
        HTRY(hipMemsetAsync(s->omega, 0, N * sizeof(float4), s->st));
        HTRY(hipMemsetAsync(s->xom, 0, N * sizeof(float4), s->st));
        HTRY(hipMemsetAsync(s->normal, 0, N * sizeof(float4), s->st));
        HTRY(hipMemsetAsync(s->corr, 0, 3 * (size_t)N * sizeof(unsigned long long), s->st));
        HTRY(hipMemsetAsync(s->dflag, 0, (size_t)N, s->st));
    }
    HTRY(hipMemsetAsync(s->ranges, 0xFF, (size_t)s->num_cells * sizeof(int2), s->st));

    s->nd = (int)dp.size() / 2;
    if (s->nd > 0) {
        TRY(s->upload(&s->dpairs, (const int2 *)dp.data(), s->nd));
        TRY(s->upload(&s->dparams, (const float2 *)dpar.data(), s->nd));
        TRY(s->alloc(&s->kbuf, s->nd));
        TRY(build_csr(s, s->dcsr, 2, dp.data(), s->nd, N, 1));
    }
    s->nb = (int)bd.size() / 4;
    if (s->nb > 0) {
        TRY(s->upload(&s->bends, (const int4 *)bd.data(), s->nb));
        TRY(s->alloc(&s->bbuf, 4 * (size_t)s->nb));
        TRY(build_csr(s, s->bcsr, 4, bd.data(), s->nb, N, 1));
    }
    s->nf = (int)fc.size() / 3;
    if (s->nf > 0) {
        std::vector<int4> f4v(s->nf);
        for (int f = 0; f < s->nf; ++f) f4v[f] = make_int4(fc[3 * f], fc[3 * f + 1], fc[3 * f + 2], 0);
        TRY(s->upload(&s->faces, f4v.data(), s->nf));
        TRY(s->alloc(&s->fbuf, s->nf));
        TRY(build_csr(s, s->fcsr, 3, fc.data(), s->nf, N, 3));
    }
    s->nr = (int)rr.size() / 2;
    if (s->nr > 0) {
        TRY(s->upload(&s->rranges, (const int2 *)rr.data(), s->nr));
        TRY(s->upload(&s->rrest, rest.data(), rest.size()));
        std::vector<float4> q(s->nr, make_float4(0.f, 0.f, 0.f, 1.f));
        TRY(s->upload(&s->rrot, q.data(), q.size()));
        TRY(s->alloc(&s->rcm, s->nr));
        HTRY(hipMemsetAsync(s->rcm, 0, s->nr * sizeof(float4), s->st));
    }
    s->ni = (int)im.size();
    if (s->ni > 0) TRY(s->upload(&s->immov, im.data(), s->ni));
    {  // may the shape-matching launch also apply the pins?
        std::vector<char> in_body(N, 0);
        for (int b = 0; b < s->nr; ++b)
            for (int i = rr[2 * b]; i < rr[2 * b + 1]; ++i) in_body[i] = 1;
        s->pins_apart_from_rigid = true;
        for (int id : im) s->pins_apart_from_rigid = s->pins_apart_from_rigid && !in_body[id];
    }
    HTRY(hipStreamSynchronize(s->st));
#undef HTRY
#undef TRY
    *out = s;
    return BP_OK;
}

}  // namespace

extern "C" {

int bp_abi_version(void) { return BP_ABI_VERSION; }

const char *bp_last_error_string(void) { return g_last_error.c_str(); }

bp_status bp_default_params(bp_params *p) {
    if (!p) return fail(BP_ERR_INVALID_ARG, "out is NULL");
    std::memset(p, 0, sizeof(*p));
    p->iterations = 2;
    p->stabilization_passes = 2;
    p->gravity[0] = 0.0f;
    p->gravity[1] = -9.8f;
    p->gravity[2] = 0.0f;
    p->fluid_rest_density = 1200.0f;  // main.cu:223 preset value
    p->solid_density_scale = 2.0f;
    p->lambda_mode = 0;
    p->lambda_denominator = 15000.0f;
    p->lambda_relaxation = 300.0f;
    p->akinci = 1;
    p->scorr_k = 0.0001f;
    p->scorr_n = 4;
    p->scorr_dq = 0.03f;
    p->delta_clamp = 5.0f;
    p->vorticity_eps = 0.001f;
    p->tension_gamma = 0.5f;
    p->xsph_c = 0.0002f;
    p->friction_static = 0.5f;
    p->friction_dynamic = 0.0f;
    p->mass_scaling_k = 2.0f;
    p->sleep_epsilon = 0.0f;
    p->wind_enabled = 1;
    p->drag_coeff = 10.0f;
    p->lift_coeff = 100.0f;
    p->bend_stiffness = 0.8f;
    p->shape_matching_iterations = 20;
    p->nan_guard = 0;
    return BP_OK;
}

bp_status bp_solver_create(const bp_scene_desc *sc, const bp_params *params, void *hip_stream, bp_solver **out) {
    if (!sc) return fail(BP_ERR_INVALID_ARG, "scene is NULL");
    return create_impl(sc, params, 1, hip_stream, out);
}

bp_status bp_batch_create(const bp_scene_desc *scenes, const bp_params *params, int num_scenes, void *hip_stream,
                          bp_solver **out) {
    return create_impl(scenes, params, num_scenes, hip_stream, out);
}

bp_status bp_batch_scene_range(bp_solver *s, int scene, int *first_original_id, int *count) {
    if (!s || !first_original_id || !count) return fail(BP_ERR_INVALID_ARG, "NULL argument");
    if (scene < 0 || scene >= s->nscenes) return fail(BP_ERR_INVALID_ARG, "scene out of range");
    *first_original_id = s->poff[scene];
    *count = s->poff[scene + 1] - s->poff[scene];
    return BP_OK;
}

bp_status bp_solver_destroy(bp_solver *s) {
    if (!s) return BP_OK;
    if (s->st) (void)hipStreamSynchronize(s->st); else (void)hipDeviceSynchronize();
    s->drop_graph();
    adjoint_release(s);
    for (void *p : s->allocs) (void)hipFree(p);
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    if (s->own_stream) (void)hipStreamDestroy(s->st);
    delete s;
    return BP_OK;
}

bp_status bp_solver_set_params(bp_solver *s, const bp_params *p) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    adjoint_invalidate(s);
    bp_status st = check_params(p);
    if (st != BP_OK) return st;
    if (s->nscenes != 1) return fail(BP_ERR_UNSUPPORTED, "per-scene parameters of a batch are fixed at creation");
    s->P = *p;
    set_scene_const(s->hsc[0], *p, s->hsc[0].K.h);
    s->drop_graph();
    return BP_OK;
}

bp_status bp_solver_step(bp_solver *s, int num_substeps, float dt, int picked, const float *ppos,
                         const float *pvel) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    adjoint_invalidate(s);
    if (num_substeps <= 0) return fail(BP_ERR_INVALID_ARG, "num_substeps must be > 0");
    if (s->N <= 0) return BP_OK;  // particlesolver.cuh:778
    const float subdt = dt / (float)num_substeps;  // :780
    const bool pick = picked >= 0 && picked < s->N;  // :790
    if (pick && !ppos) return fail(BP_ERR_INVALID_ARG, "picked position is NULL");
    if (!pick) {
        bp_status st = run_substeps(s, num_substeps, subdt);
        if (st != BP_OK) return st;
    } else {
        const float inv_dt = 1.0f / subdt;
        const f3 gstep = gravity_step(s->P, subdt);
        const f3 pp = mk3(ppos[0], ppos[1], ppos[2]);
        for (int k = 0; k < num_substeps; ++k) {
            bp_status st = substep(s, subdt, inv_dt, gstep, true, picked, pp, !s->fuse);
            if (st != BP_OK) return st;
        }
        // :1066-1071: keep the solved position, set the picked velocity
        const f3 pv = pvel ? mk3(pvel[0], pvel[1], pvel[2]) : mk3(0.f, 0.f, 0.f);
        TIMED(F_PICK, launch_pick(s->x.cur(), s->v.cur(), s->map_to_new, picked, pp, 1, pv, s->st));
    }
    HIP_TRY(hipGetLastError());
    if (s->P.nan_guard) {  // scanned on the device; 4 bytes come back
        int first = s->N;
        HIP_TRY(hipMemcpyAsync(s->guard_slot, &first, sizeof(int), hipMemcpyHostToDevice, s->st));
        launch_first_nonfinite(s->x.cur(), s->N, s->guard_slot, s->st);
        HIP_TRY(hipMemcpyAsync(&first, s->guard_slot, sizeof(int), hipMemcpyDeviceToHost, s->st));
        HIP_TRY(hipStreamSynchronize(s->st));
        if (first < s->N) {
            int orig_id = -1;
            HIP_TRY(hipMemcpy(&orig_id, s->orig.cur() + first, sizeof(int), hipMemcpyDeviceToHost));
            return fail(BP_ERR_NOT_FINITE, guard_message(first, orig_id));
        }
    }
    return BP_OK;
}

bp_status bp_solver_set_graph(bp_solver *s, int enable) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    s->use_graph = enable != 0;
    if (!s->use_graph) s->drop_graph();
    return BP_OK;
}

bp_status bp_solver_set_fusion(bp_solver *s, int enable) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    s->fuse = enable != 0;
    s->drop_graph();
    return BP_OK;
}

bp_status bp_solver_set_block_sizes(bp_solver *s, const int sizes[6]) {
    if (!s || !sizes) return fail(BP_ERR_INVALID_ARG, "NULL argument");
    for (int f = 0; f < 6; ++f)
        if (sizes[f] != 64 && sizes[f] != 128 && sizes[f] != 256 && sizes[f] != 512)
            return fail(BP_ERR_INVALID_ARG, "workgroup sizes must be 64, 128, 256 or 512");
    s->blk_lambda = sizes[0];
    s->blk_position = sizes[1];
    s->blk_post_a = sizes[2];
    s->blk_post_b = sizes[3];
    s->blk_xsph = sizes[4];
    s->blk_pp = sizes[5];
    s->drop_graph();
    return BP_OK;
}

bp_status bp_solver_set_walk(bp_solver *s, int lambda_staged, int position_flat, int omega_normal_staged,
                             int vorticity_tension_flat) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    // 1 = the default region size
    const int cl = lambda_staged == 1 ? kStageCapDefault : lambda_staged;
    const int co = omega_normal_staged == 1 ? kStageCapDefault : omega_normal_staged;
    if ((cl != 0 && !stage_cap_valid(cl)) || (co != 0 && !stage_cap_valid(co)))
        return fail(BP_ERR_INVALID_ARG, "staged walks take 0 (per lane), 1 (default region), 128, 256 or 512");
    s->staged_lambda = cl;
    s->staged_post_a = co;
    s->flat_position = position_flat != 0;
    s->flat_post_b = vorticity_tension_flat != 0;
    s->drop_graph();
    return BP_OK;
}

bp_status bp_solver_set_stage_screen(bp_solver *s, int lambda_candidates, int omega_normal_candidates) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    auto ok = [](int c) { return c == 4 || c == 8; };
    if (!ok(lambda_candidates) || !ok(omega_normal_candidates))
        return fail(BP_ERR_INVALID_ARG, "the staged walks screen 4 or 8 candidates per batch");
    s->stage_sb_lambda = lambda_candidates;
    s->stage_sb_post_a = omega_normal_candidates;
    s->drop_graph();
    return BP_OK;
}

bp_status bp_solver_set_partner_fetch(bp_solver *s, int position_partners) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    if (position_partners != 2 && position_partners != 4)
        return fail(BP_ERR_INVALID_ARG, "the position pass fetches 2 or 4 partners per round trip");
    s->pos_partners = position_partners;
    s->drop_graph();
    return BP_OK;
}

namespace {
// GPU time of `count` substeps (HIP events on the handle's stream)
bp_status time_substeps(bp_solver *s, int count, float dt, float *ms) {
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, s->st));
    bp_status st = run_substeps(s, count, dt);
    HIP_TRY(hipEventRecord(b, s->st));
    HIP_TRY(hipEventSynchronize(b));
    HIP_TRY(hipEventElapsedTime(ms, a, b));
    HIP_TRY(hipEventDestroy(a));
    HIP_TRY(hipEventDestroy(b));
    return st;
}
}  // namespace

bp_status bp_solver_autotune(bp_solver *s, float dt, int substeps_per_trial, char *report, int report_len) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    adjoint_invalidate(s);
    if (substeps_per_trial <= 0) return fail(BP_ERR_INVALID_ARG, "substeps_per_trial must be > 0");
    const int cand[4] = {64, 128, 256, 512};
    const int fams[6] = {F_LAMBDA, F_POSITION, F_POST_A, F_POST_B, F_XSPH, F_PP};
    float best_ms[6];
    int best_blk[6], best_flat[6];
    for (int f = 0; f < 6; ++f) {
        best_ms[f] = 1e30f;
        best_blk[f] = 256;
        best_flat[f] = (f == 0 || f == 2) ? kStageCapDefault : 1;
    }
    // Every trial starts from the caller's state (the passes cost more as a
    // scene evolves, so trials run one after another on an advancing state
    // would favour the early ones), and that state is restored at the end.
    std::vector<float> keep_x(3 * (size_t)s->N), keep_v(3 * (size_t)s->N), keep_q(4 * (size_t)s->nr + 4),
        keep_cm(3 * (size_t)s->nr + 3);
    std::vector<int> keep_map(s->N);
    bp_status st = bp_get_state(s, keep_x.data(), keep_v.data(), keep_map.data());
    if (st == BP_OK && s->nr > 0) st = bp_get_rigid_state(s, keep_q.data(), keep_cm.data());
    if (st != BP_OK) return st;
    auto restore = [&]() {
        bp_status r = bp_set_state(s, keep_x.data(), keep_v.data(), keep_map.data());
        if (r == BP_OK && s->nr > 0) r = bp_set_rigid_state(s, keep_q.data());
        return r;
    };
    const bool g = s->use_graph;
    s->use_graph = false;
    st = bp_solver_enable_kernel_timing(s, 1);
    if (st != BP_OK) return st;
    // every workgroup size with each walk form of the lambda, position,
    // omega/normal and vorticity/tension passes (the others have one form).
    // Form 0-6: the lambda and omega/normal passes per lane (0) or staged
    // with each region size (kStageCaps) screening 4 (1-3) or 8 (4-6)
    // candidates per batch; the position pass per chunk (0, 1, 4) or flat
    // (2, 3, 5, 6), fetching 2 (0, 2, 4, 6) or 4 (1, 3, 5) partners per round
    // trip; the vorticity/tension pass per chunk or flat as the position pass.
    // Two rounds over the candidates; a pass keeps the candidate with the
    // lowest mean launch time summed over both rounds.
    constexpr int kForms = 7, kTrials = 4 * kForms, kRounds = 2;
    auto form_cap = [&](int form) { return form ? kStageCaps[(form - 1) % 3] : 0; };
    auto form_sb = [](int form) { return form >= 4 ? 8 : 4; };
    auto form_flat = [](int form) { return (form == 2 || form == 3 || form == 5 || form == 6) ? 1 : 0; };
    auto form_partners = [](int form) { return (form == 1 || form == 3 || form == 5) ? 4 : 2; };
    float sum_ms[kTrials][6] = {};
    bool seen[kTrials][6] = {};
    for (int round = 0; round < kRounds && st == BP_OK; ++round)
        for (int t = 0; t < kTrials; ++t) {
            const int c = t / kForms, form = t % kForms;
            s->blk_lambda = s->blk_position = s->blk_post_a = s->blk_post_b = s->blk_xsph = s->blk_pp = cand[c];
            s->staged_lambda = s->staged_post_a = form_cap(form);
            s->stage_sb_lambda = s->stage_sb_post_a = form_sb(form);
            s->flat_position = s->flat_post_b = form_flat(form);
            s->pos_partners = form_partners(form);
            st = restore();
            if (st != BP_OK) break;
            for (int f = 0; f < F_COUNT; ++f) { s->fam_ms[f] = 0.f; s->fam_n[f] = 0; }
            st = run_substeps(s, substeps_per_trial, dt);
            if (st == BP_OK) st = collect_timing(s);
            if (st != BP_OK) break;
            for (int f = 0; f < 6; ++f) {
                const int fam = fams[f];
                if (s->fam_n[fam] == 0) continue;
                sum_ms[t][f] += s->fam_ms[fam] / (float)s->fam_n[fam];
                seen[t][f] = true;
            }
        }
    int best_partners = 2, best_sb[2] = {4, 4};
    for (int t = 0; t < kTrials && st == BP_OK; ++t) {
        const int c = t / kForms, form = t % kForms;
        for (int f = 0; f < 6; ++f)
            if (seen[t][f] && sum_ms[t][f] < best_ms[f]) {
                best_ms[f] = sum_ms[t][f];
                best_blk[f] = cand[c];
                best_flat[f] = (f == 0 || f == 2) ? form_cap(form) : form_flat(form);  // region size or flat flag
                if (f == 0) best_sb[0] = form_sb(form);
                if (f == 2) best_sb[1] = form_sb(form);
                if (f == 1) best_partners = form_partners(form);
            }
    }
    s->blk_lambda = best_blk[0];
    s->blk_position = best_blk[1];
    s->blk_post_a = best_blk[2];
    s->blk_post_b = best_blk[3];
    s->blk_xsph = best_blk[4];
    s->blk_pp = best_blk[5];
    s->staged_lambda = best_flat[0];
    s->flat_position = best_flat[1];
    s->staged_post_a = best_flat[2];
    s->flat_post_b = best_flat[3];
    s->pos_partners = best_partners;
    s->stage_sb_lambda = best_sb[0];
    s->stage_sb_post_a = best_sb[1];
    (void)bp_solver_enable_kernel_timing(s, 0);
    // iteration fusion: the epilogue passes (next iteration's plane
    // collisions, the velocity/position update, the lambda pass's inputs) in
    // the producing kernels, or every pass its own launch; best of two trials
    float fused_ms[2] = {1e30f, 1e30f};
    for (int trial = 0; trial < 2 && st == BP_OK; ++trial)
        for (int f = 0; f < 2 && st == BP_OK; ++f) {
            s->fuse = f == 1;
            float ms = 0.f;
            st = restore();
            if (st == BP_OK) st = time_substeps(s, substeps_per_trial, dt, &ms);
            if (ms < fused_ms[f]) fused_ms[f] = ms;
        }
    s->fuse = fused_ms[1] <= fused_ms[0];
    s->use_graph = g;
    s->drop_graph();
    if (st != BP_OK) return st;
    st = restore();
    if (st != BP_OK) return st;
    if (report && report_len > 0)
        std::snprintf(report, (size_t)report_len,
                      "{\"fluid_lambda\": %d, \"fluid_position\": %d, \"omega_normal\": %d, "
                      "\"vorticity_tension\": %d, \"xsph\": %d, \"particle_particle\": %d, "
                      "\"iteration_fusion\": %d, \"lambda_walk\": \"%s\", \"position_walk\": \"%s\", "
                      "\"omega_normal_walk\": \"%s\", \"vorticity_tension_walk\": \"%s\", "
                      "\"lambda_stage_records\": %d, \"omega_normal_stage_records\": %d, "
                      "\"position_partners\": %d, \"lambda_stage_screen\": %d, \"omega_normal_stage_screen\": %d}",
                      best_blk[0], best_blk[1], best_blk[2], best_blk[3], best_blk[4], best_blk[5], s->fuse ? 1 : 0,
                      s->staged_lambda ? "staged" : "global", s->flat_position ? "flat" : "chunk",
                      s->staged_post_a ? "staged" : "global", s->flat_post_b ? "flat" : "chunk", s->staged_lambda,
                      s->staged_post_a, s->pos_partners, s->stage_sb_lambda, s->stage_sb_post_a);
    return BP_OK;
}

bp_status bp_solver_sync(bp_solver *s) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    HIP_TRY(hipStreamSynchronize(s->st));
    return BP_OK;
}

namespace {
bp_status read_int(bp_solver *s, const int *dev, int idx, int *out) {
    HIP_TRY(hipMemcpyAsync(out, dev + idx, sizeof(int), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return BP_OK;
}
}  // namespace

bp_status bp_query_original_id(bp_solver *s, int slot, int *out) {
    if (!s || !out) return fail(BP_ERR_INVALID_ARG, "NULL argument");
    if (slot < 0 || slot >= s->N) return fail(BP_ERR_INVALID_ARG, "slot out of range");
    return read_int(s, s->orig.cur(), slot, out);
}

bp_status bp_query_slot(bp_solver *s, int original_id, int *out) {
    if (!s || !out) return fail(BP_ERR_INVALID_ARG, "NULL argument");
    if (original_id < 0 || original_id >= s->N) return fail(BP_ERR_INVALID_ARG, "original id out of range");
    return read_int(s, s->map_to_new, original_id, out);
}

bp_status bp_get_particle(bp_solver *s, int slot, float out_pos[3]) {
    if (!s || !out_pos) return fail(BP_ERR_INVALID_ARG, "NULL argument");
    if (slot < 0 || slot >= s->N) {  // particlesolver.cuh:380 returns vec3(0)
        out_pos[0] = out_pos[1] = out_pos[2] = 0.f;
        return BP_OK;
    }
    float4 a;
    HIP_TRY(hipMemcpyAsync(&a, s->x.cur() + slot, sizeof(float4), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    out_pos[0] = a.x; out_pos[1] = a.y; out_pos[2] = a.z;
    return BP_OK;
}

bp_status bp_set_particle(bp_solver *s, int slot, const float pos[3], const float vel[3]) {
    if (!s || !pos || !vel) return fail(BP_ERR_INVALID_ARG, "NULL argument");
    adjoint_invalidate(s);
    if (slot < 0 || slot >= s->N) return BP_OK;  // particlesolver.cuh:390
    float4 a, b;
    HIP_TRY(hipMemcpyAsync(&a, s->x.cur() + slot, sizeof(float4), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    a.x = pos[0]; a.y = pos[1]; a.z = pos[2];
    b = make_float4(vel[0], vel[1], vel[2], 0.f);
    HIP_TRY(hipMemcpyAsync(s->x.cur() + slot, &a, sizeof(float4), hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipMemcpyAsync(s->v.cur() + slot, &b, sizeof(float4), hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return BP_OK;
}

bp_status bp_get_device_arrays(bp_solver *s, bp_device_view *o) {
    if (!s || !o) return fail(BP_ERR_INVALID_ARG, "NULL argument");
    o->num_particles = s->N;
    o->new_positions = (float *)s->xs.cur();
    o->positions = (float *)s->x.cur();
    o->velocities = (float *)s->v.cur();
    o->phases = s->ph.cur();
    o->map_to_original = s->orig.cur();
    o->map_to_new = s->map_to_new;
    o->rigid_rotations = (float *)s->rrot;
    o->rigid_cms = (float *)s->rcm;
    return BP_OK;
}

bp_status bp_get_state(bp_solver *s, float *pos, float *vel, int *map) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    const int N = s->N;
    std::vector<float4> hx(N), hv(N);
    if (N > 0) {
        HIP_TRY(hipMemcpyAsync(hx.data(), s->x.cur(), N * sizeof(float4), hipMemcpyDeviceToHost, s->st));
        HIP_TRY(hipMemcpyAsync(hv.data(), s->v.cur(), N * sizeof(float4), hipMemcpyDeviceToHost, s->st));
        if (map) HIP_TRY(hipMemcpyAsync(map, s->orig.cur(), N * sizeof(int), hipMemcpyDeviceToHost, s->st));
    }
    HIP_TRY(hipStreamSynchronize(s->st));
    for (int i = 0; i < N; ++i) {
        if (pos) { pos[3 * i] = hx[i].x; pos[3 * i + 1] = hx[i].y; pos[3 * i + 2] = hx[i].z; }
        if (vel) { vel[3 * i] = hv[i].x; vel[3 * i + 1] = hv[i].y; vel[3 * i + 2] = hv[i].z; }
    }
    return BP_OK;
}

bp_status bp_copy_state_original_order(bp_solver *s, float *pos, float *vel) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    const int N = s->N;
    std::vector<float> p(3 * (size_t)N), v(3 * (size_t)N);
    std::vector<int> map(N);
    bp_status st = bp_get_state(s, p.data(), v.data(), map.data());
    if (st != BP_OK) return st;
    for (int i = 0; i < N; ++i) {
        const int o = map[i];
        for (int a = 0; a < 3; ++a) {
            if (pos) pos[3 * o + a] = p[3 * i + a];
            if (vel) vel[3 * o + a] = v[3 * i + a];
        }
    }
    return BP_OK;
}

bp_status bp_set_state(bp_solver *s, const float *pos, const float *vel, const int *map) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    adjoint_invalidate(s);
    const int N = s->N;
    if (N == 0) return BP_OK;
    std::vector<float4> hx(N), hv(N);
    std::vector<int> o(N), inv(N, -1);
    std::vector<float> hm(N);
    std::vector<int> hp(N);
    HIP_TRY(hipMemcpyAsync(hx.data(), s->x.cur(), N * sizeof(float4), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipMemcpyAsync(hv.data(), s->v.cur(), N * sizeof(float4), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipMemcpyAsync(o.data(), s->orig.cur(), N * sizeof(int), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    if (map) {
        for (int i = 0; i < N; ++i) {
            if (map[i] < 0 || map[i] >= N || inv[map[i]] >= 0)
                return fail(BP_ERR_INVALID_ARG, "map_to_original is not a permutation");
            inv[map[i]] = i;
            o[i] = map[i];
        }
    }
    for (int i = 0; i < N; ++i) {
        if (pos) hx[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], hx[i].w);
        if (vel) hv[i] = make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.f);
        hm[i] = s->mass_by_orig[o[i]];
        hp[i] = s->phase_by_orig[o[i]];
        inv[o[i]] = i;
    }
    HIP_TRY(hipMemcpyAsync(s->x.cur(), hx.data(), N * sizeof(float4), hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipMemcpyAsync(s->v.cur(), hv.data(), N * sizeof(float4), hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipMemcpyAsync(s->orig.cur(), o.data(), N * sizeof(int), hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipMemcpyAsync(s->map_to_new, inv.data(), N * sizeof(int), hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipMemcpyAsync(s->m.cur(), hm.data(), N * sizeof(float), hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipMemcpyAsync(s->ph.cur(), hp.data(), N * sizeof(int), hipMemcpyHostToDevice, s->st));
    std::vector<int> hs;
    if (s->nscenes > 1) {
        hs.resize(N);
        for (int i = 0; i < N; ++i) hs[i] = s->scene_of_orig[o[i]];
        HIP_TRY(hipMemcpyAsync(s->sid.cur(), hs.data(), N * sizeof(int), hipMemcpyHostToDevice, s->st));
    }
    HIP_TRY(hipStreamSynchronize(s->st));
    return BP_OK;
}

bp_status bp_get_rigid_state(bp_solver *s, float *quats, float *cms) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    if (s->nr == 0) return BP_OK;
    std::vector<float4> q(s->nr), c(s->nr);
    HIP_TRY(hipMemcpyAsync(q.data(), s->rrot, s->nr * sizeof(float4), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipMemcpyAsync(c.data(), s->rcm, s->nr * sizeof(float4), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    for (int b = 0; b < s->nr; ++b) {
        if (quats) { quats[4 * b] = q[b].x; quats[4 * b + 1] = q[b].y; quats[4 * b + 2] = q[b].z; quats[4 * b + 3] = q[b].w; }
        if (cms) { cms[3 * b] = c[b].x; cms[3 * b + 1] = c[b].y; cms[3 * b + 2] = c[b].z; }
    }
    return BP_OK;
}

// Column-major 4x4 of a rigid body: rotation of the unit quaternion (x, y, z, w),
// translation = centre of mass (mapMatrices, main.cu:284-291; the terms are
// summed in the order of extract_rotation_matrix4, cudamatrix.cuh:175-199).
static void rigid_matrix(const float *q, const float *cm, float *m) {
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float xx = x * x, yy = y * y, zz = z * z;
    m[0] = 1.0f - 2.0f * yy - 2.0f * zz;
    m[1] = 2.0f * x * y + 2.0f * w * z;
    m[2] = 2.0f * x * z - 2.0f * w * y;
    m[3] = 0.0f;
    m[4] = 2.0f * x * y - 2.0f * w * z;
    m[5] = 1.0f - 2.0f * xx - 2.0f * zz;
    m[6] = 2.0f * y * z + 2.0f * w * x;
    m[7] = 0.0f;
    m[8] = 2.0f * x * z + 2.0f * w * y;
    m[9] = 2.0f * y * z - 2.0f * w * x;
    m[10] = 1.0f - 2.0f * xx - 2.0f * yy;
    m[11] = 0.0f;
    m[12] = cm[0];
    m[13] = cm[1];
    m[14] = cm[2];
    m[15] = 1.0f;
}

bp_status bp_stream_copy_bandwidth(size_t bytes, int reps, float *best_ms) {
    if (!best_ms || reps <= 0 || bytes == 0 || bytes % 16 != 0)
        return fail(BP_ERR_INVALID_ARG, "bytes must be a positive multiple of 16, reps > 0, best_ms non-NULL");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(BP_ERR_NO_DEVICE, "no HIP device visible");
    float4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void **)&a, bytes);
    if (e == hipSuccess) e = hipMalloc((void **)&b, bytes);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, bytes, nullptr);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float best = 1e30f;
    const size_t n = bytes / 16;
    for (int t = 0; t < 16 && e == hipSuccess; ++t) {
        const int blocks = 1024 << (t >> 2), variant = t & 3;
        launch_stream_copy(a, b, n, blocks, variant, nullptr);  // warm-up
        for (int r = 0; r < reps && e == hipSuccess; ++r) {
            e = hipEventRecord(e0, nullptr);
            launch_stream_copy(a, b, n, blocks, variant, nullptr);
            if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess && ms < best) best = ms;
        }
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return fail(BP_ERR_HIP, std::string("stream copy: ") + hipGetErrorString(e));
    *best_ms = best;
    return BP_OK;
}

bp_status bp_map_particle_infos(bp_solver *s, void *device_records) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    if (s->N == 0) return BP_OK;
    if (!device_records) return fail(BP_ERR_INVALID_ARG, "device_records is NULL");
    launch_map_particle_infos(s->x.cur(), s->orig.cur(), s->group, (float4 *)device_records, s->N, s->st);
    HIP_TRY(hipGetLastError());
    return BP_OK;
}

bp_status bp_export_frame(bp_solver *s, float *particles, float *matrices) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    const int N = s->N;
    if (particles && N > 0) {
        if (!s->frame_buf) {
            bp_status r = s->alloc(&s->frame_buf, N);
            if (r != BP_OK) return r;
        }
        launch_map_particle_infos(s->x.cur(), s->orig.cur(), s->group, s->frame_buf, N, s->st);
        HIP_TRY(hipMemcpyAsync(particles, s->frame_buf, N * sizeof(float4), hipMemcpyDeviceToHost, s->st));
        HIP_TRY(hipStreamSynchronize(s->st));
    }
    if (matrices && s->nr > 0) {
        std::vector<float> q(4 * (size_t)s->nr), c(3 * (size_t)s->nr);
        bp_status r = bp_get_rigid_state(s, q.data(), c.data());
        if (r != BP_OK) return r;
        for (int b = 0; b < s->nr; ++b) rigid_matrix(&q[4 * (size_t)b], &c[3 * (size_t)b], matrices + 16 * (size_t)b);
    }
    return BP_OK;
}

bp_status bp_set_rigid_state(bp_solver *s, const float *quats) {
    if (!s || !quats) return fail(BP_ERR_INVALID_ARG, "NULL argument");
    adjoint_invalidate(s);
    if (s->nr == 0) return BP_OK;
    std::vector<float4> q(s->nr);
    for (int b = 0; b < s->nr; ++b) q[b] = make_float4(quats[4 * b], quats[4 * b + 1], quats[4 * b + 2], quats[4 * b + 3]);
    HIP_TRY(hipMemcpyAsync(s->rrot, q.data(), s->nr * sizeof(float4), hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return BP_OK;
}

bp_status bp_debug_fetch(bp_solver *s, int which, void *out) {
    if (!s || !out) return fail(BP_ERR_INVALID_ARG, "NULL argument");
    const int N = s->N;
    if (N == 0) return BP_OK;
    switch (which) {
    case 0: HIP_TRY(hipMemcpyAsync(out, s->density, N * sizeof(float), hipMemcpyDeviceToHost, s->st)); break;
    case 1: HIP_TRY(hipMemcpyAsync(out, s->lambda, N * sizeof(float), hipMemcpyDeviceToHost, s->st)); break;
    case 2:
    case 3: {
        const int *src = (which == 2) ? s->sorted_keys : s->perm;
        if (!src) return fail(BP_ERR_INVALID_ARG, "no substep has run yet");
        HIP_TRY(hipMemcpyAsync(out, src, N * sizeof(int), hipMemcpyDeviceToHost, s->st));
        break;
    }
    case 4:
    case 5: {
        std::vector<float4> t(N);
        HIP_TRY(hipMemcpyAsync(t.data(), which == 4 ? s->omega : s->normal, N * sizeof(float4),
                               hipMemcpyDeviceToHost, s->st));
        HIP_TRY(hipStreamSynchronize(s->st));
        float *o = (float *)out;
        for (int i = 0; i < N; ++i) { o[3 * i] = t[i].x; o[3 * i + 1] = t[i].y; o[3 * i + 2] = t[i].z; }
        break;
    }
    case 6: *(float *)out = adjoint_max_delivered(s); return BP_OK;
    default: return fail(BP_ERR_INVALID_ARG, "unknown debug buffer");
    }
    HIP_TRY(hipStreamSynchronize(s->st));
    return BP_OK;
}

bp_status bp_neighbor_grid(const float *xs, int n, const float cell_size[3], const int grid_dims[3], int *keys,
                           int *perm, int *cell_start, int *cell_end) {
    if (n < 0 || (n > 0 && !xs) || !cell_size || !grid_dims) return fail(BP_ERR_INVALID_ARG, "invalid arguments");
    OneScene one{};
    GridDesc &G = one.c.G;
    for (int a = 0; a < 3; ++a) {
        if (!is_pow2(grid_dims[a]) || grid_dims[a] < 4 || grid_dims[a] > 1024)
            return fail(BP_ERR_INVALID_ARG, "grid_dims must be powers of two in [4, 1024]");
        G.cs[a] = cell_size[a];
        G.dims[a] = grid_dims[a];
        G.mask[a] = (uint32_t)grid_dims[a] - 1u;
    }
    G.stride_x = (uint32_t)(grid_dims[1] * grid_dims[2]);
    G.shift_y = (uint32_t)ilog2((unsigned)grid_dims[2]);
    G.shift_x = (uint32_t)ilog2((unsigned)(grid_dims[1] * grid_dims[2]));
    const int cells = grid_dims[0] * grid_dims[1] * grid_dims[2];
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(BP_ERR_NO_DEVICE, "no HIP device visible");
    std::vector<float4> hx(n);
    for (int i = 0; i < n; ++i) hx[i] = make_float4(xs[3 * i], xs[3 * i + 1], xs[3 * i + 2], 0.f);
    float4 *dx = nullptr, *dv = nullptr;
    int *k[2] = {nullptr, nullptr}, *vv[2] = {nullptr, nullptr}, *counts = nullptr, *totals = nullptr;
    int2 *ranges = nullptr;
    const size_t nn = n > 0 ? (size_t)n : 1;
    hipError_t e = hipSuccess;
    auto A = [&](void **p, size_t bytes) { if (e == hipSuccess) e = hipMalloc(p, bytes); };
    A((void **)&dx, nn * sizeof(float4));
    A((void **)&dv, nn * sizeof(float4));
    A((void **)&k[0], nn * sizeof(int));
    A((void **)&k[1], nn * sizeof(int));
    A((void **)&vv[0], nn * sizeof(int));
    A((void **)&vv[1], nn * sizeof(int));
    A((void **)&counts, radix_counts_len(n > 0 ? n : 1) * sizeof(int));
    A((void **)&totals, 256 * sizeof(int));
    A((void **)&ranges, (size_t)cells * sizeof(int2));
    bp_status result = BP_OK;
    auto chk = [&](hipError_t r) { if (e == hipSuccess) e = r; };
    if (e == hipSuccess) {
        hipStream_t st = nullptr;
        int *sk = nullptr, *pm = nullptr;
        if (n > 0) {
            chk(hipMemcpyAsync(dx, hx.data(), n * sizeof(float4), hipMemcpyHostToDevice, st));
            chk(hipMemsetAsync(dv, 0, n * sizeof(float4), st));
        }
        chk(hipMemsetAsync(ranges, 0xFF, (size_t)cells * sizeof(int2), st));
        if (n > 0) {
            // x* = x + 0*0 + 0 is x exactly: the hash sees the given positions
            launch_predict_hash(dx, dv, k[0], n, 0.0f, mk3(0.f, 0.f, 0.f), one, nullptr, ranges, st);
            launch_radix_sort(k[0], k[1], vv[0], vv[1], counts, totals, n, ilog2((unsigned)cells), &sk, &pm, st);
            launch_find_ranges(sk, n, ranges, st);
            if (keys) chk(hipMemcpyAsync(keys, sk, n * sizeof(int), hipMemcpyDeviceToHost, st));
            if (perm) chk(hipMemcpyAsync(perm, pm, n * sizeof(int), hipMemcpyDeviceToHost, st));
        }
        std::vector<int2> hr;
        if (cell_start || cell_end) {
            hr.resize(cells);
            chk(hipMemcpyAsync(hr.data(), ranges, (size_t)cells * sizeof(int2), hipMemcpyDeviceToHost, st));
        }
        chk(hipGetLastError());
        chk(hipStreamSynchronize(st));
        for (int c = 0; c < (int)hr.size(); ++c) {
            if (cell_start) cell_start[c] = hr[c].x;
            if (cell_end) cell_end[c] = hr[c].y;
        }
    }
    if (e != hipSuccess) result = fail(BP_ERR_HIP, std::string("neighbor grid: ") + hipGetErrorString(e));
    for (void *p : {(void *)dx, (void *)dv, (void *)k[0], (void *)k[1], (void *)vv[0], (void *)vv[1],
                    (void *)counts, (void *)totals, (void *)ranges})
        if (p) (void)hipFree(p);
    return result;
}

bp_status bp_solver_timed_substeps(bp_solver *s, int substeps, float dt, float *out_ms) {
    if (!s || !out_ms) return fail(BP_ERR_INVALID_ARG, "NULL argument");
    adjoint_invalidate(s);
    if (substeps <= 0) return fail(BP_ERR_INVALID_ARG, "substeps must be > 0");
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, s->st));
    if (s->N > 0) {
        bp_status st = run_substeps(s, substeps, dt);
        if (st != BP_OK) return st;
    }
    HIP_TRY(hipEventRecord(b, s->st));
    HIP_TRY(hipEventSynchronize(b));
    HIP_TRY(hipEventElapsedTime(out_ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return collect_timing(s);
}

bp_status bp_solver_enable_kernel_timing(bp_solver *s, int enable) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    s->timing = enable != 0;
    if (s->timing && s->ev_pool.empty()) {
        s->ev_pool.resize(8192);
        for (auto &e : s->ev_pool) HIP_TRY(hipEventCreate(&e));
    }
    for (int f = 0; f < F_COUNT; ++f) { s->fam_ms[f] = 0.f; s->fam_n[f] = 0; }
    s->recs.clear();
    s->ev_next = 0;
    return BP_OK;
}

bp_status bp_solver_kernel_times(bp_solver *s, int max, char *names, float *total_ms, int *launches, int *count) {
    if (!s || !count) return fail(BP_ERR_INVALID_ARG, "NULL argument");
    bp_status st = collect_timing(s);
    if (st != BP_OK) return st;
    int c = 0;
    for (int f = 0; f < F_COUNT; ++f) {
        if (s->fam_n[f] == 0) continue;
        if (c < max) {
            if (names) std::snprintf(names + 32 * c, 32, "%s", kFamilyNames[f]);
            if (total_ms) total_ms[c] = s->fam_ms[f];
            if (launches) launches[c] = s->fam_n[f];
        }
        ++c;
    }
    *count = c;
    return BP_OK;
}

}  // extern "C"
