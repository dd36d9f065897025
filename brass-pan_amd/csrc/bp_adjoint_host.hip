// bp_adjoint_host.hip -- differentiable rollout of the particle solver:
// checkpointed forward, recorded replay of each checkpoint segment and the
// reverse sweep through the adjoint kernels (bp_adjoint.hip).
//
// bp_forward_checkpointed keeps the full particle state (x, v, mass, phase,
// slot order) entering every `every`-th substep, and records the last
// segment's substeps on the tape as it runs them.  bp_backward walks the
// segments last to first: unless the tape already holds a segment, it restores
// the segment's checkpoint and replays its substeps with the same kernels
// (bit-identical), recording the state before each pass on the tape; then it
// runs the adjoint of those substeps in reverse.  Memory: one segment of tape
// (about 300 B per particle per substep at 4 iterations, plus the lambda
// passes' screen records, 64 B per particle per iteration) and the
// checkpoints.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>
#include <vector>

#include "bp_solver_impl.h"

using namespace bp;

namespace bp_host {

struct Adjoint {
    struct State {  // solver state entering a substep
        float4 *x = nullptr, *v = nullptr;
        float *m = nullptr;
        int *ph = nullptr, *orig = nullptr, *sid = nullptr;
    };
    struct Slot {  // recorded forward of one substep
        float4 *x_in = nullptr, *v_in = nullptr, *x1 = nullptr, *xsw = nullptr, *xf = nullptr;
        int *keys = nullptr, *perm = nullptr, *orig = nullptr, *sid = nullptr;
        float *m = nullptr;
        std::vector<float4 *> xp, xl, xd;
        std::vector<float *> lam;
        std::vector<uint32_t *> rec;  // per iteration: the lambda pass's screen record of xl
    };
    int every = 0, total = 0, iters = 0;
    float subdt = 0.f;
    bool recorded = false;
    std::vector<State> cps;  // cps[c]: entering substep c * every
    State last;               // after the rollout
    std::vector<Slot> tape;
    int tape_seg = -1;  // the segment whose substeps the tape holds (-1: none)
    // adjoint buffers (slot order of the substep being reversed)
    float4 *gX[2] = {nullptr, nullptr}, *gV[2] = {nullptr, nullptr};
    float4 *gxs = nullptr, *gx1 = nullptr, *gv1 = nullptr, *gown = nullptr;
    float *gw = nullptr, *glam = nullptr, *cacc = nullptr;
    float4 *prec = nullptr;  // projection adjoints' record: (lambda | density adjoint, m, cell, -)
    uint32_t *ccell = nullptr;
    int *mtn = nullptr, *contacts = nullptr;
    unsigned long long *corr = nullptr;  // 6 delivered components per slot, then 2 guard words
    float max_delivered = 0.f;            // largest |delivered term| of the last backward, seed units
    float4 *cb = nullptr;
    float2 *wb = nullptr;
    float4 *fbar = nullptr, *facc = nullptr;
    float *seed = nullptr, *outbuf = nullptr, *theta = nullptr;
    int *poff = nullptr, *foff = nullptr;
    Csr fcsr_c;  // faces per vertex, entries 3 * face + corner
    std::vector<void *> allocs;

    template <class T> bp_status alloc(T **p, size_t count) {
        *p = nullptr;
        if (count == 0) count = 1;
        const hipError_t e = hipMalloc((void **)p, count * sizeof(T));
        if (e != hipSuccess) return fail(BP_ERR_HIP, std::string("adjoint hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back((void *)*p);
        return BP_OK;
    }
    bp_status alloc_state(State &st, int n, bool batched) {
        bp_status r;
        if ((r = alloc(&st.x, n)) != BP_OK || (r = alloc(&st.v, n)) != BP_OK || (r = alloc(&st.m, n)) != BP_OK ||
            (r = alloc(&st.ph, n)) != BP_OK || (r = alloc(&st.orig, n)) != BP_OK)
            return r;
        if (batched && (r = alloc(&st.sid, n)) != BP_OK) return r;
        return BP_OK;
    }
};

void adjoint_invalidate(bp_solver *s) {
    if (s && s->adj) s->adj->recorded = false;
}

float adjoint_max_delivered(const bp_solver *s) { return (s && s->adj) ? s->adj->max_delivered : 0.0f; }

void adjoint_release(bp_solver *s) {
    if (!s || !s->adj) return;
    for (void *p : s->adj->allocs) (void)hipFree(p);
    delete s->adj;
    s->adj = nullptr;
}

namespace {

bp_status supported(const bp_solver *s) {
    const bp_params &P = s->P;
    if (s->N <= 0) return fail(BP_ERR_INVALID_ARG, "no particles");
    if (s->nr > 0) return fail(BP_ERR_UNSUPPORTED, "the adjoint does not differentiate shape matching (rigid bodies)");
    if (s->nb > 0) return fail(BP_ERR_UNSUPPORTED, "the adjoint does not differentiate bending constraints");
    if (P.akinci || P.vorticity_eps != 0.0f)
        return fail(BP_ERR_UNSUPPORTED, "the adjoint does not differentiate vorticity confinement or Akinci tension");
    int maxp = 0;
    for (const SceneConst &c : s->hsc) {
        if (c.F.lambda_mode != 0) return fail(BP_ERR_UNSUPPORTED, "the adjoint needs lambda_mode 0");
        maxp = std::max(maxp, c.pl.n);
    }
    if (P.stabilization_passes * maxp > 64)
        return fail(BP_ERR_UNSUPPORTED, "stabilization_passes * planes must be <= 64 for the adjoint");
    return BP_OK;
}

bp_status copy_state(bp_solver *s, Adjoint::State &dst, bool save) {
    const size_t N = s->N;
    hipStream_t st = s->st;
    auto cp = [&](void *to, const void *from, size_t bytes) {
        return hipMemcpyAsync(to, from, bytes, hipMemcpyDeviceToDevice, st);
    };
    if (save) {
        HIP_TRY(cp(dst.x, s->x.cur(), N * sizeof(float4)));
        HIP_TRY(cp(dst.v, s->v.cur(), N * sizeof(float4)));
        HIP_TRY(cp(dst.m, s->m.cur(), N * sizeof(float)));
        HIP_TRY(cp(dst.ph, s->ph.cur(), N * sizeof(int)));
        HIP_TRY(cp(dst.orig, s->orig.cur(), N * sizeof(int)));
        if (dst.sid) HIP_TRY(cp(dst.sid, s->sid.cur(), N * sizeof(int)));
        return BP_OK;
    }
    HIP_TRY(cp(s->x.cur(), dst.x, N * sizeof(float4)));
    HIP_TRY(cp(s->v.cur(), dst.v, N * sizeof(float4)));
    HIP_TRY(cp(s->m.cur(), dst.m, N * sizeof(float)));
    HIP_TRY(cp(s->ph.cur(), dst.ph, N * sizeof(int)));
    HIP_TRY(cp(s->orig.cur(), dst.orig, N * sizeof(int)));
    if (dst.sid) HIP_TRY(cp(s->sid.cur(), dst.sid, N * sizeof(int)));
    launch_adj_inverse(s->orig.cur(), s->map_to_new, s->N, st);
    // the cell table is rebuilt from scratch by the next substep
    HIP_TRY(hipMemsetAsync(s->ranges, 0xFF, (size_t)s->num_cells * sizeof(int2), st));
    s->sk_idx = -1;
    s->drop_graph();
    return BP_OK;
}

bp_status ensure(bp_solver *s, int total, int every) {
    if (s->adj && s->adj->iters != s->P.iterations && !s->adj->tape.empty()) adjoint_release(s);
    if (!s->adj) s->adj = new Adjoint();
    Adjoint &A = *s->adj;
    const int N = s->N;
    const bool batched = s->nscenes > 1;
    const int iters = s->P.iterations;
    const int ncp = (total + every - 1) / every;
    bp_status r;
    if (!A.gX[0]) {  // buffers that do not depend on the rollout length
        for (int k = 0; k < 2; ++k) {
            if ((r = A.alloc(&A.gX[k], N)) != BP_OK || (r = A.alloc(&A.gV[k], N)) != BP_OK) return r;
        }
        if ((r = A.alloc(&A.gxs, N)) != BP_OK || (r = A.alloc(&A.gx1, N)) != BP_OK ||
            (r = A.alloc(&A.gv1, N)) != BP_OK || (r = A.alloc(&A.gown, N)) != BP_OK ||
            (r = A.alloc(&A.gw, N)) != BP_OK || (r = A.alloc(&A.glam, N)) != BP_OK ||
            (r = A.alloc(&A.prec, N)) != BP_OK || (r = A.alloc(&A.cacc, N)) != BP_OK ||
            (r = A.alloc(&A.ccell, N)) != BP_OK || (r = A.alloc(&A.mtn, N)) != BP_OK ||
            (r = A.alloc(&A.contacts, 1)) != BP_OK || (r = A.alloc(&A.corr, 6 * (size_t)N + 2)) != BP_OK ||
            (r = A.alloc(&A.cb, s->nd)) != BP_OK || (r = A.alloc(&A.wb, s->nd)) != BP_OK ||
            (r = A.alloc(&A.fbar, 4 * (size_t)s->nf)) != BP_OK || (r = A.alloc(&A.facc, 2 * (size_t)s->nf)) != BP_OK ||
            (r = A.alloc(&A.seed, 3 * (size_t)N)) != BP_OK || (r = A.alloc(&A.outbuf, 3 * (size_t)N)) != BP_OK ||
            (r = A.alloc(&A.theta, 6 * (size_t)s->nscenes)) != BP_OK)
            return r;
        HIP_TRY(hipMemsetAsync(A.corr, 0, (6 * (size_t)N + 2) * sizeof(unsigned long long), s->st));
        if ((r = A.alloc_state(A.last, N, batched)) != BP_OK) return r;
        // per-scene slot and face ranges
        std::vector<int> foff(s->nscenes + 1, 0);
        {
            std::vector<int> hf(4 * (size_t)s->nf);
            if (s->nf > 0)
                HIP_TRY(hipMemcpy(hf.data(), s->faces, 4 * (size_t)s->nf * sizeof(int), hipMemcpyDeviceToHost));
            for (int f = 0; f < s->nf; ++f) foff[s->scene_of_orig[hf[4 * f]] + 1]++;
            for (int k = 0; k < s->nscenes; ++k) foff[k + 1] += foff[k];
            // faces (x, y, z corners) per vertex, entries 3 f + corner, in face order
            std::vector<int> ids(3 * (size_t)s->nf);
            for (int f = 0; f < s->nf; ++f)
                for (int c = 0; c < 3; ++c) ids[3 * f + c] = hf[4 * f + c];
            std::vector<int> cnt(N, 0);
            for (int id : ids) cnt[id]++;
            std::vector<int> verts, offs(1, 0), pos(N, -1);
            for (int p = 0; p < N; ++p)
                if (cnt[p] > 0) {
                    pos[p] = (int)verts.size();
                    verts.push_back(p);
                    offs.push_back(offs.back() + cnt[p]);
                }
            std::vector<int> fill(offs.begin(), offs.end() - 1), ents(ids.size());
            for (size_t e = 0; e < ids.size(); ++e) ents[fill[pos[ids[e]]]++] = (int)e;
            A.fcsr_c.nv = (int)verts.size();
            if ((r = A.alloc(&A.fcsr_c.verts, verts.size())) != BP_OK ||
                (r = A.alloc(&A.fcsr_c.offs, offs.size())) != BP_OK ||
                (r = A.alloc(&A.fcsr_c.ents, ents.size())) != BP_OK)
                return r;
            if (!verts.empty()) {
                HIP_TRY(hipMemcpy(A.fcsr_c.verts, verts.data(), verts.size() * sizeof(int), hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy(A.fcsr_c.ents, ents.data(), ents.size() * sizeof(int), hipMemcpyHostToDevice));
            }
            HIP_TRY(hipMemcpy(A.fcsr_c.offs, offs.data(), offs.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        if ((r = A.alloc(&A.poff, s->nscenes + 1)) != BP_OK || (r = A.alloc(&A.foff, s->nscenes + 1)) != BP_OK)
            return r;
        HIP_TRY(hipMemcpy(A.poff, s->poff.data(), (s->nscenes + 1) * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(A.foff, foff.data(), (s->nscenes + 1) * sizeof(int), hipMemcpyHostToDevice));
        // pins must be distinct (the adjoint of a repeated pin would count it twice)
        if (s->ni > 0) {
            std::vector<int> im(s->ni);
            HIP_TRY(hipMemcpy(im.data(), s->immov, s->ni * sizeof(int), hipMemcpyDeviceToHost));
            if (std::set<int>(im.begin(), im.end()).size() != im.size())
                return fail(BP_ERR_UNSUPPORTED, "the adjoint needs distinct pinned particles");
        }
    }
    while ((int)A.cps.size() < ncp) {
        A.cps.emplace_back();
        if ((r = A.alloc_state(A.cps.back(), N, batched)) != BP_OK) return r;
    }
    {
        while ((int)A.tape.size() < every) {
            A.tape.emplace_back();
            Adjoint::Slot &T = A.tape.back();
            if ((r = A.alloc(&T.x_in, N)) != BP_OK || (r = A.alloc(&T.v_in, N)) != BP_OK ||
                (r = A.alloc(&T.x1, N)) != BP_OK || (r = A.alloc(&T.xsw, N)) != BP_OK ||
                (r = A.alloc(&T.xf, (size_t)N + kWalkPad)) != BP_OK || (r = A.alloc(&T.keys, N)) != BP_OK ||
                (r = A.alloc(&T.perm, N)) != BP_OK || (r = A.alloc(&T.orig, N)) != BP_OK ||
                (r = A.alloc(&T.m, N)) != BP_OK)
                return r;
            if (batched && (r = A.alloc(&T.sid, N)) != BP_OK) return r;
            T.xp.assign(iters, nullptr);
            T.xl.assign(iters, nullptr);
            T.xd.assign(iters, nullptr);
            T.lam.assign(iters, nullptr);
            T.rec.assign(iters, nullptr);
            for (int it = 0; it < iters; ++it)
                if ((r = A.alloc(&T.xp[it], N)) != BP_OK || (r = A.alloc(&T.xl[it], (size_t)N + kWalkPad)) != BP_OK ||
                    (r = A.alloc(&T.xd[it], N)) != BP_OK || (r = A.alloc(&T.lam[it], N)) != BP_OK ||
                    (s->has_fluid && (r = A.alloc(&T.rec[it], (size_t)kRecChunks * N)) != BP_OK))
                    return r;
        }
    }
    A.iters = iters;
    return BP_OK;
}

// hooks that record substep t of a segment into tape slot t
SubstepHooks tape_hooks(Adjoint &A, int t) {
    Adjoint::Slot &T = A.tape[t];
    SubstepHooks H;
    H.x_in = T.x_in;
    H.v_in = T.v_in;
    H.keys = T.keys;
    H.perm = T.perm;
    H.x1 = T.x1;
    H.xsw = T.xsw;
    H.m = T.m;
    H.orig = T.orig;
    H.sid = T.sid;
    H.xp = T.xp.data();
    H.xl = T.xl.data();
    H.xd = T.xd.data();
    H.lam = T.lam.data();
    H.rec = T.rec.data();
    H.xf = T.xf;
    H.contacts = A.contacts;
    return H;
}

template <class SP>
void reverse_substep(bp_solver *s, Adjoint &A, const Adjoint::Slot &T, int &cur, const SP &sp) {
    const int N = s->N;
    const bp_params &P = s->P;
    hipStream_t st = s->st;
    const float dt = A.subdt, inv_dt = 1.0f / A.subdt;
    const bool wind_on = P.wind_enabled && s->nf > 0;
    (void)hipMemsetAsync(s->ranges, 0xFF, (size_t)s->num_cells * sizeof(int2), st);
    launch_find_ranges(T.keys, N, s->ranges, st);
    launch_adj_inverse(T.orig, A.mtn, N, st);

    // XSPH + velocity/position update
    launch_adj_cells(T.xf, A.ccell, N, sp, st);
    launch_adj_xsph(T.xf, T.x1, A.gX[cur], A.gV[cur], A.ccell, T.keys, s->ranges, A.gv1, A.gown, A.cacc, A.corr,
                    A.gxs, A.gx1, A.gw, inv_dt, P.sleep_epsilon * P.sleep_epsilon, N, sp, s->blk_xsph, st);
    for (int it = P.iterations - 1; it >= 0; --it) {
        if (s->ni > 0) launch_adj_immovable(A.gxs, A.gx1, s->immov, A.mtn, s->ni, st);
        if (s->nd > 0)
            launch_adj_distance(T.xd[it], T.x1, s->dpairs, s->dparams, A.mtn, A.gxs, A.gw, A.cb, A.wb, s->nd,
                                s->dcsr.verts, s->dcsr.offs, s->dcsr.ents, s->dcsr.nv, st);
        if (s->has_fluid) {
            launch_adj_projection(T.xl[it], T.m, T.lam[it], A.prec, T.keys, s->ranges, A.gxs, A.gown, A.glam,
                                  A.corr, T.rec[it], s->srec, N, sp, s->blk_position, st);
        }
        if (s->any_planes) launch_adj_planes(T.xp[it], T.x1, A.gxs, A.gx1, N, sp, st);
    }
    (void)hipMemsetAsync(A.gv1, 0, (size_t)N * sizeof(float4), st);
    if (wind_on)
        launch_adj_wind(T.xsw, T.v_in, T.perm, T.m, s->faces, A.mtn, A.gxs, A.gv1, A.fbar, A.facc, s->nf, sp, dt,
                        A.fcsr_c.verts, A.fcsr_c.offs, A.fcsr_c.ents, A.fcsr_c.nv, st);
    const bool need_w = s->has_solid || s->nd > 0 || s->nb > 0;
    launch_adj_reorder(T.perm, T.x_in, T.x1, A.gxs, A.gx1, A.gv1, A.gw, A.gX[1 - cur], A.gV[1 - cur], N, dt,
                       need_w ? P.mass_scaling_k : 0.0f, P.stabilization_passes, sp, st);
    cur = 1 - cur;
}

}  // namespace
}  // namespace bp_host

using namespace bp_host;

extern "C" {

bp_status bp_forward_checkpointed(bp_solver *s, int substeps, float subdt, int checkpoint_every) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    if (substeps <= 0 || checkpoint_every <= 0) return fail(BP_ERR_INVALID_ARG, "substeps and checkpoint_every must be > 0");
    if (!(subdt > 0.0f)) return fail(BP_ERR_INVALID_ARG, "subdt must be > 0");
    bp_status r = supported(s);
    if (r != BP_OK) return r;
    if ((r = ensure(s, substeps, checkpoint_every)) != BP_OK) return r;
    Adjoint &A = *s->adj;
    A.recorded = false;
    A.every = checkpoint_every;
    A.total = substeps;
    A.subdt = subdt;
    A.tape_seg = -1;
    s->drop_graph();
    const float inv_dt = 1.0f / subdt;
    const f3 gstep = gravity_step(s->P, subdt);
    // contacts are counted over the whole rollout (bp_backward refuses it if
    // any occurred); the last segment is recorded on the tape as it runs
    HIP_TRY(hipMemsetAsync(A.contacts, 0, sizeof(int), s->st));
    const int last_seg = (substeps - 1) / checkpoint_every;
    for (int k = 0; k < substeps; ++k) {
        if (k % checkpoint_every == 0 && (r = copy_state(s, A.cps[k / checkpoint_every], true)) != BP_OK) return r;
        SubstepHooks H;
        if (k / checkpoint_every == last_seg)
            H = tape_hooks(A, k - last_seg * checkpoint_every);
        else
            H.contacts = A.contacts;
        if ((r = substep(s, subdt, inv_dt, gstep, false, -1, mk3(0.f, 0.f, 0.f), true, &H)) != BP_OK) return r;
    }
    if ((r = copy_state(s, A.last, true)) != BP_OK) return r;
    HIP_TRY(hipGetLastError());
    A.tape_seg = last_seg;
    A.recorded = true;
    return BP_OK;
}

bp_status bp_backward(bp_solver *s, const float *grad_x, const float *grad_v, float *grad_x0, float *grad_v0,
                      float *grad_theta) {
    if (!s) return fail(BP_ERR_INVALID_ARG, "solver is NULL");
    if (!s->adj || !s->adj->recorded) return fail(BP_ERR_INVALID_ARG, "no recorded rollout (bp_forward_checkpointed)");
    Adjoint &A = *s->adj;
    const int N = s->N;
    hipStream_t st = s->st;
    // the adjoint is linear in the seeds: scale them to unit maximum so the
    // fixed-point deliveries keep their resolution, and undo it at the end
    double mx = 0.0;
    for (size_t i = 0; i < 3 * (size_t)N; ++i) {
        if (grad_x) mx = std::max(mx, (double)std::fabs(grad_x[i]));
        if (grad_v) mx = std::max(mx, (double)std::fabs(grad_v[i]));
    }
    if (!std::isfinite(mx)) return fail(BP_ERR_NOT_FINITE, "non-finite seed");
    const float scale = (mx > 0.0) ? (float)(1.0 / mx) : 1.0f;
    const int *orig_T = A.last.orig;
    if (grad_x) HIP_TRY(hipMemcpyAsync(A.seed, grad_x, 3 * (size_t)N * sizeof(float), hipMemcpyHostToDevice, st));
    launch_adj_seed(grad_x ? A.seed : nullptr, orig_T, A.gX[0], scale, N, st);
    HIP_TRY(hipStreamSynchronize(st));
    if (grad_v) HIP_TRY(hipMemcpyAsync(A.seed, grad_v, 3 * (size_t)N * sizeof(float), hipMemcpyHostToDevice, st));
    launch_adj_seed(grad_v ? A.seed : nullptr, orig_T, A.gV[0], scale, N, st);
    HIP_TRY(hipMemsetAsync(A.cacc, 0, (size_t)N * sizeof(float), st));
    if (s->nf > 0) HIP_TRY(hipMemsetAsync(A.facc, 0, 2 * (size_t)s->nf * sizeof(float4), st));
    // guard words of the fixed-point deliveries: overflow count, largest |term|
    HIP_TRY(hipMemsetAsync(A.corr + 6 * (size_t)N, 0, 2 * sizeof(unsigned long long), st));

    const float inv_dt = 1.0f / A.subdt;
    const f3 gstep = gravity_step(s->P, A.subdt);
    int cur = 0;
    const int nseg = (A.total + A.every - 1) / A.every;
    bp_status r = BP_OK;
    for (int seg = nseg - 1; seg >= 0 && r == BP_OK; --seg) {
        const int k0 = seg * A.every;
        const int cnt = std::min(A.every, A.total - k0);
        if (seg != A.tape_seg) {  // replay the segment from its checkpoint onto the tape
            A.tape_seg = -1;
            if ((r = copy_state(s, A.cps[seg], false)) != BP_OK) break;
            for (int t = 0; t < cnt && r == BP_OK; ++t) {
                const SubstepHooks H = tape_hooks(A, t);
                r = substep(s, A.subdt, inv_dt, gstep, false, -1, mk3(0.f, 0.f, 0.f), true, &H);
            }
            if (r == BP_OK) A.tape_seg = seg;
        }
        for (int t = cnt - 1; t >= 0 && r == BP_OK; --t) {
            const Adjoint::Slot &T = A.tape[t];
            if (s->nscenes == 1)
                reverse_substep(s, A, T, cur, OneScene{s->hsc[0]});
            else
                reverse_substep(s, A, T, cur, SceneTable{s->dsc, T.sid});
        }
    }
    // leave the solver at the end of the rollout
    bp_status r2 = copy_state(s, A.last, false);
    if (r != BP_OK) return r;
    if (r2 != BP_OK) return r2;
    HIP_TRY(hipGetLastError());
    int contacts = 0;
    unsigned long long guard[2] = {0ull, 0ull};
    HIP_TRY(hipMemcpyAsync(&contacts, A.contacts, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(guard, A.corr + 6 * (size_t)N, sizeof(guard), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    {
        const uint32_t bits = (uint32_t)guard[1];
        float mx_term;
        std::memcpy(&mx_term, &bits, sizeof(float));
        A.max_delivered = mx_term / scale;
    }
    if (guard[0] > 0)
        return fail(BP_ERR_NOT_FINITE, "adjoint term out of the fixed-point delivery range (" +
                                           std::to_string(guard[0]) + " terms above 2^17 after seed scaling)");
    if (contacts > 0)
        return fail(BP_ERR_UNSUPPORTED, "particle-particle contact occurred in the rollout; the adjoint does not "
                                        "differentiate contact (" + std::to_string(contacts) + " particle passes)");
    const float inv_scale = 1.0f / scale;
    const int *orig_0 = A.cps[0].orig;
    if (grad_x0) {
        launch_adj_unseed(A.gX[cur], orig_0, A.outbuf, inv_scale, N, st);
        HIP_TRY(hipMemcpyAsync(grad_x0, A.outbuf, 3 * (size_t)N * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    if (grad_v0) {
        launch_adj_unseed(A.gV[cur], orig_0, A.outbuf, inv_scale, N, st);
        HIP_TRY(hipMemcpyAsync(grad_v0, A.outbuf, 3 * (size_t)N * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    if (grad_theta) {
        launch_adj_scene_reduce(A.facc, A.cacc, A.poff, A.foff, A.theta, s->nscenes, st);
        std::vector<float> th(6 * (size_t)s->nscenes);
        HIP_TRY(hipMemcpyAsync(th.data(), A.theta, th.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (size_t i = 0; i < th.size(); ++i) grad_theta[i] = th[i] * inv_scale;
    }
    HIP_TRY(hipGetLastError());
    return BP_OK;
}

}  // extern "C"
