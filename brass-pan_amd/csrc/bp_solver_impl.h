// bp_solver_impl.h -- the solver handle (struct bp_solver) and host helpers
// shared by the C ABI (bp_solver.hip) and the adjoint (bp_adjoint_host.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/brasspan.h"
#include "bp_kernels.h"

namespace bp_host {

bp_status fail(bp_status st, const std::string &msg);  // sets bp_last_error_string

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return fail(BP_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_));          \
    } while (0)

template <class T> struct DBuf {
    T *b[2] = {nullptr, nullptr};
    int c = 0;
    T *cur() const { return b[c]; }
    T *alt() const { return b[1 - c]; }
    void flip() { c = 1 - c; }
};

struct Csr {
    int nv = 0;
    int *verts = nullptr, *offs = nullptr, *ents = nullptr;
};

enum Family {
    F_PREDICT, F_SORT, F_RANGES, F_REORDER, F_WIND, F_PLANES, F_PP, F_LAMBDA, F_POSITION, F_DISTANCE,
    F_BENDING, F_SHAPE, F_IMMOVABLE, F_FINISH, F_POST_A, F_POST_B, F_XSPH, F_PICK, F_COUNT
};


inline bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
inline int ilog2(unsigned v) {
    int r = 0;
    while ((1u << r) < v) ++r;
    return r;
}

struct Adjoint;  // differentiable rollout state (bp_adjoint_host.hip)

}  // namespace bp_host

struct bp_solver {
    int device = 0;
    hipStream_t st = nullptr;
    bool own_stream = false;
    bp_params P{};
    int N = 0;
    int nscenes = 1;
    int num_cells = 0, sort_bits = 0;
    std::vector<bp::SceneConst> hsc;       // per-scene kernel constants
    bp::SceneConst *dsc = nullptr;         // device copy (batched solver)
    std::vector<int> poff;             // scene s owns original ids [poff[s], poff[s+1])
    std::vector<int> scene_of_orig;
    bool has_solid = false, has_fluid = false, any_planes = false;
    bool one_phase = false;  // every particle has the same fluid phase (the same-phase tests always pass)

    bp_host::DBuf<float4> xs, x, v;
    bp_host::DBuf<float> m;
    bp_host::DBuf<int> ph, orig, sid;
    int *map_to_new = nullptr;
    int *guard_slot = nullptr;  // the non-finite guard's answer (one int)

    int *keys[2] = {nullptr, nullptr};
    int *vals[2] = {nullptr, nullptr};
    int *counts = nullptr, *totals = nullptr;
    int sk_idx = -1;  // keys buffer holding the last sorted keys
    int *sorted_keys = nullptr, *perm = nullptr;
    int2 *ranges = nullptr;

    float *lambda = nullptr, *density = nullptr;
    float4 *lam_rec = nullptr;  // lambda pass -> position pass: (lambda, m, current cell, -) per slot
    bp::LambdaIn lin;           // the lambda pass's inputs besides x* (fluid scenes)
    uint32_t *ccell = nullptr;
    uint32_t *srec = nullptr;  // screen record of the lambda and omega/normal passes (kRecChunks x N masks)
    int *group = nullptr;      // render group id by original id (bp_scene_desc::group_ids, else 0)
    float4 *frame_buf = nullptr;  // bp_export_frame staging, allocated on first use
    float4 *sum_tmp = nullptr;
    unsigned long long *corr = nullptr;  // fixed-point pair corrections, kept zero between passes
    uint8_t *dflag = nullptr;  // slots that received a correction, one byte each (kept zero)
    float4 *omega = nullptr, *normal = nullptr;  // normal.w: the density it was computed with
    float4 *xom = nullptr;                        // (x, |omega|) from the omega/normal pass
    uint32_t *cell_stamp = nullptr, *stamp = nullptr;  // cells near a solid (fluid + solid scenes)
    uint32_t *fluid_stamp = nullptr;                    // cells holding a fluid particle (same scenes)

    int nd = 0;
    int2 *dpairs = nullptr;
    float2 *dparams = nullptr;
    float4 *kbuf = nullptr;
    bp_host::Csr dcsr;
    int nb = 0;
    int4 *bends = nullptr;
    float4 *bbuf = nullptr;
    bp_host::Csr bcsr;
    int nf = 0;
    int4 *faces = nullptr;
    float4 *fbuf = nullptr;
    bp_host::Csr fcsr;
    int nr = 0;
    int2 *rranges = nullptr;
    float4 *rrest = nullptr, *rrot = nullptr, *rcm = nullptr;
    int ni = 0;
    int *immov = nullptr;
    bool pins_apart_from_rigid = false;  // no pinned particle lies in a rigid body

    std::vector<float> mass_by_orig;
    std::vector<int> phase_by_orig;
    std::vector<void *> allocs;

    // neighbour-pass workgroup sizes (bp_solver_autotune)
    int blk_lambda = 128, blk_position = 128, blk_post_a = 128, blk_post_b = 128, blk_xsph = 128, blk_pp = 128;
    // walk forms (bp_solver_set_walk, bp_solver_autotune): the lambda and
    // omega/normal passes' candidates staged in an LDS region of that many
    // records per wave and column offset (128, 256 or 512) or read from
    // global memory (0); the position and vorticity/tension passes flat (1,
    // bodies over the whole walk) or per chunk (0)
    int staged_lambda = bp::kStageCapDefault, flat_position = 1, staged_post_a = bp::kStageCapDefault,
        flat_post_b = 1;
    // partners the position pass fetches per round trip (2 or 4)
    int pos_partners = 2;
    // candidates per screen batch of the staged lambda and omega/normal walks
    // (4 or 8; bp_solver_set_stage_screen)
    int stage_sb_lambda = 4, stage_sb_post_a = 4;

    // hipGraph of two substeps (the ping-pong buffers return to their start)
    bool use_graph = false;
    bool fuse = true;  // iteration fusion: epilogue passes inside the producing kernels
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    float graph_dt = 0.f;
    int graph_sig = -1;  // buffer parities the graph was captured with (buffer_sig)

    // per-family kernel timing (bp_solver_enable_kernel_timing)
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_next = 0;
    struct Rec {
        int fam;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    float fam_ms[bp_host::F_COUNT] = {};
    int fam_n[bp_host::F_COUNT] = {};

    template <class T> bp_status alloc(T **p, size_t count) {
        *p = nullptr;
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void **)p, count * sizeof(T));
        if (e != hipSuccess) return bp_host::fail(BP_ERR_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        allocs.push_back((void *)*p);
        return BP_OK;
    }
    template <class T> bp_status upload(T **p, const T *host, size_t count) {
        bp_status s = alloc(p, count);
        if (s != BP_OK) return s;
        if (count && host) {
            const hipError_t e = hipMemcpyAsync(*p, host, count * sizeof(T), hipMemcpyHostToDevice, st);
            if (e != hipSuccess) return bp_host::fail(BP_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
        }
        return BP_OK;
    }

    void tbeg(int fam, hipEvent_t *out) {
        *out = nullptr;
        if (!timing) return;
        if (ev_next + 2 > ev_pool.size()) return;
        *out = ev_pool[ev_next];
        (void)hipEventRecord(*out, st);
        ev_next += 1;
        (void)fam;
    }
    void tend(int fam, hipEvent_t a) {
        if (!timing || !a) return;
        hipEvent_t b = ev_pool[ev_next++];
        (void)hipEventRecord(b, st);
        recs.push_back(Rec{fam, a, b});
    }
    // which half of every ping-pong pair is current: a captured graph holds
    // raw pointers, so it replays only from the parities it was captured at
    int buffer_sig() const { return x.c | xs.c << 1 | v.c << 2 | (sk_idx + 1) << 3; }
    // checkpoints, tape and adjoint buffers of bp_forward_checkpointed / bp_backward
    bp_host::Adjoint *adj = nullptr;

    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }
};

#define TIMED(fam, stmt)      \
    do {                      \
        hipEvent_t ev_;       \
        s->tbeg(fam, &ev_);   \
        stmt;                 \
        s->tend(fam, ev_);    \
    } while (0)

namespace bp_host {

// Snapshot destinations of one recorded substep (nullptr: not recorded).
struct SubstepHooks {
    float4 *x_in = nullptr, *v_in = nullptr;  // state entering the substep (old slot order)
    int *keys = nullptr, *perm = nullptr;     // sorted build cells, permutation
    float4 *x1 = nullptr, *xsw = nullptr;     // x (w = inverse scaled mass) and x* after reorder
    float *m = nullptr;
    int *orig = nullptr, *sid = nullptr;
    float4 **xp = nullptr, **xl = nullptr, **xd = nullptr;  // per iteration: before planes, before lambda, before distance
    float **lam = nullptr;                    // per iteration
    float4 *xf = nullptr;                     // x* after the last iteration
    int *contacts = nullptr;                  // counts particles with a particle-particle contact
    uint32_t **rec = nullptr;                 // per iteration: where the lambda pass writes its screen record
};

// One substep of ParticleSolver::update.  `no_fuse` keeps every pass a
// separate kernel (the form the adjoint records); `hk` snapshots state.
bp_status substep(bp_solver *s, float subdt, float inv_dt, bp::f3 gstep, bool pick, int picked, bp::f3 ppos,
                  bool no_fuse = false, const SubstepHooks *hk = nullptr);
bp::f3 gravity_step(const bp_params &P, float subdt);
void adjoint_release(bp_solver *s);
// a recorded rollout no longer matches the solver state (bp_backward refuses it)
void adjoint_invalidate(bp_solver *s);
// largest |term| the last bp_backward delivered in fixed point (seed units; 0 if none ran)
float adjoint_max_delivered(const bp_solver *s);

}  // namespace bp_host
