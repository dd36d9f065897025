// bp_kernels.h -- host launchers of the gfx950 kernels (one per pass family).
#pragma once
#include "bp_internal.h"

namespace bp {

// Kernels that depend on the scene take a parameter source SP: OneScene
// (plain solver, constants by value) or SceneTable (batched solver).

// bp_grid.hip
void launch_map_particle_infos(const float4 *x, const int *orig, const int *group, float4 *out, int n,
                               hipStream_t st);
void launch_first_nonfinite(const float4 *x, int n, int *first, hipStream_t st);
void launch_stream_copy(const float4 *src, float4 *dst, size_t n, int blocks, int variant, hipStream_t st);
template <class SP>
void launch_predict_hash(const float4 *x, const float4 *v, int *keys, int n, float dt, f3 gstep, const SP &sp,
                         const int *old_sorted_keys, int2 *ranges, hipStream_t st);
size_t radix_counts_len(int n);
void launch_radix_sort(int *keys_a, int *keys_b, int *vals_a, int *vals_b, int *counts, int *totals, int n, int bits,
                       int **out_keys, int **out_vals, hipStream_t st);
void launch_find_ranges(const int *sorted_keys, int n, int2 *ranges, hipStream_t st);
template <class SP>
void launch_reorder(const int *perm, const float4 *x_in, const float4 *v_in, const float *m_in, const int *ph_in,
                    const int *o_in, const int *sid_in, float4 *xs_out, float4 *x_out, float4 *v_out, float *m_out,
                    int *ph_out, int *o_out, int *sid_out, int *map_to_new, int n, float dt, f3 gstep, float mass_k,
                    int need_w, int stab_passes, int collide, const LambdaIn &L, const SP &sp, hipStream_t st);

// bp_fluid.hip (`block`: threads per workgroup of the neighbour pass, <= kMaxBlock).
// `srec` (kRecChunks x n candidate masks) is written by the lambda and
// omega/normal passes and replayed by the position pass of the same iteration
// and by the vorticity/tension and XSPH passes.  `staged` (lambda and
// omega/normal): 0 = per-lane walk, else the staged walk's region size in
// records per wave and column offset (one of kStageCaps); `stage_sb`: the
// staged walk's screen batch, 4 or 8 candidates.
template <class SP>
void launch_lambda(const float4 *xs, const float *m, const int2 *ranges, float *lambda, float *density,
                   float4 *lam_rec, uint32_t *srec, const LambdaIn &L, int lin_ready, int n, const SP &sp,
                   int block, int staged, int stage_sb, hipStream_t st);
template <class SP>
void launch_position(const float4 *xs, const float4 *lam_rec, const int *build_keys, const int2 *ranges,
                     const uint32_t *srec, const uint32_t *fluid_stamp, const uint32_t *stamp, float4 *sum_tmp,
                     unsigned long long *corr, uint8_t *dflag, float4 *xs_out, float4 *x_io,
                     float4 *v_out, int n, const SP &sp, const Epilogue &E, int block, int flat, int partners,
                     hipStream_t st);
template <class SP>
void launch_post_a(const float4 *xs, const float4 *v, const float *density, const int2 *ranges, float4 *omega,
                   float4 *xom, float4 *normal, uint32_t *ccell, uint32_t *srec, const LambdaIn *one_phase,
                   int planes_ready, int n, const SP &sp, int do_omega, int do_normal, int block, int staged,
                   int stage_sb, hipStream_t st);
template <class SP>
void launch_post_b(const float4 *xs, float4 *v, const float4 *omega, const float4 *xom, const float4 *normal,
                   const float *density, const int2 *ranges, const uint32_t *ccell, const uint32_t *srec, int n,
                   const SP &sp, int do_vort, int do_akinci, float dt, int block, int flat, hipStream_t st);
template <class SP>
void launch_xsph(const float4 *xs, const float4 *v, const int2 *ranges, const uint32_t *ccell, const uint32_t *srec,
                 float4 *v_out, int n, const SP &sp, int use_ccell, int block, hipStream_t st);

// bp_constraints.hip
template <class SP>
void launch_plane_collide(float4 *xs, const float4 *x, int n, const SP &sp, hipStream_t st);
template <class SP>
void launch_particle_particle(const float4 *xs, const float4 *x, const int2 *ranges, float4 *xs_out, int n,
                              const SP &sp, int block, int *contacts, const uint32_t *cell_stamp,
                              const uint32_t *stamp, const LambdaIn &L, const float *m, hipStream_t st);
// advances the substep stamp and marks the cells whose 27-cell block holds a
// solid (cell_stamp) and those holding a fluid particle (fluid_stamp)
template <class SP>
void launch_mark_cells(const int *sorted_keys, const int *ph, int n, uint32_t *stamp, uint32_t *cell_stamp,
                       uint32_t *fluid_stamp, const SP &sp, hipStream_t st);
void launch_distance(float4 *xs, const float4 *x, const int2 *pairs, const float2 *params, const int *map_to_new,
                     float4 *kbuf, int nc, const int *verts, const int *offs, const int *ents, int nv,
                     hipStream_t st);
void launch_bending(float4 *xs, const float4 *x, const int4 *bends, const int *map_to_new, float4 *bbuf, int nb,
                    float kb, const int *verts, const int *offs, const int *ents, int nv, hipStream_t st);
template <class SP>
void launch_wind(float4 *xs, const float4 *v, const float *m, const int4 *faces, const int *map_to_new,
                 float4 *fbuf, int nf, const SP &sp, float dt, const int *verts, const int *offs, const int *ents,
                 int nv, hipStream_t st);
// With npins > 0 the same launch also applies the pins (launch_immovable); the
// caller guarantees that no pinned particle belongs to a rigid body.
void launch_shape_matching(float4 *xs, const int *map_to_new, const float4 *rest, const int2 *ranges, float4 *rot,
                           float4 *cms, int nr, int iters, hipStream_t st, const float4 *pin_x = nullptr,
                           const int *pin_ids = nullptr, int npins = 0);
void launch_immovable(float4 *xs, const float4 *x, const int *ids, const int *map_to_new, int ni, hipStream_t st);
void launch_pick(float4 *pos, float4 *v, const int *map_to_new, int orig, f3 p, int keep_pos, f3 vel,
                 hipStream_t st);
void launch_finish(const float4 *xs, float4 *x, float4 *v, const float *vw, int n, float inv_dt, float eps2,
                   hipStream_t st);

// bp_adjoint.hip (adjoint passes; g* buffers hold adjoints in slot order)
void launch_adj_inverse(const int *orig, int *map_to_new, int n, hipStream_t st);
void launch_adj_seed(const float *seed, const int *orig, float4 *g, float scale, int n, hipStream_t st);
void launch_adj_unseed(const float4 *g, const int *orig, float *out, float inv_scale, int n, hipStream_t st);
template <class SP>
void launch_adj_cells(const float4 *xs, uint32_t *ccell, int n, const SP &sp, hipStream_t st);
template <class SP>
void launch_adj_xsph(const float4 *xf, const float4 *x1, const float4 *gx2, const float4 *gv3, const uint32_t *ccell,
                     const int *keys, const int2 *ranges, float4 *gv2_own, float4 *gx_own, float *cacc,
                     unsigned long long *corr, float4 *gxs, float4 *gx1, float *gw, float inv_dt, float eps2, int n,
                     const SP &sp, int block, hipStream_t st);
template <class SP>
void launch_adj_projection(const float4 *xs, const float *m, const float *lambda, float4 *prec, const int *keys,
                           const int2 *ranges, float4 *gxs, float4 *gx_own, float *glam_own,
                           unsigned long long *corr, const uint32_t *trec, uint32_t *srec, int n, const SP &sp,
                           int block, hipStream_t st);
template <class SP>
void launch_adj_planes(const float4 *xp, const float4 *x1, float4 *gxs, float4 *gx1, int n, const SP &sp,
                       hipStream_t st);
void launch_adj_immovable(float4 *gxs, float4 *gx1, const int *ids, const int *map_to_new, int ni, hipStream_t st);
void launch_adj_distance(const float4 *xd, const float4 *x1, const int2 *pairs, const float2 *params,
                         const int *map_to_new, float4 *gxs, float *gw, float4 *cb, float2 *wb, int nc,
                         const int *verts, const int *offs, const int *ents, int nv, hipStream_t st);
template <class SP>
void launch_adj_wind(const float4 *xsw, const float4 *v_in, const int *perm, const float *m, const int4 *faces,
                     const int *map_to_new, float4 *gxs, float4 *gv1, float4 *fbar, float4 *facc, int nf, const SP &sp,
                     float dt, const int *verts, const int *offs, const int *ents, int nv, hipStream_t st);
template <class SP>
void launch_adj_reorder(const int *perm, const float4 *x_in, const float4 *x1, const float4 *gxs, const float4 *gx1,
                        const float4 *gv1, const float *gw, float4 *gx_old, float4 *gv_old, int n, float dt,
                        float mass_k, int stab_passes, const SP &sp, hipStream_t st);
void launch_adj_scene_reduce(const float4 *facc, const float *cacc, const int *poff, const int *foff, float *out,
                             int nscenes, hipStream_t st);

}  // namespace bp
