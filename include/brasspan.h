/*
 * brasspan.h -- C ABI of the MI355X (gfx950) particle solver substep.
 *
 * This is the drop-in boundary for brass-pan's `ParticleSolver` / `Scene`
 * pair.  Every entry point below replaces one member of the reference's
 * `struct ParticleSolver` (brasstower/particlesolver.cuh) and is called by
 * the C++ shim in brass-pan_amd/cpp/brasspan/particlesolver.h, which keeps the
 * reference's class and method names.  Plain pointers and sizes only: no HIP,
 * torch or glm types cross this boundary (the stream is passed as void*).
 *
 * Conventions
 *  - Vectors are packed float[3] (or int[2] / int[3] / int[4]) in host memory.
 *  - "slot" = index of a particle in the solver's current sorted order (the
 *    reference's "new id"); "original id" = index in the scene description.
 *  - Every call returns bp_status; on failure bp_last_error_string() gives a
 *    message.  The reference prints and continues (cuda/helper.cuh:282-290)
 *    or throws (particlesolver.cuh:398-402); this ABI reports instead.
 *  - One host thread per solver handle.  Different handles are independent.
 */
#ifndef BRASSPAN_H_
#define BRASSPAN_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BP_ABI_VERSION 1

typedef enum bp_status {
    BP_OK = 0,
    BP_ERR_INVALID_ARG = 1,   /* bad pointer, size or parameter           */
    BP_ERR_CAPACITY = 2,      /* e.g. rigid body with more than 64 particles */
    BP_ERR_HIP = 3,           /* a HIP runtime call failed                */
    BP_ERR_NOT_FINITE = 4,    /* NaN/Inf found by the optional guard      */
    BP_ERR_NO_DEVICE = 5,     /* no gfx950 device visible                 */
    BP_ERR_UNSUPPORTED = 6
} bp_status;

/*
 * Solver parameters.  bp_default_params() fills every field with the
 * literal the reference hard-codes; the comment gives where.
 */
typedef struct bp_params {
    int   iterations;              /* 2: particlesolver.cuh:862                  */
    int   stabilization_passes;    /* 2: particlesolver.cuh:822                  */
    float gravity[3];              /* (0,-9.8,0): particlesolver.cuh:56          */
    float fluid_rest_density;      /* OldSceneFormat::fluidRestDensity (:338)    */
    float solid_density_scale;     /* 2.0: particlesolver.cuh:909                */
    int   lambda_mode;             /* 0 = constant denominator (fluid.cuh:76,153),
                                      1 = -C/(sum|grad C|^2 + eps) (fluid.cuh:155) */
    float lambda_denominator;      /* 15000: fluid.cuh:153                       */
    float lambda_relaxation;       /* 300 (eps of mode 1): particlesolver.cuh:910 */
    int   akinci;                  /* 1: particlesolver.cuh:785 (clamps C>=0, turns
                                      sCorr off, runs normal + tension passes)   */
    float scorr_k;                 /* 1e-4: particlesolver.cuh:926               */
    int   scorr_n;                 /* 4:    particlesolver.cuh:927               */
    float scorr_dq;                /* 0.03 (x h): fluid.cuh:208                  */
    float delta_clamp;             /* 5.0: fluid.cuh:214                         */
    float vorticity_eps;           /* 0.001: particlesolver.cuh:1017 (0 = skip)  */
    float tension_gamma;           /* 0.5: particlesolver.cuh:1043               */
    float xsph_c;                  /* 0.0002: particlesolver.cuh:1055 (0 = skip) */
    float friction_static;         /* 0.5: cubrasstower.cuh:18                   */
    float friction_dynamic;        /* 0.0: cubrasstower.cuh:19                   */
    float mass_scaling_k;          /* 2: cubrasstower.cuh:20                     */
    float sleep_epsilon;           /* 0: cubrasstower.cuh:21                     */
    int   wind_enabled;            /* 1: particlesolver.cuh:838-858              */
    float wind[3];                 /* (0,0,0): the reference has no wind vector  */
    float drag_coeff;              /* 10: wind.cuh:40                            */
    float lift_coeff;              /* 100: wind.cuh:43                           */
    float bend_stiffness;          /* 0.8: bending.cuh:65                        */
    int   shape_matching_iterations; /* 20: shapematching.cuh:72                */
    int   nan_guard;               /* 0: optional non-finite check after a step
                                      (reference CHECK_NAN, cubrasstower.cuh:15) */
} bp_params;

/*
 * Scene description.  Replaces `struct Scene` (scene.h:455-536) on the C side.
 * Host pointers; the solver copies everything at create time and keeps no
 * pointer afterwards (the reference copies once too, particlesolver.cuh:158-214).
 */
typedef struct bp_scene_desc {
    int          num_particles;
    const float *positions;        /* [3N] scene.h:507                          */
    const float *velocities;       /* [3N] or NULL = zero (particlesolver.cuh:167) */
    const float *masses;           /* [N]  scene.h:508; each >= 0 (else       */
                                   /*      BP_ERR_INVALID_ARG)                  */
    const int   *phases;           /* [N]  scene.h:509: fluid < 0, solid > 0    */
    const int   *group_ids;        /* [N] or NULL (render-only)                 */
    float        particle_radius;  /* solidRadius, scene.h:500                  */
    float        kernel_radius;    /* h, scene.h:501                            */
    float        cell_size[3];     /* particlesolver.cuh:150                    */
    int          grid_dims[3];     /* per-axis powers of two (reference: 512^3) */
    int          num_planes;       /* scene.h:513                               */
    const float *plane_origins;    /* [3*num_planes]                            */
    const float *plane_normals;    /* [3*num_planes]                            */
    int          num_distance;     /* scene.h:525-526                           */
    const int   *distance_pairs;   /* [2*nd] original ids                       */
    const float *distance_params;  /* [2*nd] (rest length, stiffness)           */
    int          num_bendings;     /* scene.h:528                               */
    const int   *bendings;         /* [4*nb] original ids                       */
    int          num_faces;        /* scene.h:516                               */
    const int   *faces;            /* [3*nf] original ids                        */
    int          num_rigid;        /* scene.h:522                               */
    const int   *rigid_ranges;     /* [2*nr] [begin,end) original ids, <=64 each */
    int          rigid_rest_len;   /* number of entries in rigid_rest_positions */
    const float *rigid_rest_positions; /* [3*rigid_rest_len] indexed by original
                                      id, centre of mass at 0 (scene.h:521)     */
    int          num_immovables;   /* scene.h:530                               */
    const int   *immovables;       /* [ni] original ids                         */
} bp_scene_desc;

/* Device view of the solver state, sorted order (bp_get_device_arrays). */
typedef struct bp_device_view {
    int    num_particles;
    float *new_positions;  /* float4[N]: x* (xyz, w = phase as int bits)       */
    float *positions;      /* float4[N]: x  (xyz, w = inverse scaled mass)     */
    float *velocities;     /* float4[N]: v  (xyz, w = 0)                       */
    int   *phases;         /* int[N]                                           */
    int   *map_to_original;/* int[N]: slot -> original id                      */
    int   *map_to_new;     /* int[N]: original id -> slot                      */
    float *rigid_rotations;/* float4[nr] quaternions (x,y,z,w)                 */
    float *rigid_cms;      /* float4[nr]                                       */
} bp_device_view;

typedef struct bp_solver bp_solver;

/* Library + parameter helpers ------------------------------------------- */
int         bp_abi_version(void);
const char *bp_last_error_string(void);
/* Fills the reference literals (see bp_params). */
bp_status   bp_default_params(bp_params *out);

/* Lifetime ----------------------------------------------------------------- */
/* Replaces ParticleSolver(const std::shared_ptr<Scene>&), particlesolver.cuh:148-215.
 * `hip_stream` is a hipStream_t or NULL (the null stream). */
bp_status bp_solver_create(const bp_scene_desc *scene, const bp_params *params,
                           void *hip_stream, bp_solver **out);
/* The reference has no destructor (particlesolver.cuh:1074). */
bp_status bp_solver_destroy(bp_solver *s);

/* Step --------------------------------------------------------------------- */
/* Replaces ParticleSolver::update(numSubTimeStep, deltaTime, pickedOriginalParticleId,
 * pickedParticlePosition, pickedParticleVelocity), particlesolver.cuh:772-1072.
 * `picked_pos` / `picked_vel` may be NULL when picked_original_id < 0.
 * Asynchronous on the solver's stream. */
bp_status bp_solver_step(bp_solver *s, int num_substeps, float dt,
                         int picked_original_id, const float *picked_pos,
                         const float *picked_vel);
/* Waits for the solver's stream. */
bp_status bp_solver_sync(bp_solver *s);
/* Changes parameters between steps (same scene). */
bp_status bp_solver_set_params(bp_solver *s, const bp_params *params);

/* Queries ------------------------------------------------------------------ */
/* queryOriginalParticleId, particlesolver.cuh:352-358 */
bp_status bp_query_original_id(bp_solver *s, int slot, int *out_original_id);
/* queryNewParticleId, particlesolver.cuh:360-366 */
bp_status bp_query_slot(bp_solver *s, int original_id, int *out_slot);
/* getParticlePosition, particlesolver.cuh:378-386 (slot index) */
bp_status bp_get_particle(bp_solver *s, int slot, float out_pos[3]);
/* setParticle, particlesolver.cuh:388-393 (slot index; sets x and v) */
bp_status bp_set_particle(bp_solver *s, int slot, const float pos[3], const float vel[3]);
/* Public device members devPositions etc., particlesolver.cuh:1075-1148 */
bp_status bp_get_device_arrays(bp_solver *s, bp_device_view *out);
/* Headless export in original-id order: pos/vel [3N] host, either may be NULL. */
bp_status bp_copy_state_original_order(bp_solver *s, float *pos, float *vel);
/* Full state in slot order, host buffers.  pos/vel [3N]; map [N] slot -> original id. */
bp_status bp_get_state(bp_solver *s, float *pos, float *vel, int *map_to_original);
bp_status bp_set_state(bp_solver *s, const float *pos, const float *vel,
                       const int *map_to_original);
/* The renderer boundary, headless (main.cu:273-291; the reference maps a GL
 * buffer and launches mapParticleInfos and mapMatrices into it).  One record
 * per particle in original-id order: {x, y, z, group id}, the group id's int
 * bits in the fourth float (bp_scene_desc::group_ids, 0 when absent).  One
 * column-major 4x4 per rigid body: the rotation of its quaternion and its
 * centre of mass as translation (extract_rotation_matrix4,
 * cudamatrix.cuh:175-199).
 * bp_export_frame: particles [4N] and matrices [16 * rigid bodies] are host
 * buffers, either may be NULL.  bp_map_particle_infos: the particle records
 * into a device buffer of 16 N bytes on the solver's device (a mapped
 * graphics buffer, say), ordered on the solver's stream. */
bp_status bp_export_frame(bp_solver *s, float *particles, float *matrices);
bp_status bp_map_particle_infos(bp_solver *s, void *device_records);
/* Rigid-body warm-start state: quats [4*nr] (x,y,z,w), cms [3*nr]; either may be NULL. */
bp_status bp_get_rigid_state(bp_solver *s, float *quats, float *cms);
bp_status bp_set_rigid_state(bp_solver *s, const float *quats);

/* Intermediate buffers of the last substep, slot order, for parity checks.
 * which: 0 = densities [N], 1 = lambdas [N], 2 = sorted cell keys [N],
 *        3 = sort permutation (previous slot of each new slot) [N],
 *        4 = omegas [3N], 5 = normals [3N],
 *        6 = the largest |term| the last bp_backward delivered between slots
 *            in 2^-40 fixed point, in seed units [1] (bp_backward fails with
 *            BP_ERR_NOT_FINITE when a term leaves that format's range). */
bp_status bp_debug_fetch(bp_solver *s, int which, void *host_out);

/* Stand-alone neighbour grid (updateGrid, particlesolver.cuh:712-765):
 * hash (grid.cuh:30-47), stable sort (particlesolver.cuh:721-738) and cell
 * ranges (grid.cuh:70-99) of n positions xs[3n] (host).  Outputs (host):
 * keys[n] sorted cell ids, perm[n] sorted slot values, cell_start/cell_end
 * [dims product] (-1 for empty cells).  Any output may be NULL. */
bp_status bp_neighbor_grid(const float *xs, int n, const float cell_size[3],
                           const int grid_dims[3], int *keys, int *perm,
                           int *cell_start, int *cell_end);

/* bp_neighbor_grid without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops).  Device buffers throughout: positions = float4[n] (x, y, z,
 * w ignored), keys[n] and perm[n] (required when n > 0), cell_start and
 * cell_end [dims product] (may be NULL), results equal to bp_neighbor_grid's.
 * Asynchronous on hip_stream (NULL = default stream); no hipMalloc, hipFree or
 * memset.  At most 2^20 positions (BP_ERR_CAPACITY beyond).  The sort's
 * scratch is one fixed array per GPU (about 9.4 MB, held until the process
 * ends) that keeps the last call's contents: calls that overlap on one GPU
 * (two streams, two threads) race and must be serialised by the caller. */
bp_status bp_neighbor_grid_noalloc(const float *positions, int n, const float cell_size[3],
                                   const int grid_dims[3], int *keys, int *perm,
                                   int *cell_start, int *cell_end, void *hip_stream);

/* Batched solver (no counterpart in the reference, which runs one scene):
 * `num_scenes` independent scenes laid end to end in one solver, each with its
 * own grid, kernel radius, planes and per-scene parameters (rest density, wind,
 * XSPH, friction, ...); every pass runs once over all scenes.  Parameters that
 * shape the pass sequence (iterations, pass enables, gravity, solid and cloth
 * constants) must agree.  Original ids are global: scene s owns
 * [first, first + count) from bp_batch_scene_range.  All other entry points
 * (step, state, queries) act on the whole batch. */
bp_status bp_batch_create(const bp_scene_desc *scenes, const bp_params *params, int num_scenes,
                          void *hip_stream, bp_solver **out);
bp_status bp_batch_scene_range(bp_solver *s, int scene, int *first_original_id, int *count);

/* Replays substeps from a captured hipGraph (two substeps per graph, captured
 * on first use and whenever dt changes).  No counterpart in the reference
 * (which launches every kernel each frame, particlesolver.cuh:796-1062). */
bp_status bp_solver_set_graph(bp_solver *s, int enable);

/* ---------------------------------------------------------------- adjoint
 * Differentiable rollout (no counterpart in the reference).  Differentiated:
 * prediction, wind (drag/lift with the wind vector), plane collisions, the
 * PBF projection (lambda_mode 0, sCorr, clamp), distance constraints, pins,
 * the velocity/position update and XSPH.  Scenes with rigid bodies, bending
 * constraints, vorticity or Akinci tension are refused (BP_ERR_UNSUPPORTED),
 * and so is a rollout in which particle-particle contact occurs.  Neighbour
 * sets are those of the forward passes, held fixed.
 *
 * bp_forward_checkpointed advances `substeps` substeps of size `subdt`
 * (every pass a separate kernel, no picking) and keeps the full particle
 * state entering every `checkpoint_every`-th substep.  Device memory: the
 * checkpoints plus one segment of tape, about (300 + 64 x iterations) B per
 * particle per substep of a segment. */
bp_status bp_forward_checkpointed(bp_solver *s, int substeps, float subdt, int checkpoint_every);

/* Reverse sweep over the last bp_forward_checkpointed rollout: each segment
 * is replayed from its checkpoint (recording the state before every pass),
 * except the last, which the forward recorded as it ran it, then
 * differentiated in reverse.  Inputs: dL/dx_T, dL/dv_T (float3 per
 * particle, original order; either may be NULL = 0).  Outputs (each may be
 * NULL): dL/dx_0, dL/dv_0 (float3, original order) and, per scene s,
 * grad_theta[6 s .. 6 s + 5] = dL/d(wind x, wind y, wind z, drag_coeff,
 * lift_coeff, xsph_c).  The solver is left at the end of the rollout; the
 * rollout can be differentiated again with other seeds.  Any state change
 * (step, set_state, set_params, ...) discards it. */
bp_status bp_backward(bp_solver *s, const float *grad_x, const float *grad_v, float *grad_x0, float *grad_v0,
                      float *grad_theta);

/* Iteration fusion (default on): the next iteration's plane collisions, the
 * velocity/position update and the lambda pass's inputs are computed inside
 * the kernels that produce their operands; off, every pass is its own launch.
 * Results are bit-identical either way.  No counterpart in the reference. */
bp_status bp_solver_set_fusion(bp_solver *s, int enable);

/* Walk forms of the neighbour passes (default 1 for all four).  Lambda and
 * omega/normal passes: 128, 256 or 512 = each wave stages its candidates'
 * records in an LDS region of that many records per column offset (about
 * eight per occupied cell; a column offset whose candidates do not fit reads
 * global memory) and screens and fetches partners from there, 1 = the same
 * with 256, 0 = every lane reads its candidates from global memory; any other
 * value is BP_ERR_INVALID_ARG.  Position and vorticity/tension passes:
 * 1 = the candidate walk records each 32-candidate chunk's partner mask and
 * the pair terms run after the walk over all partners, 0 = the pair terms of
 * each chunk run right after its screen.  Same pairs in the same order, so
 * results are bit-identical; only the speed differs (bp_solver_autotune
 * chooses).  No counterpart in the reference. */
bp_status bp_solver_set_walk(bp_solver *s, int lambda_staged, int position_flat, int omega_normal_staged,
                             int vorticity_tension_flat);

/* Partners the position pass fetches together per round trip to memory, 2
 * (the default) or 4; anything else is BP_ERR_INVALID_ARG.  The pass mostly
 * waits on these loads; more in flight per wave costs registers.  Same pairs
 * in the same order, so results are bit-identical (bp_solver_autotune
 * chooses).  No counterpart in the reference. */
bp_status bp_solver_set_partner_fetch(bp_solver *s, int position_partners);

/* Candidates the staged lambda and omega/normal walks screen per batch from
 * LDS, 4 (the default) or 8; anything else is BP_ERR_INVALID_ARG.  Four take
 * fewer registers, so more waves fit, which pays on large workloads; eight
 * issue more work per batch, which pays when every wave is resident anyway.
 * Same pairs in the same order, so results are bit-identical
 * (bp_solver_autotune chooses).  No counterpart in the reference. */
bp_status bp_solver_set_stage_screen(bp_solver *s, int lambda_candidates, int omega_normal_candidates);

/* Chooses the workgroup size of each neighbour pass among 64..512 and the
 * walk forms of the lambda, position, omega/normal and vorticity/tension
 * passes (bp_solver_set_walk, including the staged walks' region sizes) by
 * timing `substeps_per_trial` substeps of size dt per candidate, and
 * iteration fusion on or off (best of two trials each).  Every trial starts
 * from the particle and rigid state the call started from, and that state is
 * restored at the end (results do not depend on any of these choices).
 * Writes a JSON object {pass name: size, "iteration_fusion": 0|1,
 * "lambda_walk"/"omega_normal_walk": "staged"|"global",
 * "position_walk"/"vorticity_tension_walk": "flat"|"chunk",
 * "lambda_stage_records"/"omega_normal_stage_records": 0|128|256|512 (the
 * staged region size, 0 for the per-lane walk), "position_partners": 2|4
 * (bp_solver_set_partner_fetch), "lambda_stage_screen"/
 * "omega_normal_stage_screen": 4|8 (bp_solver_set_stage_screen)} into
 * `report` (may be NULL; 1024 bytes hold it).  No counterpart in the
 * reference. */
bp_status bp_solver_autotune(bp_solver *s, float dt, int substeps_per_trial, char *report, int report_len);

/* Sets the neighbour passes' workgroup sizes, each 64, 128, 256 or 512, in
 * the order lambda, position, omega/normal, vorticity/tension, XSPH,
 * particle-particle (the sizes bp_solver_autotune reports), e.g. to reuse one
 * tuning across runs.  Results do not depend on them.  No counterpart in the
 * reference. */
bp_status bp_solver_set_block_sizes(bp_solver *s, const int sizes[6]);

/* Timing helper for benchmarks: runs `substeps` substeps of size dt and
 * returns the elapsed GPU time in ms measured with HIP events on the
 * solver's stream (bracketing all launches). */
bp_status bp_solver_timed_substeps(bp_solver *s, int substeps, float dt, float *out_ms);

/* HBM bandwidth check for roofline reporting (no counterpart in the
 * reference): a device-to-device copy of `bytes` (a multiple of 16) by this
 * library's streaming kernel on the default stream, at 1024, 2048, 4096 and
 * 8192 workgroups of 256, 4 or 8 loads in flight per lane, cached or
 * non-temporal, best of `reps` each after one warm-up, timed with HIP
 * events.  *best_ms = the fastest copy; bytes read plus written over it
 * is the bandwidth.  Allocates and frees its two buffers. */
bp_status bp_stream_copy_bandwidth(size_t bytes, int reps, float *best_ms);

/* Per-kernel timing over the last bp_solver_timed_substeps call when
 * profiling is enabled with bp_solver_enable_kernel_timing(s, 1): names and
 * total ms per kernel family, `max` entries.  Returns the count in *count. */
bp_status bp_solver_enable_kernel_timing(bp_solver *s, int enable);
bp_status bp_solver_kernel_times(bp_solver *s, int max, char *names /*max*32*/,
                                 float *total_ms, int *launches, int *count);

#ifdef __cplusplus
}
#endif
#endif /* BRASSPAN_H_ */
