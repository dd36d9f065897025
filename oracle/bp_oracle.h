/*
 * bp_oracle.h -- CPU oracle of brass-pan's particle-solver substep.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline -- never as part of the product path.
 *
 * A host restatement, in the reference's operation order, of
 * ParticleSolver::update (brasstower/particlesolver.cuh:772-1072) with the
 * commented-out projection and wind blocks (:838-858, :881-993) restored, and
 * of the kernels under brasstower/kernel/.  Templated on the real type: the
 * FP32 instance is the parity oracle, the FP64 instance is the error
 * reference, and a forward-mode dual-number instance gives exact FP64
 * derivatives of a rollout (or_rollout_tangent).  Grid hashing is always done
 * in FP32 on the (rounded) position,
 * exactly as the reference does, so both instances see the same neighbour
 * candidate sets.
 *
 * Two projection modes:
 *   mode 0 "scatter": fluidPosition's pair loop (fluid.cuh:161-228) as
 *          written -- the lower sorted index i owns pair (i,j), both sides are
 *          accumulated immediately (atomics serialised in visit order).
 *   mode 1 "gather":  every particle sums its own side over its own search
 *          cells, with the reference's pair-ownership rule (the pair exists
 *          only if the lower-index particle is not a solid) and the
 *          reference's candidate criterion for partners below it.  All
 *          neighbour sums visit cells in the GPU kernels' order (y, x, then
 *          the z run) instead of the reference's (z, y, x).
 */
#ifndef BP_ORACLE_H_
#define BP_ORACLE_H_

#include "../include/brasspan.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_solver or_solver;

or_solver *or_create(const bp_scene_desc *scene, const bp_params *params, int fp64, int mode);
void or_destroy(or_solver *s);
int  or_num_particles(or_solver *s);
/* ParticleSolver::update semantics; returns 0 on success. */
int  or_step(or_solver *s, int substeps, float dt, int picked_original_id,
             const float *picked_pos, const float *picked_vel);
void or_set_params(or_solver *s, const bp_params *params);
/* State in slot order, as doubles.  pos/vel [3N], map [N] slot -> original id. */
void or_get_state(or_solver *s, double *pos, double *vel, int *map);
void or_set_state(or_solver *s, const double *pos, const double *vel, const int *map);
void or_get_rigid(or_solver *s, double *quats, double *cms);
void or_set_rigid(or_solver *s, const double *quats);
/* which: 0 densities, 1 lambdas, 4 omegas (3N), 5 normals (3N) */
void or_debug_fetch(or_solver *s, int which, double *out);
/* which: 2 sorted keys, 3 permutation */
void or_debug_fetch_int(or_solver *s, int which, int *out);

/* Differentiable-rollout checks.  Loss L = sum_i |x_i(T) - target_i|^2 / norm
 * after `substeps` substeps of `subdt` from the scene's initial state (FP64,
 * reference order).  or_rollout_loss: theta (6: wind xyz, drag, lift, xsph_c;
 * NULL = params) and v0 (3N, original order; NULL = scene) override the
 * inputs, for central finite differences.  or_rollout_tangent: the FP64
 * derivative by forward-mode dual numbers -- dloss[k] = dL/dtheta_k for
 * k < 6, and dloss[6 + j] = dL/dv0 . vdirs[j] for j < ndirs_v <= 2. */
double or_rollout_loss(const bp_scene_desc *scene, const bp_params *params, int substeps, float subdt,
                       const double *target, double norm, const double *theta, const double *v0);
double or_rollout_tangent(const bp_scene_desc *scene, const bp_params *params, int substeps, float subdt,
                          const double *target, double norm, const double *vdirs, int ndirs_v, double *dloss);

/* updateGrid stand-alone (grid.cuh:30-99, particlesolver.cuh:712-739). */
/* OpenMP threads for the per-particle loops (results do not depend on it);
 * returns the count now in effect. n <= 0 only queries. */
int or_set_threads(int n);

void or_neighbor_grid(const float *xs, int n, const float cell_size[3],
                      const int grid_dims[3], int *keys, int *perm,
                      int *cell_start, int *cell_end);
/* calcGridPos + calcGridAddress (grid.cuh:9-28) of one position. */
int  or_cell_address(const float p[3], const float cell_size[3], const int grid_dims[3]);

/* SPH kernels (fluid.cuh:3-74) for closed-form checks.  fp64 selects the twin. */
double or_poly6(double r2, double h, int fp64);
void   or_grad_spiky(const double v[3], double r2, double h, int fp64, double out[3]);
double or_akinci_c(double r, double h, int fp64);
/* Rotation extraction of shapeMatchingAlphaOne (shapematching.cuh:64-91):
 * A = 3 columns (9 doubles, column-major), q in/out (x,y,z,w). */
void   or_extract_rotation(const double A[9], double q[4], int iterations, int fp64);

#ifdef __cplusplus
}
#endif
#endif
