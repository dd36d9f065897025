// particlesolver.h -- brass-pan's `struct ParticleSolver` API
// (brasstower/particlesolver.cuh:146-1148) over the C ABI of libbrasspan.so.
// Header-only; link with -lbrasspan.  Errors throw std::runtime_error with
// the library's message (the reference prints and continues,
// cuda/helper.cuh:282-290).
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/brasspan.h"
#include "scene.h"

struct ParticleSolver {
    // particlesolver.cuh:148-215.  `params` defaults to the reference literals;
    // `hipStream` (a hipStream_t) defaults to the null stream.
    explicit ParticleSolver(const std::shared_ptr<Scene> &scene, const bp_params *params = nullptr,
                            void *hipStream = nullptr) {
        bp_params p;
        if (params) p = *params; else check(bp_default_params(&p));
        const Scene &s = *scene;
        planeO.clear();
        planeN.clear();
        for (const Plane &pl : s.planes) {
            planeO.insert(planeO.end(), {pl.origin.x, pl.origin.y, pl.origin.z});
            planeN.insert(planeN.end(), {pl.normal.x, pl.normal.y, pl.normal.z});
        }
        std::vector<int> faces3;
        for (const int3 &f : s.faces) faces3.insert(faces3.end(), {f.x, f.y, f.z});
        bp_scene_desc d{};
        d.num_particles = (int)s.positions.size();
        d.positions = reinterpret_cast<const float *>(s.positions.data());
        d.masses = s.masses.data();
        d.phases = s.phases.data();
        d.group_ids = s.groupIds.empty() ? nullptr : s.groupIds.data();
        d.particle_radius = s.particleRadius;
        d.kernel_radius = s.fluidKernelRadius;
        d.cell_size[0] = s.cellSize.x; d.cell_size[1] = s.cellSize.y; d.cell_size[2] = s.cellSize.z;
        d.grid_dims[0] = s.gridDims.x; d.grid_dims[1] = s.gridDims.y; d.grid_dims[2] = s.gridDims.z;
        d.num_planes = (int)s.planes.size();
        d.plane_origins = planeO.data();
        d.plane_normals = planeN.data();
        d.num_distance = (int)s.distancePairs.size();
        d.distance_pairs = reinterpret_cast<const int *>(s.distancePairs.data());
        d.distance_params = reinterpret_cast<const float *>(s.distanceParams.data());
        d.num_bendings = (int)s.bendingConstraints.size();
        d.bendings = reinterpret_cast<const int *>(s.bendingConstraints.data());
        d.num_faces = (int)s.faces.size();
        d.faces = faces3.data();
        d.num_rigid = (int)s.rigidbodyParticleIdRanges.size();
        d.rigid_ranges = reinterpret_cast<const int *>(s.rigidbodyParticleIdRanges.data());
        d.rigid_rest_len = (int)s.rigidbodyInitialPositions.size();
        d.rigid_rest_positions = reinterpret_cast<const float *>(s.rigidbodyInitialPositions.data());
        d.num_immovables = (int)s.immovableConstraints.size();
        d.immovables = s.immovableConstraints.data();
        check(bp_solver_create(&d, &p, hipStream, &handle));
        numParticles = (size_t)d.num_particles;
        solidRadius = s.particleRadius;
        fluidKernelRadius = s.fluidKernelRadius;
    }
    ~ParticleSolver() { bp_solver_destroy(handle); }
    ParticleSolver(const ParticleSolver &) = delete;
    ParticleSolver &operator=(const ParticleSolver &) = delete;

    // particlesolver.cuh:772-776
    void update(const int numSubTimeStep, const float deltaTime, const int pickedOriginalParticleId = -1,
                const glm::vec3 &pickedParticlePosition = glm::vec3(0.0f),
                const glm::vec3 &pickedParticleVelocity = glm::vec3(0.0f)) {
        const float pp[3] = {pickedParticlePosition.x, pickedParticlePosition.y, pickedParticlePosition.z};
        const float pv[3] = {pickedParticleVelocity.x, pickedParticleVelocity.y, pickedParticleVelocity.z};
        check(bp_solver_step(handle, numSubTimeStep, deltaTime, pickedOriginalParticleId, pp, pv));
    }

    // particlesolver.cuh:352-366
    int queryOriginalParticleId(int newParticleId) {
        int o = -1;
        check(bp_query_original_id(handle, newParticleId, &o));
        return o;
    }
    int queryNewParticleId(int oldParticleId) {
        int o = -1;
        check(bp_query_slot(handle, oldParticleId, &o));
        return o;
    }
    // particlesolver.cuh:378-386 (slot index)
    glm::vec3 getParticlePosition(const int particleIndex) {
        float p[3];
        check(bp_get_particle(handle, particleIndex, p));
        return glm::vec3(p[0], p[1], p[2]);
    }
    // particlesolver.cuh:388-393 (slot index; sets x and v)
    void setParticle(const int particleIndex, const glm::vec3 &position, const glm::vec3 &velocity) {
        const float p[3] = {position.x, position.y, position.z}, v[3] = {velocity.x, velocity.y, velocity.z};
        check(bp_set_particle(handle, particleIndex, p, v));
    }
    // Device views of the sorted state (the reference's public dev* members).
    bp_device_view deviceArrays() {
        bp_device_view v{};
        check(bp_get_device_arrays(handle, &v));
        return v;
    }
    // Positions and velocities in original-id order (headless export).
    void copyStateOriginalOrder(std::vector<float3> &pos, std::vector<float3> &vel) {
        pos.resize(numParticles);
        vel.resize(numParticles);
        check(bp_copy_state_original_order(handle, reinterpret_cast<float *>(pos.data()),
                                           reinterpret_cast<float *>(vel.data())));
    }
    void synchronize() { check(bp_solver_sync(handle)); }

    size_t numParticles = 0;
    float solidRadius = 0.f;
    float fluidKernelRadius = 0.f;
    bp_solver *handle = nullptr;

  private:
    std::vector<float> planeO, planeN;
    static void check(bp_status s) {
        if (s != BP_OK) throw std::runtime_error(std::string("brasspan: ") + bp_last_error_string());
    }
};
