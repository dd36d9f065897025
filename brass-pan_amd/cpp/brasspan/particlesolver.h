// particlesolver.h -- brass-pan's `struct ParticleSolver` API
// (brasstower/particlesolver.cuh:146-1148) over the C ABI of libbrasspan.so.
// Header-only; link with -lbrasspan.  Errors throw std::runtime_error with
// the library's message (the reference prints and continues,
// cuda/helper.cuh:282-290).
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/brasspan.h"
#include "oldscene.h"
#include "scene.h"

// Flat C-ABI description of a Scene (keeps the arrays it points into).
struct SceneDesc {
    explicit SceneDesc(const Scene &s) {
        for (const Plane &pl : s.planes) {
            planeO.insert(planeO.end(), {pl.origin.x, pl.origin.y, pl.origin.z});
            planeN.insert(planeN.end(), {pl.normal.x, pl.normal.y, pl.normal.z});
        }
        for (const int3 &f : s.faces) faces3.insert(faces3.end(), {f.x, f.y, f.z});
        d.num_particles = (int)s.positions.size();
        d.positions = reinterpret_cast<const float *>(s.positions.data());
        d.masses = s.masses.data();
        d.phases = s.phases.data();
        d.group_ids = s.groupIds.empty() ? nullptr : s.groupIds.data();
        d.particle_radius = s.particleRadius;
        d.kernel_radius = s.fluidKernelRadius;
        const float3 cs = s.effectiveCellSize();
        d.cell_size[0] = cs.x; d.cell_size[1] = cs.y; d.cell_size[2] = cs.z;
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
    }
    SceneDesc(const SceneDesc &) = delete;
    SceneDesc &operator=(const SceneDesc &) = delete;
    bp_scene_desc d{};

  private:
    std::vector<float> planeO, planeN;
    std::vector<int> faces3;
};

inline void bp_check(bp_status s) {
    if (s != BP_OK) throw std::runtime_error(std::string("brasspan: ") + bp_last_error_string());
}

// What the plain and the batched solver share: state export, graphs, the
// block-size autotuner and the differentiable rollout (no counterparts in
// the reference).
struct SolverHandle {
    SolverHandle() = default;
    SolverHandle(const SolverHandle &) = delete;
    SolverHandle &operator=(const SolverHandle &) = delete;
    ~SolverHandle() { bp_solver_destroy(handle); }

    // Positions and velocities in original-id order (headless export).
    void copyStateOriginalOrder(std::vector<float3> &pos, std::vector<float3> &vel) {
        pos.resize(numParticles);
        vel.resize(numParticles);
        bp_check(bp_copy_state_original_order(handle, reinterpret_cast<float *>(pos.data()),
                                              reinterpret_cast<float *>(vel.data())));
    }
    // The renderer boundary, headless (main.cu:273-291): one record per
    // particle in original-id order, {x, y, z, group id} with the id's int bits
    // in w, and one column-major 4x4 per rigid body (rotation + centre of mass).
    void exportFrame(std::vector<float4> &particles, std::vector<float> &matrices) {
        particles.resize(numParticles);
        matrices.resize(16 * (size_t)numRigidBodies);
        bp_check(bp_export_frame(handle, reinterpret_cast<float *>(particles.data()),
                                 matrices.empty() ? nullptr : matrices.data()));
    }
    // The particle records into a device buffer of 16 * numParticles bytes.
    void mapParticleInfos(void *deviceRecords) { bp_check(bp_map_particle_infos(handle, deviceRecords)); }
    // Writes the frame file read by brass-pan_amd/frames.py: a 32-byte header
    // ("BPFR", u32 version 1, u32 particles, u32 bodies, u64 frame, f64 time),
    // then the particle records, then the matrices (little endian).
    void writeFrame(const std::string &path, uint64_t frame, double time) {
        std::vector<float4> particles;
        std::vector<float> matrices;
        exportFrame(particles, matrices);
        unsigned char header[32];
        const uint32_t version = 1, n = (uint32_t)numParticles, nb = (uint32_t)numRigidBodies;
        std::memcpy(header, "BPFR", 4);
        std::memcpy(header + 4, &version, 4);
        std::memcpy(header + 8, &n, 4);
        std::memcpy(header + 12, &nb, 4);
        std::memcpy(header + 16, &frame, 8);
        std::memcpy(header + 24, &time, 8);
        std::FILE *f = std::fopen(path.c_str(), "wb");
        if (!f) throw std::runtime_error("brasspan: cannot open " + path);
        const bool ok = std::fwrite(header, 1, sizeof(header), f) == sizeof(header) &&
                        std::fwrite(particles.data(), sizeof(float4), particles.size(), f) == particles.size() &&
                        std::fwrite(matrices.data(), sizeof(float), matrices.size(), f) == matrices.size();
        if (std::fclose(f) != 0 || !ok) throw std::runtime_error("brasspan: cannot write " + path);
    }
    void synchronize() { bp_check(bp_solver_sync(handle)); }
    // Replays pairs of substeps from a captured hipGraph.
    void setGraph(bool enable) { bp_check(bp_solver_set_graph(handle, enable ? 1 : 0)); }
    // Iteration fusion (epilogue passes inside their producing kernels); bit-identical either way.
    void setFusion(bool enable) { bp_check(bp_solver_set_fusion(handle, enable ? 1 : 0)); }
    // Chooses the neighbour-pass workgroup sizes and iteration fusion; returns the JSON report.
    std::string autotune(float subDeltaTime, int substepsPerTrial = 2) {
        char buf[512] = {0};
        bp_check(bp_solver_autotune(handle, subDeltaTime, substepsPerTrial, buf, (int)sizeof(buf)));
        return std::string(buf);
    }
    // Differentiable rollout: `substeps` substeps of `subDeltaTime`, full
    // state kept every `checkpointEvery` substeps.
    void forwardCheckpointed(int substeps, float subDeltaTime, int checkpointEvery = 8) {
        bp_check(bp_forward_checkpointed(handle, substeps, subDeltaTime, checkpointEvery));
    }
    // Adjoint of the last rollout.  gradX/gradV: dL/dx_T, dL/dv_T (original
    // order; gradV may be empty).  Outputs dL/dx_0, dL/dv_0 and, per scene,
    // dL/d(wind x, y, z, drag, lift, xsph c).
    void backward(const std::vector<float3> &gradX, const std::vector<float3> &gradV, std::vector<float3> &gradX0,
                  std::vector<float3> &gradV0, std::vector<float> &gradTheta) {
        if (gradX.size() != numParticles || (!gradV.empty() && gradV.size() != numParticles))
            throw std::invalid_argument("brasspan: gradient seeds need one float3 per particle");
        gradX0.resize(numParticles);
        gradV0.resize(numParticles);
        gradTheta.resize(6 * (size_t)numScenes);
        bp_check(bp_backward(handle, reinterpret_cast<const float *>(gradX.data()),
                             gradV.empty() ? nullptr : reinterpret_cast<const float *>(gradV.data()),
                             reinterpret_cast<float *>(gradX0.data()), reinterpret_cast<float *>(gradV0.data()),
                             gradTheta.data()));
    }

    size_t numParticles = 0;
    int numScenes = 1;
    int numRigidBodies = 0;
    bp_solver *handle = nullptr;
};

struct ParticleSolver : SolverHandle {
    // particlesolver.cuh:148-215.  `params` defaults to the reference literals;
    // `hipStream` (a hipStream_t) defaults to the null stream.
    explicit ParticleSolver(const std::shared_ptr<Scene> &scene, const bp_params *params = nullptr,
                            void *hipStream = nullptr) {
        bp_params p;
        if (params) p = *params; else check(bp_default_params(&p));
        init(*scene, p, hipStream);
    }
    // particlesolver.cuh:218-342: the legacy scene format, assembled as the
    // reference's constructor assembles it (oldscene.h), with its fluid rest
    // density (:338; a scene without fluid leaves it unset, and only fluid
    // reads it); the scene's counters advance as the parts are added.
    explicit ParticleSolver(const std::shared_ptr<OldSceneFormat> &oldScene, const bp_params *params = nullptr,
                            void *hipStream = nullptr) {
        bp_params p;
        if (params) p = *params; else check(bp_default_params(&p));
        if (oldScene->fluidRestDensity > 0.0f || !oldScene->fluids.empty())
            p.fluid_rest_density = oldScene->fluidRestDensity;
        init(assembleOldScene(*oldScene), p, hipStream);
    }

    void init(const Scene &s, const bp_params &p, void *hipStream) {
        SceneDesc d(s);
        check(bp_solver_create(&d.d, &p, hipStream, &handle));
        numParticles = (size_t)d.d.num_particles;
        numRigidBodies = d.d.num_rigid;
        solidRadius = s.particleRadius;
        fluidKernelRadius = s.fluidKernelRadius;
    }

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

    float solidRadius = 0.f;
    float fluidKernelRadius = 0.f;

  private:
    static void check(bp_status s) { bp_check(s); }
};

// Many independent scenes advanced together (bp_batch_create): one launch per
// pass serves every scene.  No counterpart in the reference.
struct BatchParticleSolver : SolverHandle {
    BatchParticleSolver(const std::vector<std::shared_ptr<Scene>> &scenes, const std::vector<bp_params> &params,
                        void *hipStream = nullptr) {
        if (scenes.size() != params.size() || scenes.empty())
            throw std::invalid_argument("brasspan: one bp_params per scene");
        std::vector<std::unique_ptr<SceneDesc>> keep;
        std::vector<bp_scene_desc> ds;
        for (const auto &sc : scenes) {
            keep.emplace_back(new SceneDesc(*sc));
            ds.push_back(keep.back()->d);
        }
        bp_check(bp_batch_create(ds.data(), params.data(), (int)ds.size(), hipStream, &handle));
        numScenes = (int)ds.size();
        for (const auto &d : ds) {
            numParticles += (size_t)d.num_particles;
            numRigidBodies += d.num_rigid;
        }
    }
    // ParticleSolver::update for all scenes (no picking).
    void update(const int numSubTimeStep, const float deltaTime) {
        bp_check(bp_solver_step(handle, numSubTimeStep, deltaTime, -1, nullptr, nullptr));
    }
    // Scene k owns original ids [first, first + count).
    void sceneRange(int k, int &first, int &count) { bp_check(bp_batch_scene_range(handle, k, &first, &count)); }
};
