// scene.h -- host scene description with brass-pan's `struct Scene` API
// (brasstower/scene.h:455-536, brasstower/scene.cpp), header-only.
//
// Same member and method names as the reference.  Differences, each a fix of
// a reference gap:
//  * addFluidBlock / addRigidBox / addRope / makeImmovable are declared but
//    not implemented in the reference (scene.h:470-496); implemented here.
//  * addCloth (scene.cpp:4-128) indexes particles in storage order, offsets
//    them by the particles already present, pushes masses and group ids, and
//    returns the [begin, end) range (the reference returns int2()).
//  * rigidbodyInitialPositions is indexed by original particle id for any
//    insertion order (the reference needs rigid bodies first).
//  * cellSize / gridDims make the hash grid explicit (the reference hard-codes
//    512^3 cells, particlesolver.cuh:151).  cellSize left at zero follows the
//    reference: particleRadius * 2.3f when the solver is built (:150).
//  * addCloth takes a trailing dihedralBending flag (default off) that adds
//    the reference's commented-out hinge generator (scene.cpp:108-123).
#pragma once
#include <hip/hip_vector_types.h>

#include <cmath>
#include <vector>

#include "glm_compat.h"

// Plane (scene.h:13-37) without the render-only model matrix.
struct Plane {
    glm::vec3 origin;
    glm::vec3 normal;
    Plane(const glm::vec3 &o, const glm::vec3 &n) : origin(o), normal(n) {}
};

struct Scene {
    Scene() {}

    // ---- builders (all return the [begin, end) particle range) -------------
    int2 addCloth(const float3 &startPosition, const float3 &stepX, const float3 &stepY, const int numJointX,
                  const int numJointY, const float massPerParticle, const float kStiffness, const float kBending,
                  const bool isSelfCollidable = true, const bool dihedralBending = false) {
        const float lx = len(stepX), ly = len(stepY);
        const float ld = len(make_float3(stepX.x + stepY.x, stepX.y + stepY.y, stepX.z + stepY.z));
        const int base = (int)positions.size();
        const int group = groupIdCounter++;
        auto P = [&](int x, int y) { return base + x * numJointY + y; };
        for (int x = 0; x < numJointX; x++)
            for (int y = 0; y < numJointY; y++) {
                positions.push_back(make_float3(startPosition.x + stepY.x * y + stepX.x * x,
                                                startPosition.y + stepY.y * y + stepX.y * x,
                                                startPosition.z + stepY.z * y + stepX.z * x));
                masses.push_back(massPerParticle);
                phases.push_back(isSelfCollidable ? solidPhaseCounter++ : solidPhaseCounter);
                groupIds.push_back(group);
            }
        if (!isSelfCollidable) solidPhaseCounter++;
        for (int x = 0; x < numJointX; x++)
            for (int y = 0; y < numJointY; y++) {
                const int p1 = P(x, y), p2 = P(x + 1, y), p3 = P(x, y + 1), p4 = P(x + 1, y + 1);
                if (x < numJointX - 1 && y < numJointY - 1) {
                    faces.push_back(make_int3(p1, p2, p3));
                    faces.push_back(make_int3(p2, p3, p4));
                }
                if (x < numJointX - 1) distance(p1, p2, lx, kStiffness);
                if (y < numJointY - 1) distance(p1, p3, ly, kStiffness);
                if (x < numJointX - 1 && y < numJointY - 1) {
                    distance(p1, p4, ld, kStiffness);
                    distance(p2, p3, ld, kStiffness);
                }
                if (x > 0 && y > 0 && x < numJointX - 1 && y < numJointY - 1) {  // "hack bending"
                    for (int j = -1; j <= 1; j++) distance(P(x - 1, y + j), P(x + 1, y + j), lx * 2, kBending);
                    for (int i = -1; i <= 1; i++) distance(P(x + i, y - 1), P(x + i, y + 1), ly * 2, kBending);
                }
                if (dihedralBending && x < numJointX - 1 && y < numJointY - 1)  // hinge across each quad
                    bendingConstraints.push_back((x + y) % 2 ? make_int4(p3, p2, p1, p4) : make_int4(p1, p4, p3, p2));
            }
        return make_int2(base, (int)positions.size());
    }

    int2 addFluidBlock(const uint3 &dimension, const float3 &startPosition, const glm::vec3 &stepX,
                       const glm::vec3 &stepY, const float massPerParticle) {
        // lattice along stepX (i), stepY (j) and stepX x stepY scaled to |stepX| (k)
        glm::vec3 n(stepX.y * stepY.z - stepX.z * stepY.y, stepX.z * stepY.x - stepX.x * stepY.z,
                    stepX.x * stepY.y - stepX.y * stepY.x);
        const float nl = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
        const float sl = std::sqrt(stepX.x * stepX.x + stepX.y * stepX.y + stepX.z * stepX.z);
        const glm::vec3 stepZ(n.x / nl * sl, n.y / nl * sl, n.z / nl * sl);
        const int begin = (int)positions.size();
        const int group = groupIdCounter++;
        const int phase = fluidPhaseCounter--;
        for (unsigned i = 0; i < dimension.x; i++)
            for (unsigned j = 0; j < dimension.y; j++)
                for (unsigned k = 0; k < dimension.z; k++) {
                    positions.push_back(make_float3(startPosition.x + stepX.x * i + stepY.x * j + stepZ.x * k,
                                                    startPosition.y + stepX.y * i + stepY.y * j + stepZ.y * k,
                                                    startPosition.z + stepX.z * i + stepY.z * j + stepZ.z * k));
                    masses.push_back(massPerParticle);
                    phases.push_back(phase);
                    groupIds.push_back(group);
                }
        return make_int2(begin, (int)positions.size());
    }

    // scene.cpp:130-146
    int2 addGranularsBlock(const uint3 &dimension, const float3 &startPosition, const float3 &step,
                           const float massPerParticle) {
        const int begin = (int)positions.size();
        const int group = groupIdCounter++;
        for (unsigned i = 0; i < dimension.x; i++)
            for (unsigned j = 0; j < dimension.y; j++)
                for (unsigned k = 0; k < dimension.z; k++) {
                    positions.push_back(make_float3(startPosition.x + step.x * i, startPosition.y + step.y * j,
                                                    startPosition.z + step.z * k));
                    masses.push_back(massPerParticle);
                    phases.push_back(solidPhaseCounter++);
                    groupIds.push_back(group);
                }
        return make_int2(begin, (int)positions.size());
    }

    // Box of dimension particles spanned by stepX, stepY (and their normal),
    // rest offsets about the centre of mass (at most 64 particles per body).
    int2 addRigidBox(const glm::ivec3 &dimension, const glm::vec3 &startPosition, const glm::vec3 &stepX,
                     const glm::vec3 &stepY, const float massPerParticle) {
        const uint3 d = make_uint3((unsigned)dimension.x, (unsigned)dimension.y, (unsigned)dimension.z);
        const int begin = (int)positions.size();
        const int2 r = addFluidBlock(d, make_float3(startPosition.x, startPosition.y, startPosition.z), stepX, stepY,
                                     massPerParticle);
        fluidPhaseCounter++;  // not a fluid block: undo the phase draw
        const int phase = solidPhaseCounter++;
        double cx = 0, cy = 0, cz = 0;
        for (int i = r.x; i < r.y; i++) { cx += positions[i].x; cy += positions[i].y; cz += positions[i].z; }
        const double n = (double)(r.y - r.x);
        if (rigidbodyInitialPositions.size() < (size_t)r.y) rigidbodyInitialPositions.resize(r.y, make_float3(0, 0, 0));
        for (int i = r.x; i < r.y; i++) {
            phases[i] = phase;
            rigidbodyInitialPositions[i] = make_float3((float)(positions[i].x - cx / n),
                                                       (float)(positions[i].y - cy / n),
                                                       (float)(positions[i].z - cz / n));
        }
        rigidbodyParticleIdRanges.push_back(make_int2(begin, r.y));
        return r;
    }

    int2 addRope(const float3 &startPosition, const float3 &step, const int numJoint, const float massPerParticle) {
        const int begin = (int)positions.size();
        const int group = groupIdCounter++;
        for (int i = 0; i < numJoint; i++) {
            positions.push_back(make_float3(startPosition.x + step.x * i, startPosition.y + step.y * i,
                                            startPosition.z + step.z * i));
            masses.push_back(massPerParticle);
            phases.push_back(solidPhaseCounter++);
            groupIds.push_back(group);
            if (i > 0) distance(begin + i - 1, begin + i, len(step), 1.0f);
        }
        return make_int2(begin, (int)positions.size());
    }

    void makeImmovable(const int id) { immovableConstraints.push_back(id); }

    size_t numParticles() { return positions.size(); }

    float particleRadius = 0.05f;
    // scene.h:501: the product is taken in double, then stored as float
    float fluidKernelRadius = 2.3 * particleRadius;
    // Hash-cell edge.  Left at zero it follows the radius when the solver is
    // built, particleRadius * 2.3f in float as particlesolver.cuh:150; any
    // positive value set by the caller is used as it is.
    float3 cellSize = make_float3(0.f, 0.f, 0.f);
    int3 gridDims = make_int3(512, 512, 512);

    float3 effectiveCellSize() const {
        if (cellSize.x > 0.f && cellSize.y > 0.f && cellSize.z > 0.f) return cellSize;
        const float c = particleRadius * 2.3f;
        return make_float3(c, c, c);
    }

    std::vector<float3> positions;
    std::vector<float> masses;
    std::vector<int> phases;
    std::vector<int> groupIds;
    std::vector<Plane> planes;
    std::vector<int3> faces;
    std::vector<float3> rigidbodyInitialPositions;
    std::vector<int2> rigidbodyParticleIdRanges;
    std::vector<int2> distancePairs;
    std::vector<float2> distanceParams;
    std::vector<int4> bendingConstraints;
    std::vector<int> immovableConstraints;
    int groupIdCounter = 1;
    int solidPhaseCounter = 1;
    int fluidPhaseCounter = -1;

  private:
    static float len(const float3 &v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
    void distance(int a, int b, float rest, float k) {
        distancePairs.push_back(make_int2(a, b));
        distanceParams.push_back(make_float2(rest, k));
    }
};
