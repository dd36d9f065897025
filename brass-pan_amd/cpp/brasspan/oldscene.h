// oldscene.h -- brass-pan's legacy scene format (`struct OldSceneFormat` and
// its parts, brasstower/scene.h:39-435) and the assembly the legacy
// constructor ParticleSolver(const std::shared_ptr<OldSceneFormat> &)
// performs (particlesolver.cuh:218-342 with addRigidBody / addGranulars /
// addFluids / addNoodles / addCloth, :395-710), as a conversion to `Scene`.
// Header-only.
//
// Kept from the reference:
//  * insertion order rigid bodies, granulars, fluids, ropes, cloths; one group
//    id per part counting from 0 (addNewGroup, :343-350);
//  * phases: a rigid body takes the solid counter and advances it by one;
//    granulars, ropes and cloths (self-colliding, the default) give every
//    particle its own solid phase; fluids take -1, -2, ... per part;
//  * constraint, bending, face and pin indices are local to their part and
//    offset by the particles already present;
//  * the capacity checks, with the reference's conditions (a part must leave
//    the total strictly below numMax*; the rope's distance check compares the
//    particle count, :515), throwing std::runtime_error with its messages;
//  * the grid: cell size and kernel radius 2.3 r, 512^3 cells (:218-225);
//  * the scene's counters (numParticles, numRigidBodies, ...) advance as the
//    parts are added.
// Differences:
//  * setDevArr_counterIncrement hands out the per-particle solid phases with
//    atomicAdd, so which particle gets which value is not defined there; here
//    they go in particle order (one of the reference's possible outcomes);
//  * render-only members (meshes, colours' use, camera, light) are not kept;
//  * constraint containers use HIP's int2 / int3 / int4 / float2.
#pragma once
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "scene.h"

namespace oldscene_detail {
inline glm::vec3 add(const glm::vec3 &a, const glm::vec3 &b) { return glm::vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline glm::vec3 sub(const glm::vec3 &a, const glm::vec3 &b) { return glm::vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline glm::vec3 mul(const glm::vec3 &a, const glm::vec3 &b) { return glm::vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline glm::vec3 scale(const glm::vec3 &a, float s) { return glm::vec3(a.x * s, a.y * s, a.z * s); }
inline float length(const glm::vec3 &a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// startPosition + stepSize * (i, j, k) over a dimension.x x .y x .z lattice, i slowest
inline std::vector<glm::vec3> lattice(const glm::ivec3 &dim, const glm::vec3 &start, const glm::vec3 &step) {
    std::vector<glm::vec3> out;
    for (int i = 0; i < dim.x; i++)
        for (int j = 0; j < dim.y; j++)
            for (int k = 0; k < dim.z; k++)
                out.push_back(add(start, mul(step, glm::vec3((float)i, (float)j, (float)k))));
    return out;
}
}  // namespace oldscene_detail

// scene.h:39-91
struct RigidBody {
    size_t numParticles = 0;
    std::vector<glm::vec3> positions;
    std::vector<glm::vec3> positions_CM_Origin;  // positions about the centre of mass
    glm::vec3 CM;
    glm::vec3 color;
    float massPerParticle = 0.f;

    static std::shared_ptr<RigidBody> CreateRigidBox(const glm::vec3 &color, const glm::ivec3 &dimension,
                                                     const glm::vec3 &startPosition, const glm::vec3 &stepSize,
                                                     const float massPerParticle) {
        using namespace oldscene_detail;
        auto r = std::make_shared<RigidBody>();
        r->color = color;
        r->massPerParticle = massPerParticle;
        r->positions = lattice(dimension, startPosition, stepSize);
        glm::vec3 cm(0.0f);
        for (const glm::vec3 &p : r->positions) cm = add(cm, p);
        cm = scale(cm, 1.0f / (float)r->positions.size());
        for (const glm::vec3 &p : r->positions) r->positions_CM_Origin.push_back(sub(p, cm));
        r->CM = cm;
        r->numParticles = r->positions.size();
        return r;
    }
};

// scene.h:93-120
struct Granulars {
    std::vector<glm::vec3> positions;
    float massPerParticle = 0.f;
    static std::shared_ptr<Granulars> CreateGranularsBlock(const glm::ivec3 &dimension,
                                                           const glm::vec3 &startPosition,
                                                           const glm::vec3 &stepSize, const float massPerParticle) {
        auto r = std::make_shared<Granulars>();
        r->massPerParticle = massPerParticle;
        r->positions = oldscene_detail::lattice(dimension, startPosition, stepSize);
        return r;
    }
};

// scene.h:122-147
struct Fluid {
    std::vector<glm::vec3> positions;
    float massPerParticle = 0.f;
    static std::shared_ptr<Fluid> CreateFluidBlock(const glm::ivec3 &dimension, const glm::vec3 &startPosition,
                                                   const glm::vec3 &stepSize, const float massPerParticle) {
        auto r = std::make_shared<Fluid>();
        r->massPerParticle = massPerParticle;
        r->positions = oldscene_detail::lattice(dimension, startPosition, stepSize);
        return r;
    }
};

// scene.h:149-185: a straight chain, links to the next joint (stiffness 1)
// and to the one after (twice the length, stiffness 0.1)
struct Rope {
    std::vector<glm::vec3> positions;
    std::vector<int2> distancePairs;
    std::vector<float2> distanceParams;  // (distance, stiffness)
    float massPerParticle = 0.f;
    static std::shared_ptr<Rope> CreateRope(const glm::vec3 &startPosition, const glm::vec3 &endPosition,
                                            const int numJoint, const float massPerParticle) {
        using namespace oldscene_detail;
        auto r = std::make_shared<Rope>();
        r->massPerParticle = massPerParticle;
        const glm::vec3 diff = sub(endPosition, startPosition);
        const float d = length(diff) / float(numJoint - 1);
        for (int i = 0; i < numJoint; i++)
            r->positions.push_back(add(startPosition, scale(diff, float(i) / float(numJoint - 1))));
        for (int i = 0; i < numJoint - 1; i++) {
            r->distancePairs.push_back(make_int2(i, i + 1));
            r->distanceParams.push_back(make_float2(d, 1.0f));
        }
        for (int i = 1; i < numJoint - 1; i++) {
            r->distancePairs.push_back(make_int2(i - 1, i + 1));
            r->distanceParams.push_back(make_float2(d * 2.0f, 0.1f));
        }
        return r;
    }
};

// scene.h:187-308: particle (x, y) at startPosition + y offsetX + x offsetY,
// stored x-major; constraint indices y * numJointX + x as the reference writes
// them (for a square cloth the two agree)
struct Cloth {
    std::vector<glm::vec3> positions;
    std::vector<int2> distancePairs;
    std::vector<float2> distanceParams;
    std::vector<int4> bendings;
    std::vector<int3> faces;
    std::vector<int> immovables;
    float massPerParticle = 0.f;
    static std::shared_ptr<Cloth> CreateCloth(const glm::vec3 &startPosition, const glm::vec3 &offsetX,
                                              const glm::vec3 &offsetY, const int numJointX, const int numJointY,
                                              const float massPerParticle, const bool firstCorner = false,
                                              const bool secondCorner = false, const bool thirdCorner = false,
                                              const bool fourthCorner = false) {
        using namespace oldscene_detail;
        const float stiffness = 0.2f, stiffness2 = 0.1f;
        auto r = std::make_shared<Cloth>();
        r->massPerParticle = massPerParticle;
        const float lx = length(offsetX), ly = length(offsetY), ld = length(add(offsetX, offsetY));
        if (firstCorner) r->immovables.push_back(0);
        if (secondCorner) r->immovables.push_back(numJointX - 1);
        if (thirdCorner) r->immovables.push_back(numJointX * (numJointY - 1));
        if (fourthCorner) r->immovables.push_back(numJointX * numJointY - 1);
        auto link = [&](int a, int b, float rest, float k) {
            r->distancePairs.push_back(make_int2(a, b));
            r->distanceParams.push_back(make_float2(rest, k));
        };
        for (int x = 0; x < numJointX; x++)
            for (int y = 0; y < numJointY; y++) {
                r->positions.push_back(add(startPosition, add(scale(offsetX, (float)y), scale(offsetY, (float)x))));
                const int p1 = y * numJointX + x, p2 = y * numJointX + x + 1;
                const int p3 = (y + 1) * numJointX + x, p4 = (y + 1) * numJointX + x + 1;
                if (x < numJointX - 1 && y < numJointY - 1) {
                    r->faces.push_back(make_int3(p1, p2, p3));
                    r->faces.push_back(make_int3(p2, p3, p4));
                }
                if (x < numJointX - 1) link(p1, p2, lx, stiffness);
                if (y < numJointY - 1) link(p1, p3, ly, stiffness);
                if (x < numJointX - 1 && y < numJointY - 1) {
                    link(p1, p4, ld, stiffness);
                    link(p2, p3, ld, stiffness);
                }
                if (x > 0 && y > 0 && x < numJointX - 1 && y < numJointY - 1) {  // "hack bending"
                    for (int i = -1; i <= 1; i++) link((y + i) * numJointX + x - 1, (y + i) * numJointX + x + 1,
                                                       lx * 2, stiffness2);
                    for (int j = -1; j <= 1; j++) link((y - 1) * numJointX + x + j, (y + 1) * numJointX + x + j,
                                                       ly * 2, stiffness2);
                }
            }
        return r;
    }
};

// scene.h:403-435 without the render-only camera and light
struct OldSceneFormat {
    std::vector<Plane> planes;
    std::vector<std::shared_ptr<RigidBody>> rigidBodies;
    std::vector<std::shared_ptr<Granulars>> granulars;
    std::vector<std::shared_ptr<Fluid>> fluids;
    std::vector<std::shared_ptr<Rope>> ropes;
    std::vector<std::shared_ptr<Cloth>> clothes;
    float fluidRestDensity = 0.f;
    size_t numParticles = 0, numMaxParticles = 0;
    size_t numRigidBodies = 0, numMaxRigidBodies = 0;
    size_t numDistancePairs = 0, numMaxDistancePairs = 0;
    size_t numBendings = 0, numMaxBendings = 0;
    size_t numWindFaces = 0, numMaxWindFaces = 0;
    size_t numImmovables = 0, numMaxImmovables = 0;
    float radius = 0.f;
};

// The particle system the legacy constructor assembles, as a Scene.
inline Scene assembleOldScene(OldSceneFormat &o) {
    using namespace oldscene_detail;
    auto fail = [](const char *what) { throw std::runtime_error(std::string("brasspan: ") + what); };
    Scene s;
    s.particleRadius = o.radius;
    s.fluidKernelRadius = 2.3f * o.radius;
    s.cellSize = make_float3(o.radius * 2.3f, o.radius * 2.3f, o.radius * 2.3f);
    s.gridDims = make_int3(512, 512, 512);
    s.planes = o.planes;
    s.groupIdCounter = 0;
    int solid = 1, fluid = -1;
    auto room = [&](size_t n) {
        if (o.numParticles + n >= o.numMaxParticles) fail("num particles exceed num max particles");
    };
    auto push = [&](const std::vector<glm::vec3> &pos, float mass, int group, int phase, bool perParticle) {
        for (size_t i = 0; i < pos.size(); i++) {
            s.positions.push_back(make_float3(pos[i].x, pos[i].y, pos[i].z));
            s.masses.push_back(mass);
            s.phases.push_back(perParticle ? phase + (int)i : phase);
            s.groupIds.push_back(group);
        }
    };
    auto links = [&](const std::vector<int2> &pairs, const std::vector<float2> &params, int base) {
        for (size_t i = 0; i < pairs.size(); i++) {
            s.distancePairs.push_back(make_int2(pairs[i].x + base, pairs[i].y + base));
            s.distanceParams.push_back(params[i]);
        }
    };
    for (const auto &rb : o.rigidBodies) {  // addRigidBody, :424-473
        room(rb->positions.size());
        if (o.numRigidBodies + 1 >= o.numMaxRigidBodies) fail("num rigid bodies exceed num max rigid bodies");
        glm::vec3 cm(0.0f);
        for (const glm::vec3 &p : rb->positions_CM_Origin) cm = add(cm, p);
        cm = scale(cm, 1.0f / (float)rb->positions_CM_Origin.size());
        if (length(cm) >= 1e-5f) fail("expected Center of Mass at the origin");
        const int base = (int)o.numParticles;
        push(rb->positions, rb->massPerParticle, s.groupIdCounter++, solid++, false);
        if (s.rigidbodyInitialPositions.size() < (size_t)base + rb->positions.size())
            s.rigidbodyInitialPositions.resize((size_t)base + rb->positions.size(), make_float3(0, 0, 0));
        for (size_t i = 0; i < rb->positions_CM_Origin.size(); i++) {
            const glm::vec3 &q = rb->positions_CM_Origin[i];
            s.rigidbodyInitialPositions[base + i] = make_float3(q.x, q.y, q.z);
        }
        s.rigidbodyParticleIdRanges.push_back(make_int2(base, base + (int)rb->positions.size()));
        o.numParticles += rb->positions.size();
        o.numRigidBodies += 1;
    }
    for (const auto &g : o.granulars) {  // addGranulars, :395-422
        room(g->positions.size());
        push(g->positions, g->massPerParticle, s.groupIdCounter++, solid, true);
        solid += (int)g->positions.size();
        o.numParticles += g->positions.size();
    }
    for (const auto &f : o.fluids) {  // addFluids, :475-503
        room(f->positions.size());
        push(f->positions, f->massPerParticle, s.groupIdCounter++, fluid--, false);
        o.numParticles += f->positions.size();
    }
    for (const auto &r : o.ropes) {  // addNoodles, :505-565 (self-colliding)
        room(r->positions.size());
        if (o.numParticles + r->distancePairs.size() >= o.numMaxDistancePairs)
            fail("num distance pairs exceed num max distance pairs");
        const int base = (int)o.numParticles;
        push(r->positions, r->massPerParticle, s.groupIdCounter++, solid, true);
        solid += (int)r->positions.size();
        links(r->distancePairs, r->distanceParams, base);
        o.numParticles += r->positions.size();
        o.numDistancePairs += r->distancePairs.size();
    }
    for (const auto &c : o.clothes) {  // addCloth, :567-710 (self-colliding)
        room(c->positions.size());
        if (o.numDistancePairs + c->distancePairs.size() >= o.numMaxDistancePairs)
            fail("num distance pairs exceed num max distance pairs");
        if (o.numBendings + c->bendings.size() >= o.numMaxBendings) fail("num bendings exceed num max bendings");
        if (o.numWindFaces + c->faces.size() >= o.numMaxWindFaces) fail("num faces exceed num max faces");
        if (o.numImmovables + c->immovables.size() >= o.numMaxImmovables)
            fail("num immovables exceed num max immovables");
        const int base = (int)o.numParticles;
        push(c->positions, c->massPerParticle, s.groupIdCounter++, solid, true);
        solid += (int)c->positions.size();
        links(c->distancePairs, c->distanceParams, base);
        for (const int4 &b : c->bendings)
            s.bendingConstraints.push_back(make_int4(b.x + base, b.y + base, b.z + base, b.w + base));
        for (const int3 &f : c->faces) s.faces.push_back(make_int3(f.x + base, f.y + base, f.z + base));
        for (int id : c->immovables) s.immovableConstraints.push_back(id + base);
        o.numImmovables += c->immovables.size();
        o.numBendings += c->bendings.size();
        o.numParticles += c->positions.size();
        o.numDistancePairs += c->distancePairs.size();
        o.numWindFaces += c->faces.size();
    }
    s.solidPhaseCounter = solid;
    s.fluidPhaseCounter = fluid;
    return s;
}
