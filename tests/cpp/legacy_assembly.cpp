// Host-only check of the legacy scene assembly (brass-pan_amd/cpp/brasspan/
// oldscene.h) against the rules of the reference's legacy constructor
// (particlesolver.cuh:218-342, 395-710).  Prints "legacy assembly ok".
// Also prints the bits of the hash-cell edge and kernel radius a Scene hands
// the solver (default radius, and radius 0.025 set after construction), and
// with a path argument writes the assembled scene there as a scene file.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "brasspan/oldscene.h"
#include "brasspan/scenefile.h"

#include <cstring>

static int failures = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);    \
            ++failures;                                                   \
        }                                                                 \
    } while (0)

static std::shared_ptr<OldSceneFormat> sample(float r) {
    auto o = std::make_shared<OldSceneFormat>();
    o->radius = r;
    o->numMaxParticles = 1000;
    o->numMaxRigidBodies = 16;
    o->numMaxDistancePairs = 1000;
    o->numMaxBendings = 10;
    o->numMaxWindFaces = 1000;
    o->numMaxImmovables = 10;
    o->fluidRestDensity = 1200.0f;
    const glm::vec3 d(2 * r);
    o->rigidBodies.push_back(RigidBody::CreateRigidBox(glm::vec3(1.f), glm::ivec3(3, 3, 3), glm::vec3(0, 1, 0), d, 2.f));
    o->rigidBodies.push_back(RigidBody::CreateRigidBox(glm::vec3(1.f), glm::ivec3(3, 4, 2), glm::vec3(1, 1, 0), d, 1.f));
    o->granulars.push_back(Granulars::CreateGranularsBlock(glm::ivec3(2, 2, 2), glm::vec3(0, 2, 0), d, 1.f));
    o->fluids.push_back(Fluid::CreateFluidBlock(glm::ivec3(3, 3, 3), glm::vec3(2, 0.5f, 0), d, 0.75f));
    o->fluids.push_back(Fluid::CreateFluidBlock(glm::ivec3(3, 3, 3), glm::vec3(3, 0.5f, 0), d, 0.75f));
    o->ropes.push_back(Rope::CreateRope(glm::vec3(0, 3, 0), glm::vec3(0, 4.8f, 0), 10, 1.f));
    o->clothes.push_back(Cloth::CreateCloth(glm::vec3(2, 3, 2), glm::vec3(0.1f, 0, 0), glm::vec3(0, 0, 0.1f), 6, 6,
                                            1.f, true, true));
    return o;
}

static unsigned bits(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    return u;
}

int main(int argc, char **argv) {
    {  // Scene defaults: h = 2.3 * radius in double at construction, cell = radius * 2.3f at solver build
        Scene a;
        Scene b;
        b.particleRadius = 0.025f;
        Scene c;
        c.cellSize = make_float3(0.1f, 0.2f, 0.3f);
        std::printf("scene defaults %08x %08x %08x %08x %08x %08x\n", bits(a.effectiveCellSize().x),
                    bits(a.fluidKernelRadius), bits(b.effectiveCellSize().x), bits(b.fluidKernelRadius),
                    bits(c.effectiveCellSize().y), bits(b.effectiveCellSize().z));
    }
    const float r = 0.05f;
    auto o = sample(r);
    const Scene s = assembleOldScene(*o);
    // order and offsets: rigid 27 + 24, granular 8, fluids 27 + 27, rope 10, cloth 36
    CHECK(s.positions.size() == 159 && o->numParticles == 159);
    CHECK(o->numRigidBodies == 2 && s.rigidbodyParticleIdRanges.size() == 2);
    CHECK(s.rigidbodyParticleIdRanges[0].x == 0 && s.rigidbodyParticleIdRanges[0].y == 27);
    CHECK(s.rigidbodyParticleIdRanges[1].x == 27 && s.rigidbodyParticleIdRanges[1].y == 51);
    // phases: one per rigid body, one per granular, rope and cloth particle, -1 / -2 per fluid block
    for (int i = 0; i < 27; i++) CHECK(s.phases[i] == 1);
    for (int i = 27; i < 51; i++) CHECK(s.phases[i] == 2);
    for (int i = 51; i < 59; i++) CHECK(s.phases[i] == 3 + (i - 51));
    for (int i = 59; i < 86; i++) CHECK(s.phases[i] == -1);
    for (int i = 86; i < 113; i++) CHECK(s.phases[i] == -2);
    for (int i = 113; i < 123; i++) CHECK(s.phases[i] == 11 + (i - 113));
    for (int i = 123; i < 159; i++) CHECK(s.phases[i] == 21 + (i - 123));
    // group ids 0.. per part (addNewGroup), masses per part
    const int firsts[] = {0, 27, 51, 59, 86, 113, 123, 159};
    for (int g = 0; g < 7; g++)
        for (int i = firsts[g]; i < firsts[g + 1]; i++) CHECK(s.groupIds[i] == g);
    CHECK(s.masses[0] == 2.f && s.masses[30] == 1.f && s.masses[60] == 0.75f && s.masses[150] == 1.f);
    // rigid rest offsets: about the body's centre, rest + CM = position
    for (int b = 0; b < 2; b++) {
        const int2 rg = s.rigidbodyParticleIdRanges[b];
        double cx = 0, cy = 0, cz = 0;
        for (int i = rg.x; i < rg.y; i++) {
            cx += s.rigidbodyInitialPositions[i].x;
            cy += s.rigidbodyInitialPositions[i].y;
            cz += s.rigidbodyInitialPositions[i].z;
        }
        CHECK(std::fabs(cx) < 1e-4 && std::fabs(cy) < 1e-4 && std::fabs(cz) < 1e-4);
        const glm::vec3 cm = o->rigidBodies[b]->CM;
        for (int i = rg.x; i < rg.y; i++)
            CHECK(std::fabs(s.rigidbodyInitialPositions[i].y + cm.y - s.positions[i].y) < 1e-6f);
    }
    // rope: 9 links to the next joint, 8 to the one after, offset by 113
    CHECK(s.distancePairs[0].x == 113 && s.distancePairs[0].y == 114);
    CHECK(std::fabs(s.distanceParams[0].x - 0.2f) < 1e-6f && s.distanceParams[0].y == 1.0f);
    CHECK(s.distancePairs[9].x == 113 && s.distancePairs[9].y == 115);
    CHECK(std::fabs(s.distanceParams[9].x - 0.4f) < 1e-6f && s.distanceParams[9].y == 0.1f);
    // cloth n x n: 2n(n-1) + 2(n-1)^2 + 6(n-2)^2 distance constraints, 2(n-1)^2 faces
    const int n = 6;
    CHECK(s.distancePairs.size() == 17 + (size_t)(2 * n * (n - 1) + 2 * (n - 1) * (n - 1) + 6 * (n - 2) * (n - 2)));
    CHECK(o->numDistancePairs == s.distancePairs.size());
    CHECK(s.faces.size() == (size_t)(2 * (n - 1) * (n - 1)) && o->numWindFaces == s.faces.size());
    CHECK(s.faces[0].x == 123 && s.faces[0].y == 124 && s.faces[0].z == 123 + n);
    for (const int2 &p : s.distancePairs) CHECK(p.x >= 113 && p.x < 159 && p.y >= 113 && p.y < 159);
    CHECK(s.immovableConstraints.size() == 2 && s.immovableConstraints[0] == 123 &&
          s.immovableConstraints[1] == 123 + n - 1);
    CHECK(o->numImmovables == 2 && o->numBendings == 0 && s.bendingConstraints.empty());
    // grid and kernel radius of the legacy constructor
    CHECK(std::fabs(s.cellSize.x - 2.3f * r) < 1e-7f && std::fabs(s.fluidKernelRadius - 2.3f * r) < 1e-7f);
    CHECK(s.gridDims.x == 512 && s.gridDims.y == 512 && s.gridDims.z == 512 && s.particleRadius == r);
    // capacity checks: the total must stay strictly below numMax*
    auto throws = [](std::shared_ptr<OldSceneFormat> q, const char *what) {
        try {
            assembleOldScene(*q);
        } catch (const std::runtime_error &e) {
            return std::string(e.what()).find(what) != std::string::npos;
        }
        return false;
    };
    auto q = sample(r);
    q->numMaxParticles = 159;
    CHECK(throws(q, "num particles exceed num max particles"));
    q = sample(r);
    q->numMaxParticles = 160;
    CHECK(!throws(q, "exceed"));
    q = sample(r);
    q->numMaxRigidBodies = 2;
    CHECK(throws(q, "num rigid bodies exceed num max rigid bodies"));
    q = sample(r);
    q->numMaxImmovables = 2;
    CHECK(throws(q, "num immovables exceed num max immovables"));
    q = sample(r);  // the rope's check compares particles + pairs (particlesolver.cuh:515)
    q->numMaxDistancePairs = 113 + 17;
    CHECK(throws(q, "num distance pairs exceed num max distance pairs"));
    q = sample(r);
    for (glm::vec3 &p : q->rigidBodies[0]->positions_CM_Origin) p.x += 0.01f;
    CHECK(throws(q, "expected Center of Mass at the origin"));
    if (argc > 1) writeSceneFile(argv[1], s, o->fluidRestDensity);
    if (failures) return 1;
    std::printf("legacy assembly ok\n");
    return 0;
}
