// legacy_scenes.cpp -- the reference's legacy scene format driven headless:
// scenes shaped like main.cu's initRigidBodiesScene, initFluidScene and
// initRopesScene (main.cu:140-255), built as OldSceneFormat and handed to
// ParticleSolver(const std::shared_ptr<OldSceneFormat> &).  Usage:
// legacy_scenes [frames] [frame file prefix] [dump]; with a prefix, each
// scene's last frame is written to <prefix><name>.bpfr, and with "dump" also
// the assembled scene to <prefix><name>.bpsc (scenefile.h).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "brasspan/particlesolver.h"
#include "brasspan/scenefile.h"

static std::shared_ptr<OldSceneFormat> base(float radius, size_t maxParticles) {
    auto s = std::make_shared<OldSceneFormat>();
    s->radius = radius;
    s->numMaxParticles = maxParticles;
    s->numMaxRigidBodies = 128;
    s->numMaxDistancePairs = 50000;
    s->numMaxBendings = 1000;
    s->numMaxWindFaces = 10000;
    s->numMaxImmovables = 10;
    s->planes.push_back(Plane(glm::vec3(0, 0, 0), glm::vec3(0, 1, 0)));
    return s;
}

int main(int argc, char **argv) {
    const int frames = argc > 1 ? std::atoi(argv[1]) : 20;
    const std::string prefix = argc > 2 ? argv[2] : "";
    const bool dump = argc > 3 && std::string(argv[3]) == "dump";
    const float r = 0.05f;
    const glm::vec3 d(2 * r);
    std::vector<std::pair<std::string, std::shared_ptr<OldSceneFormat>>> scenes;
    {  // a stack of rigid boxes (main.cu:140-165)
        auto s = base(r, 50000);
        s->rigidBodies.push_back(RigidBody::CreateRigidBox(glm::vec3(1.f), glm::ivec3(3, 3, 3),
                                                           glm::vec3(-r, r + 2, -r), d, 2.0f));
        s->rigidBodies.push_back(RigidBody::CreateRigidBox(glm::vec3(1.f), glm::ivec3(3, 4, 2),
                                                           glm::vec3(-r, r + 1, -r), d, 1.5f));
        s->rigidBodies.push_back(RigidBody::CreateRigidBox(glm::vec3(1.f), glm::ivec3(3, 4, 2),
                                                           glm::vec3(-r, r + 3, -r), d, 1.0f));
        scenes.push_back({"rigid", s});
    }
    {  // a fluid block (main.cu:188-227), smaller
        auto s = base(r, 60000);
        s->fluids.push_back(Fluid::CreateFluidBlock(glm::ivec3(16, 16, 16), glm::vec3(0.5f, r, 0.5f), d, 0.75f));
        s->fluidRestDensity = 1200.0f;
        scenes.push_back({"fluid", s});
    }
    {  // two pinned cloths and a rope (main.cu:229-255)
        auto s = base(r, 10000);
        s->ropes.push_back(Rope::CreateRope(glm::vec3(0.f, 1.0f, 0.f), glm::vec3(0.f, 3.0f, 0.f), 20, 1.0f));
        s->clothes.push_back(Cloth::CreateCloth(glm::vec3(2.f, 2.f, 2.f), glm::vec3(0.1f, 0.f, 0.f),
                                                glm::vec3(0.f, 0.f, 0.1f), 10, 10, 1.0f, true, true));
        s->clothes.push_back(Cloth::CreateCloth(glm::vec3(5.f, 2.f, 5.f), glm::vec3(0.1f, 0.f, 0.f),
                                                glm::vec3(0.f, 0.1f, 0.0f), 10, 10, 1.0f, true));
        scenes.push_back({"ropes", s});
    }
    try {
        for (auto &[name, s] : scenes) {
            if (dump && !prefix.empty()) {  // the assembly on a copy: it advances the scene's counters
                OldSceneFormat copy = *s;
                const bool fluid = s->fluidRestDensity > 0.0f || !s->fluids.empty();
                writeSceneFile(prefix + name + ".bpsc", assembleOldScene(copy), fluid ? s->fluidRestDensity : -1.0f);
            }
            bp_params p;
            bp_default_params(&p);
            ParticleSolver solver(s, &p);
            for (int f = 0; f < frames; ++f) solver.update(2, 1.0f / 60.0f);
            std::vector<float3> pos, vel;
            solver.copyStateOriginalOrder(pos, vel);
            double ymin = 1e30, ymean = 0;
            bool finite = true;
            for (const float3 &q : pos) {
                finite = finite && std::isfinite(q.x) && std::isfinite(q.y) && std::isfinite(q.z);
                ymin = std::fmin(ymin, q.y);
                ymean += q.y;
            }
            std::printf("%s particles %zu rigid %zu pairs %zu faces %zu pins %zu finite %d ymin %.4f ymean %.4f\n",
                        name.c_str(), s->numParticles, s->numRigidBodies, s->numDistancePairs, s->numWindFaces,
                        s->numImmovables, finite ? 1 : 0, ymin, ymean / pos.size());
            if (!prefix.empty()) solver.writeFrame(prefix + name + ".bpfr", (uint64_t)frames, frames / 60.0);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
