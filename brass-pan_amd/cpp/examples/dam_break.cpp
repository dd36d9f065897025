// dam_break.cpp -- headless use of the Scene / ParticleSolver API, as
// main.cu:293-435 drives the reference (without the renderer): build a fluid
// block in a box, step frames with update(2, 1/60), print the state.
// Usage: dam_break [frames] [frame file]; with a frame file, the final state
// is written there in the frame format of brass-pan_amd/frames.py (what the
// reference would hand its renderer).
#include <cstdio>
#include <memory>
#include <vector>

#include "brasspan/particlesolver.h"

int main(int argc, char **argv) {
    const int frames = argc > 1 ? std::atoi(argv[1]) : 10;
    auto scene = std::make_shared<Scene>();
    scene->particleRadius = 0.025f;
    scene->fluidKernelRadius = 0.1f;
    scene->cellSize = make_float3(0.1f, 0.1f, 0.1f);
    scene->gridDims = make_int3(64, 64, 64);
    scene->addFluidBlock(make_uint3(16, 16, 16), make_float3(0.025f, 0.025f, 0.025f), glm::vec3(0.05f, 0, 0),
                         glm::vec3(0, 0.05f, 0), 1.0f);
    const float box = 6.4f;
    scene->planes.push_back(Plane(glm::vec3(0, 0, 0), glm::vec3(1, 0, 0)));
    scene->planes.push_back(Plane(glm::vec3(box, box, box), glm::vec3(-1, 0, 0)));
    scene->planes.push_back(Plane(glm::vec3(0, 0, 0), glm::vec3(0, 1, 0)));
    scene->planes.push_back(Plane(glm::vec3(box, box, box), glm::vec3(0, -1, 0)));
    scene->planes.push_back(Plane(glm::vec3(0, 0, 0), glm::vec3(0, 0, 1)));
    scene->planes.push_back(Plane(glm::vec3(box, box, box), glm::vec3(0, 0, -1)));
    try {
        bp_params p;
        bp_default_params(&p);
        p.iterations = 4;
        p.fluid_rest_density = 4822.4414f;  // mean half-sum density of the lattice
        ParticleSolver solver(scene, &p);
        for (int f = 0; f < frames; ++f) solver.update(2, 1.0f / 60.0f);
        std::vector<float3> pos, vel;
        solver.copyStateOriginalOrder(pos, vel);
        double ymean = 0;
        for (const float3 &q : pos) ymean += q.y;
        std::printf("particles %zu frames %d mean height %.4f slot(0)->orig %d\n", solver.numParticles, frames,
                    ymean / pos.size(), solver.queryOriginalParticleId(0));
        if (argc > 2) solver.writeFrame(argv[2], (uint64_t)frames, frames / 60.0);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
