// batch_and_grad.cpp -- three flags under different winds advanced as one
// batch, then the gradient of a final-state loss with respect to each
// scene's wind, drag, lift and XSPH c.  Headless; prints a summary.
#include <cstdio>

#include "brasspan/particlesolver.h"

static std::shared_ptr<Scene> flag_scene() {
    auto scene = std::make_shared<Scene>();
    scene->particleRadius = 0.025f;
    scene->fluidKernelRadius = 0.1f;
    scene->cellSize = make_float3(0.1f, 0.1f, 0.1f);
    scene->gridDims = make_int3(32, 32, 32);
    scene->addFluidBlock(make_uint3(6, 6, 6), make_float3(0.035f, 0.035f, 0.035f), glm::vec3(0.05f, 0, 0),
                         glm::vec3(0, 0.05f, 0), 1.0f);
    const int n = 10;
    const int2 cloth = scene->addCloth(make_float3(0.2f, 1.5f, 1.6f), make_float3(0.05f, 0.f, 0.f),
                                       make_float3(0.f, 0.05f, 0.f), n, n, 0.5f, 0.2f, 0.1f, false);
    for (int y = 0; y < n; ++y) scene->makeImmovable(cloth.x + y);
    const glm::vec3 lo(0.0f), hi(3.2f);
    const glm::vec3 axes[3] = {glm::vec3(1, 0, 0), glm::vec3(0, 1, 0), glm::vec3(0, 0, 1)};
    for (const glm::vec3 &a : axes) {
        scene->planes.push_back(Plane(lo, a));
        scene->planes.push_back(Plane(hi, glm::vec3(-a.x, -a.y, -a.z)));
    }
    return scene;
}

int main() {
    std::vector<std::shared_ptr<Scene>> scenes;
    std::vector<bp_params> params;
    for (int k = 0; k < 3; ++k) {
        scenes.push_back(flag_scene());
        bp_params p;
        bp_default_params(&p);
        p.iterations = 4;
        p.akinci = 0;
        p.vorticity_eps = 0.0f;
        p.xsph_c = 2e-5f;
        p.friction_static = p.friction_dynamic = 0.0f;
        p.fluid_rest_density = 4822.44f;
        p.wind[0] = 2.0f + 2.0f * (float)k;
        params.push_back(p);
    }
    try {
        BatchParticleSolver batch(scenes, params);
        batch.forwardCheckpointed(8, 1.0f / 120.0f, 4);
        std::vector<float3> pos, vel;
        batch.copyStateOriginalOrder(pos, vel);
        // loss = mean |x_T - x_T0 - (0, 0, 0.1)|^2: pull every particle 10 cm along z
        std::vector<float3> gx(pos.size()), gx0, gv0;
        std::vector<float> gtheta;
        for (size_t i = 0; i < pos.size(); ++i) gx[i] = make_float3(0.f, 0.f, -0.2f / (float)pos.size());
        batch.backward(gx, {}, gx0, gv0, gtheta);
        for (int k = 0; k < batch.numScenes; ++k)
            std::printf("scene %d: dL/dwind = (%.3e, %.3e, %.3e) dL/ddrag = %.3e dL/dlift = %.3e dL/dxsph = %.3e\n", k,
                        gtheta[6 * k], gtheta[6 * k + 1], gtheta[6 * k + 2], gtheta[6 * k + 3], gtheta[6 * k + 4],
                        gtheta[6 * k + 5]);
        std::printf("batch_and_grad ok: %zu particles in %d scenes\n", batch.numParticles, batch.numScenes);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
