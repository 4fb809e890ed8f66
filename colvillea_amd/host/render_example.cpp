/*
 * render_example.cpp -- a C++ host driving the renderer through
 * hip_integrator.h in the reference's RenderEngine order
 * (renderengine.cpp:27-51: materials, entities, BLAS, TLAS, emitters,
 * camera, render, readback).
 *
 *   g++ -std=c++17 -I<repo> render_example.cpp -L<repo>/colvillea_amd/lib -lclpt_hip -o render_example
 *   ./render_example [width height spp]
 *
 * Prints the mean radiance and the per-depth ray counts; exit code 0 on
 * success, 2 when the renderer reports an error (e.g. no HIP device).
 */
#include "hip_integrator.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>

int main(int argc, char** argv)
{
    const uint32_t w = argc > 1 ? (uint32_t)atoi(argv[1]) : 64;
    const uint32_t h = argc > 2 ? (uint32_t)atoi(argv[2]) : 48;
    const uint32_t spp = argc > 3 ? (uint32_t)atoi(argv[3]) : 4;
    try
    {
        auto integ = clhip::HipWavefrontPathTracingIntegrator::createIntegrator(
            clhip::IntegratorType::WavefrontPathTracing, w, h);

        /* ground quad and a unit box */
        const float    quad_v[] = {-4, 0, -4, 4, 0, -4, 4, 0, 4, -4, 0, 4};
        const uint32_t quad_i[] = {0, 2, 1, 0, 3, 2};
        float          box_v[24];
        for (int c = 0; c < 8; ++c)
        {
            box_v[3 * c]     = (c & 1) ? 0.5f : -0.5f;
            box_v[3 * c + 1] = (c & 2) ? 1.0f : 0.0f;
            box_v[3 * c + 2] = (c & 4) ? 0.5f : -0.5f;
        }
        const uint32_t box_i[] = {0, 1, 3, 0, 3, 2, 4, 6, 7, 4, 7, 5, 0, 4, 5, 0, 5, 1,
                                  2, 3, 7, 2, 7, 6, 0, 2, 6, 0, 6, 4, 1, 5, 7, 1, 7, 3};
        clhip::TriangleMeshView quad{quad_v, quad_i, nullptr, nullptr, nullptr, 4, 2};
        clhip::TriangleMeshView box{box_v, box_i, nullptr, nullptr, nullptr, 8, 12};

        clhip_material diffuse{};
        diffuse.type           = CLHIP_MTL_DIFFUSE;
        diffuse.reflectance[0] = diffuse.reflectance[1] = diffuse.reflectance[2] = 0.6f;
        clhip_material metal{};
        metal.type      = CLHIP_MTL_METAL;
        metal.roughness = 0.4f;
        for (int k = 0; k < 3; ++k)
        {
            metal.reflectance[k] = 1.0f;
            metal.eta[k]         = 0.2f;
            metal.k[k]           = 3.9f;
        }
        integ->buildMaterials({diffuse, metal});
        integ->buildGeometryEntities({0u, 1u});
        integ->buildBLAS({quad, box});
        integ->buildTLAS({0u, 1u}, {0u, 1u});

        std::vector<float> sky(16 * 8 * 3, 1.0f);
        clhip_emitter      sun{};
        sun.type           = CLHIP_EMITTER_DIRECTIONAL;
        sun.intensity[0] = sun.intensity[1] = sun.intensity[2] = 5.0f;
        sun.direction[0]   = -0.3f;
        sun.direction[1]   = -0.9f;
        sun.direction[2]   = -0.3f;
        sun.angular_radius = 0.00873f;
        sun.texture        = -1;
        clhip_emitter dome{};
        dome.type    = CLHIP_EMITTER_HDRIDOME;
        dome.texture = integ->createTexture(sky.data(), 16, 8, 3);
        integ->buildEmitters({sun, dome}, 1);

        /* camera at (0, 2, 6) looking at the origin, 45 degree field of view */
        const float t = 2.0f * std::tan(0.5f * 45.0f * 3.14159265f / 180.0f);
        const float a = (float)w / (float)h;
        clhip::vec3f fwd{0.0f, -0.316228f, -0.948683f};
        clhip::vec3f ddu{t * a, 0.0f, 0.0f};
        clhip::vec3f ddv{0.0f, t * 0.948683f, t * -0.316228f};
        clhip::Camera cam{{0.0f, 2.0f, 6.0f},
                          {fwd[0] - 0.5f * ddu[0] - 0.5f * ddv[0], fwd[1] - 0.5f * ddu[1] - 0.5f * ddv[1],
                           fwd[2] - 0.5f * ddu[2] - 0.5f * ddv[2]},
                          ddu,
                          ddv};
        integ->updateCamera(cam);
        integ->setParams(clhip_params{5u, 3u, 0u, 0u});
        for (uint32_t s = 0; s < spp; ++s)
            integ->render(1); /* one progressive spp per call, as in the reference viewer loop */
        auto   fb   = integ->readbackFramebuffer();
        double m[3] = {0, 0, 0};
        for (size_t i = 0; i < (size_t)w * h; ++i)
            for (int k = 0; k < 3; ++k)
                m[k] += fb[i][k];
        const clhip_stats st = integ->counters();
        printf("mean %.6f %.6f %.6f\n", m[0] / (w * h), m[1] / (w * h), m[2] / (w * h));
        printf("rays");
        for (int b = 1; b <= 5; ++b)
            printf(" %llu", (unsigned long long)st.rays[b]);
        printf("\n");
        return 0;
    }
    catch (const clhip::Error& e)
    {
        fprintf(stderr, "clhip error %d: %s\n", e.code(), e.what());
        return 2;
    }
}
