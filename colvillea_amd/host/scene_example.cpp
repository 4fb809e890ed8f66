/*
 * scene_example.cpp -- a reference-style host program: build a Scene, hand
 * it to a RenderEngine over the HIP integrator and render progressively, as
 * Colvillea's viewer loop does (src/colvillea/CLViewer.cpp:604-738 calls
 * RenderEngine::startRendering once per frame).
 *
 *   scene_example --selftest   edit-action bookkeeping and compile outputs
 *                              (no device needed)
 *   scene_example [frames]     render through the engine and check it
 *                              against direct integrator calls on the same
 *                              data, then edit a material and a transform
 *
 * Exit status 0 = all checks passed, 1 = a check failed, 2 = error (e.g.
 * no HIP device).
 */
#include "hip_scene.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

using namespace clhip;

namespace
{
int g_fail = 0;
void expect(bool ok, const char* what)
{
    if (!ok)
    {
        std::fprintf(stderr, "check failed: %s\n", what);
        ++g_fail;
    }
}

/* a unit quad in y = 0 scaled by `half`, and an octahedron */
std::shared_ptr<TriangleMesh> ground(Scene& s, float half)
{
    return s.createTriangleMesh({-half, 0, -half, half, 0, -half, half, 0, half, -half, 0, half}, {0, 2, 1, 0, 3, 2},
                                {0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0});
}
std::shared_ptr<TriangleMesh> octahedron(Scene& s)
{
    return s.createTriangleMesh({1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1},
                                {0, 2, 4, 2, 1, 4, 1, 3, 4, 3, 0, 4, 2, 0, 5, 1, 2, 5, 3, 1, 5, 0, 3, 5});
}
std::vector<float> sky(uint32_t w, uint32_t h)
{
    std::vector<float> img(3 * (size_t)w * h);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x)
        {
            const float t = (float)y / (float)(h - 1);
            float*      p = &img[3 * ((size_t)y * w + x)];
            p[0]          = 0.4f + 0.6f * (1.0f - t);
            p[1]          = 0.5f + 0.5f * (1.0f - t);
            p[2]          = 0.9f;
        }
    return img;
}
Transform place(float x, float y, float z, float s) { return {s, 0, 0, x, 0, s, 0, y, 0, 0, s, z}; }
Camera    camera(uint32_t w, uint32_t h)
{
    Camera c;
    c.m_camera_pos = {0.0f, 2.0f, 6.0f};
    c.m_camera_d00 = {-0.5f * (float)w / (float)h, 0.3f, -1.0f};
    c.m_camera_ddu = {(float)w / (float)h, 0.0f, 0.0f};
    c.m_camera_ddv = {0.0f, -0.8f, 0.0f};
    return c;
}

struct Built
{
    std::shared_ptr<Scene>        scene;
    std::shared_ptr<Material>     floor;
    std::shared_ptr<Entity>       prop;
    std::shared_ptr<TriangleMesh> prop_mesh;
};

Built build_scene(uint32_t w, uint32_t h)
{
    Built b;
    b.scene   = Scene::createScene();
    Scene& s  = *b.scene;
    auto   g  = ground(s, 5.0f);
    auto   o  = octahedron(s);
    b.prop_mesh = o;
    b.floor   = s.createDiffuseMaterial({0.6f, 0.6f, 0.55f});
    auto m1   = s.createMetalMaterial({1, 1, 1}, 0.3f, {0.2f, 0.9f, 1.1f}, {3.9f, 2.4f, 2.2f});
    auto m2   = s.createGlassMaterial(0.1f, 1.5f);
    s.createEntity(g, b.floor);
    b.prop = s.createEntity(o, m1, place(-1.0f, 1.0f, 0.0f, 0.8f));
    s.createEntity(o, m2, place(1.2f, 0.8f, 0.5f, 0.6f));
    auto tex = s.createTexture(sky(64, 32), 64, 32, 3);
    s.createEmitter(tex);
    s.createEmitter({5.0f, 4.8f, 4.5f}, {-0.3f, -0.8f, -0.5f}, 0.0087f);
    s.setCamera(camera(w, h));
    return b;
}

int selftest()
{
    Built b = build_scene(32, 24);
    Scene& s = *b.scene;
    const SceneEditAction& a = s.editActions();
    expect(a.hasAction(Action::ShapeAdded) && a.hasAction(Action::EntityAdded) &&
               a.hasAction(Action::MaterialAdded) && a.hasAction(Action::EmitterAdded) &&
               a.hasAction(Action::TexturesAdded),
           "creation records the Added actions");
    auto blas = s.collectTriangleMeshForBLASBuilding();
    auto tlas = s.collectTriangleMeshForTLASBuilding();
    auto ents = s.compileEntity();
    auto mats = s.compileMaterials();
    auto emit = s.compileEmitters();
    expect(blas && blas->meshes.size() == 2 && blas->meshes[1].numTriangles == 8 &&
               blas->meshIndex == std::vector<uint32_t>({0, 1}),
           "BLAS source: both meshes");
    s.markBLASBuilt(blas->meshIndex);
    expect(tlas && tlas->meshOfInstance == std::vector<uint32_t>({0, 1, 1}) &&
               tlas->instanceIDs == std::vector<uint32_t>({0, 1, 2}) && tlas->transforms.size() == 36 &&
               tlas->transforms[12 + 3] == -1.0f,
           "TLAS source: mesh per instance, entity ids, transforms");
    expect(ents && *ents == std::vector<uint32_t>({0, 1, 2}), "entity -> material index");
    expect(mats && mats->size() == 3 && (*mats)[2].type == CLHIP_MTL_GLASS, "materials in creation order");
    expect(emit && emit->emitters.size() == 2 && emit->domeIndex == 0, "emitters and dome index");
    s.resetSceneEditActions();
    expect(!s.collectTriangleMeshForBLASBuilding() && !s.collectTriangleMeshForTLASBuilding() && !s.compileEntity() &&
               !s.compileMaterials() && !s.compileEmitters(),
           "nothing to compile after the actions are reset");
    b.floor->setReflectance({0.1f, 0.2f, 0.3f});
    expect(s.compileMaterials() && !s.collectTriangleMeshForBLASBuilding() && !s.compileEmitters(),
           "a material edit recompiles materials only");
    s.resetSceneEditActions();
    b.prop->setTransform(place(0, 2, 0, 1));
    auto t2 = s.collectTriangleMeshForTLASBuilding();
    expect(t2 && t2->transforms[12 + 7] == 2.0f && !s.collectTriangleMeshForBLASBuilding(),
           "a transform edit rebuilds the TLAS only");
    s.resetSceneEditActions();
    b.prop_mesh->setVertices(std::vector<float>(18, 0.5f));
    auto b3 = s.collectTriangleMeshForBLASBuilding();
    expect(b3 && b3->meshIndex == std::vector<uint32_t>({1}) && s.collectTriangleMeshForTLASBuilding(),
           "a shape edit rebuilds that mesh's BLAS only, then the TLAS");
    std::printf("selftest %s\n", g_fail ? "FAILED" : "ok");
    return g_fail ? 1 : 0;
}

double film_mean(HipWavefrontPathTracingIntegrator& it)
{
    const auto sz = it.getFilmSize();
    auto       fb = it.readbackFramebuffer();
    double     m  = 0.0;
    for (size_t i = 0; i < (size_t)sz.first * sz.second; ++i)
        m += fb[i][0] + fb[i][1] + fb[i][2];
    return m / (3.0 * sz.first * sz.second);
}

int render(int frames)
{
    const uint32_t W = 96, H = 64;
    Built          b = build_scene(W, H);
    std::shared_ptr<HipWavefrontPathTracingIntegrator> integ =
        HipWavefrontPathTracingIntegrator::createIntegrator(IntegratorType::WavefrontPathTracing, W, H);
    integ->setParams(clhip_params{5u, 0u, 0u, 0u});
    auto engine = RenderEngine::createRenderEngine(integ, b.scene);
    for (int f = 0; f < frames; ++f)
        engine->startRendering();
    engine->endRendering();
    uint32_t spp = 0;
    clhip_get_film_size(integ->native(), nullptr, nullptr, &spp);
    expect(spp == (uint32_t)frames, "progressive accumulation across frames");
    std::vector<double> a(3 * (size_t)W * H), c(3 * (size_t)W * H);
    clhip_readback_sums(integ->native(), a.data(), a.size() * sizeof(double));

    /* the same data through direct integrator calls */
    auto direct = HipWavefrontPathTracingIntegrator::createIntegrator(IntegratorType::WavefrontPathTracing, W, H);
    direct->setParams(clhip_params{5u, 0u, 0u, 0u});
    Scene& s = *b.scene;
    (void)s;
    {
        Built                         b2 = build_scene(W, H);
        Scene&                        s2 = *b2.scene;
        auto                          tex = s2.textures()[0];
        const int32_t                 tid = direct->createTexture(tex->data().data(), 64, 32, 3);
        direct->buildMaterials(*s2.compileMaterials());
        direct->buildGeometryEntities(*s2.compileEntity());
        direct->buildBLAS(s2.collectTriangleMeshForBLASBuilding()->meshes);
        auto t = *s2.collectTriangleMeshForTLASBuilding();
        direct->buildTLAS(t.meshOfInstance, t.instanceIDs, t.transforms);
        auto e = *s2.compileEmitters();
        e.emitters[0].texture = tid;
        direct->buildEmitters(e.emitters, e.domeIndex);
        direct->updateCamera(s2.collectCamera());
        direct->render((uint32_t)frames);
        direct->sync();
    }
    clhip_readback_sums(direct->native(), c.data(), c.size() * sizeof(double));
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i)
    {
        num += (a[i] - c[i]) * (a[i] - c[i]);
        den += c[i] * c[i];
    }
    expect(den > 0 && num <= 1e-20 * den, "engine frames equal one direct render of the same samples");
    const double m0 = film_mean(*integ);

    /* an edit resets the film; the next frame is sample 0 again */
    b.floor->setReflectance({0.05f, 0.05f, 0.05f});
    engine->startRendering();
    engine->endRendering();
    clhip_get_film_size(integ->native(), nullptr, nullptr, &spp);
    expect(spp == 1, "a material edit restarts the accumulation");
    const double m1 = film_mean(*integ);
    expect(m1 < m0, "a darker floor darkens the image");
    b.prop->setTransform(place(-1.0f, 3.0f, 0.0f, 0.8f));
    engine->startRendering();
    engine->endRendering();
    clhip_get_film_size(integ->native(), nullptr, nullptr, &spp);
    expect(spp == 1, "a transform edit restarts the accumulation");

    /* a shape edit rebuilds one BLAS: the octahedron, flattened, renders as
     * a fresh context given the edited scene renders it */
    clhip_stats st{};
    clhip_counters(integ->native(), &st);
    const uint64_t builds0 = st.blas_builds;
    std::vector<float> flat = {1, 0, 0, -1, 0, 0, 0, 0.2f, 0, 0, -0.2f, 0, 0, 0, 1, 0, 0, -1};
    b.prop_mesh->setVertices(flat);
    engine->startRendering();
    engine->endRendering();
    clhip_counters(integ->native(), &st);
    expect(st.blas_builds == builds0 + 1, "a shape edit rebuilds exactly one BLAS");
    std::vector<double> e1(3 * (size_t)W * H), e2(3 * (size_t)W * H);
    clhip_readback_sums(integ->native(), e1.data(), e1.size() * sizeof(double));
    {
        Built b3 = build_scene(W, H);
        b3.prop->setTransform(place(-1.0f, 3.0f, 0.0f, 0.8f));
        b3.prop_mesh->setVertices(flat);
        std::shared_ptr<HipWavefrontPathTracingIntegrator> fresh =
            HipWavefrontPathTracingIntegrator::createIntegrator(IntegratorType::WavefrontPathTracing, W, H);
        fresh->setParams(clhip_params{5u, 0u, 0u, 0u});
        auto e3 = RenderEngine::createRenderEngine(fresh, b3.scene);
        b3.floor->setReflectance({0.05f, 0.05f, 0.05f});
        e3->startRendering();
        e3->endRendering();
        clhip_readback_sums(fresh->native(), e2.data(), e2.size() * sizeof(double));
    }
    double dn = 0, dd = 0;
    for (size_t i = 0; i < e1.size(); ++i)
    {
        dn += (e1[i] - e2[i]) * (e1[i] - e2[i]);
        dd += e2[i] * e2[i];
    }
    expect(dd > 0 && dn <= 1e-20 * dd, "after one BLAS rebuild the image equals a full build's");
    std::printf("frames %d mean %.6f after-edit %.6f\n", frames, m0, m1);
    return g_fail ? 1 : 0;
}
} // namespace

int main(int argc, char** argv)
{
    try
    {
        if (argc > 1 && std::strcmp(argv[1], "--selftest") == 0)
            return selftest();
        return render(argc > 1 ? std::atoi(argv[1]) : 8);
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
}
