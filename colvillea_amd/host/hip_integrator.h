/*
 * hip_integrator.h -- C++ host class over the C ABI (include/clhip.h) with
 * the method set of core::Integrator (include/colvillea/librender/
 * integrator.h:34-103), for C++ hosts such as Colvillea's RenderEngine.
 *
 * Header-only; link with libclpt_hip.so.  Errors throw clhip::Error (the
 * reference aborts the process on device errors, cudacommon.h:13-26).
 * GL interop (register/unregister/mapFramebuffer) has no MI355X counterpart
 * and throws; readbackFramebuffer() returns the linear HDR film.
 */
#ifndef COLVILLEA_AMD_HIP_INTEGRATOR_H
#define COLVILLEA_AMD_HIP_INTEGRATOR_H

#include "../../include/clhip.h"

#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace clhip
{
class Error : public std::runtime_error
{
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), m_code(code) {}
    int code() const noexcept { return m_code; }

private:
    int m_code;
};

/* IntegratorType, integrator.h:22-32 */
enum class IntegratorType : uint32_t
{
    None                    = CLHIP_INTEGRATOR_NONE,
    WavefrontDirectLighting = CLHIP_INTEGRATOR_DIRECT_LIGHTING,
    WavefrontPathTracing    = CLHIP_INTEGRATOR_PATH_TRACING
};

using vec3f = std::array<float, 3>;

/* TriangleMesh view: the caller keeps the arrays alive during buildBLAS. */
struct TriangleMeshView
{
    const float*    vertices = nullptr;
    const uint32_t* indices  = nullptr;
    const float*    normals  = nullptr;
    const float*    tangents = nullptr;
    const float*    uvs      = nullptr;
    uint32_t        numVertices = 0, numTriangles = 0;
};

/* core::Camera (librender/nodebase/camera.h:8-27) */
struct Camera
{
    vec3f m_camera_pos, m_camera_d00, m_camera_ddu, m_camera_ddv;
};

class HipWavefrontPathTracingIntegrator
{
public:
    /* Integrator::createIntegrator (integrator.cpp:10-22) */
    static std::unique_ptr<HipWavefrontPathTracingIntegrator> createIntegrator(IntegratorType type, uint32_t width,
                                                                               uint32_t height)
    {
        return std::unique_ptr<HipWavefrontPathTracingIntegrator>(
            new HipWavefrontPathTracingIntegrator(type, width, height));
    }

    ~HipWavefrontPathTracingIntegrator() { clhip_destroy(m_ctx); }
    HipWavefrontPathTracingIntegrator(const HipWavefrontPathTracingIntegrator&)            = delete;
    HipWavefrontPathTracingIntegrator& operator=(const HipWavefrontPathTracingIntegrator&) = delete;

    IntegratorType getIntegratorType() const noexcept { return m_type; }

    /* buildBLAS (integrator.h:47) */
    void buildBLAS(const std::vector<TriangleMeshView>& trimeshes)
    {
        std::vector<clhip_mesh> m(trimeshes.size());
        for (size_t i = 0; i < m.size(); ++i)
            m[i] = clhip_mesh{trimeshes[i].vertices, trimeshes[i].indices, trimeshes[i].normals,
                              trimeshes[i].tangents, trimeshes[i].uvs,     trimeshes[i].numVertices,
                              trimeshes[i].numTriangles};
        check(clhip_build_blas(m_ctx, m.data(), (uint32_t)m.size()));
    }

    /* buildBLAS for the meshes whose BLAS needs a rebuild, as RenderEngine
     * passes them (renderengine.cpp:64-80): trimeshes[k] is mesh number
     * meshIndex[k]; the other meshes keep their trees. */
    void buildBLAS(const std::vector<uint32_t>& meshIndex, const std::vector<TriangleMeshView>& trimeshes)
    {
        if (meshIndex.size() != trimeshes.size())
            throw Error(CLHIP_E_INVALID, "buildBLAS: mismatched array sizes");
        std::vector<clhip_mesh> m(trimeshes.size());
        for (size_t i = 0; i < m.size(); ++i)
            m[i] = clhip_mesh{trimeshes[i].vertices, trimeshes[i].indices, trimeshes[i].normals,
                              trimeshes[i].tangents, trimeshes[i].uvs,     trimeshes[i].numVertices,
                              trimeshes[i].numTriangles};
        check(clhip_build_blas_subset(m_ctx, m.data(), meshIndex.data(), (uint32_t)m.size()));
    }

    /* buildTLAS (integrator.h:49-50): meshOfInstance indexes the BLAS list;
     * transforms (3x4 row-major per instance) may be empty for identity. */
    void buildTLAS(const std::vector<uint32_t>& meshOfInstance, const std::vector<uint32_t>& instanceIDs,
                   const std::vector<float>& transforms = {})
    {
        if (instanceIDs.size() != meshOfInstance.size() ||
            (!transforms.empty() && transforms.size() != 12 * meshOfInstance.size()))
            throw Error(CLHIP_E_INVALID, "buildTLAS: mismatched array sizes");
        check(clhip_build_tlas(m_ctx, meshOfInstance.data(), instanceIDs.data(),
                               transforms.empty() ? nullptr : transforms.data(), (uint32_t)meshOfInstance.size()));
    }

    /* buildMaterials / buildGeometryEntities / buildEmitters (integrator.h:52-58) */
    void buildMaterials(const std::vector<clhip_material>& materials)
    {
        check(clhip_build_materials(m_ctx, materials.data(), (uint32_t)materials.size()));
    }
    void buildGeometryEntities(const std::vector<uint32_t>& materialIndexOfEntity)
    {
        check(clhip_build_entities(m_ctx, materialIndexOfEntity.data(), (uint32_t)materialIndexOfEntity.size()));
    }
    int32_t createTexture(const float* data, uint32_t w, uint32_t h, uint32_t channels)
    {
        int32_t id = -1;
        check(clhip_create_texture(m_ctx, data, w, h, channels, &id));
        return id;
    }
    void buildEmitters(const std::vector<clhip_emitter>& emitters, int32_t domeIndex)
    {
        check(clhip_build_emitters(m_ctx, emitters.data(), (uint32_t)emitters.size(), domeIndex));
    }

    /* updateCamera (integrator.h:80) */
    void updateCamera(const Camera& c)
    {
        check(clhip_set_camera(m_ctx, c.m_camera_pos.data(), c.m_camera_d00.data(), c.m_camera_ddu.data(),
                               c.m_camera_ddv.data()));
    }
    void setParams(const clhip_params& p) { check(clhip_set_params(m_ctx, &p)); }

    /* render (integrator.h:61): one progressive sample per pixel */
    void render(uint32_t spp = 1) { check(clhip_render(m_ctx, spp)); }
    void sync() { check(clhip_sync(m_ctx)); }

    /* resize / resetIterationIndex / getFilmSize (integrator.h:64,75,77) */
    void resize(uint32_t w, uint32_t h) { check(clhip_resize(m_ctx, w, h)); }
    void resetIterationIndex() { check(clhip_reset(m_ctx)); }
    std::pair<uint32_t, uint32_t> getFilmSize() const
    {
        uint32_t w = 0, h = 0;
        check(clhip_get_film_size(m_ctx, &w, &h, nullptr));
        return {w, h};
    }

    /* readbackFramebuffer (integrator.h:73): RGBA float32, linear HDR */
    std::unique_ptr<std::array<float, 4>[]> readbackFramebuffer()
    {
        const auto sz = getFilmSize();
        const size_t n = (size_t)sz.first * sz.second;
        std::unique_ptr<std::array<float, 4>[]> out(new std::array<float, 4>[n]);
        check(clhip_readback(m_ctx, out[0].data(), n * sizeof(float) * 4));
        return out;
    }

    /* GL interop (integrator.h:66-71) */
    void unregisterFramebuffer() { throw Error(CLHIP_E_UNSUPPORTED, "GL interop is not available"); }
    void registerFramebuffer(unsigned int) { throw Error(CLHIP_E_UNSUPPORTED, "GL interop is not available"); }
    void mapFramebuffer() { throw Error(CLHIP_E_UNSUPPORTED, "GL interop is not available"); }

    clhip_stats counters()
    {
        clhip_stats s{};
        check(clhip_counters(m_ctx, &s));
        return s;
    }
    clhip_ctx* native() const noexcept { return m_ctx; }

private:
    HipWavefrontPathTracingIntegrator(IntegratorType type, uint32_t width, uint32_t height) : m_type(type)
    {
        check(clhip_create(width, height, (uint32_t)type, &m_ctx));
    }
    void check(int rc) const
    {
        if (rc != CLHIP_OK)
            throw Error(rc, clhip_last_error(m_ctx));
    }

    clhip_ctx*     m_ctx = nullptr;
    IntegratorType m_type;
};
} // namespace clhip

#endif
