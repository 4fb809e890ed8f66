/*
 * hip_scene.h -- C++ host mirror of Colvillea's scene manager and render
 * engine over the HIP integrator (hip_integrator.h):
 *
 *   SceneEditAction  include/colvillea/librender/scene.h:25-92 (same flags)
 *   Scene            scene.h:105-392, src/core/librender/scene.cpp:186-345:
 *                    create* adds nodes and records edit actions; collect* /
 *                    compile* return the integrator inputs only when the
 *                    corresponding actions are pending
 *   RenderEngine     src/core/librender/renderengine.cpp:27-121:
 *                    startRendering() = compile materials, entities, accel
 *                    structs, emitters (each only if dirty), update the
 *                    camera, reset the actions, render one progressive spp
 *
 * Differences, by design:
 *   - entities carry an optional 3x4 object-to-world transform (the
 *     reference has none, its defect D14); editing it marks EntityEdited and
 *     rebuilds the TLAS only;
 *   - BLAS validity is tracked per mesh, as the reference's TriMeshBLAS
 *     does (asdataset.h:16-124): a shape edit rebuilds that mesh's tree
 *     only (clhip_build_blas_subset), then the TLAS;
 *   - textures are uploaded by the engine when TexturesAdded is pending, and
 *     the HDRI dome refers to its texture by the integrator's texture id.
 */
#ifndef COLVILLEA_AMD_HIP_SCENE_H
#define COLVILLEA_AMD_HIP_SCENE_H

#include "hip_integrator.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <memory>
#include <optional>
#include <utility>
#include <vector>

namespace clhip
{

/* Edit actions (scene.h:25-92): bit flags accumulated until the engine has
 * pushed the scene to the integrator. */
class SceneEditAction
{
public:
    enum class EditActionType : uint32_t
    {
        None            = 0,
        ShapeEdited     = 1u << 0,
        ShapeAdded      = 1u << 1,
        ShapeRemoved    = 1u << 2,
        MaterialEdited  = 1u << 3,
        MaterialAdded   = 1u << 4,
        MaterialRemoved = 1u << 5,
        EntityEdited    = 1u << 6,
        EntityAdded     = 1u << 7,
        EntityRemoved   = 1u << 8,
        EmitterEdited   = 1u << 9,
        EmitterAdded    = 1u << 10,
        EmitterRemoved  = 1u << 11,
        TexturesEdited  = 1u << 12,
        TexturesAdded   = 1u << 13,
        TexturesRemoved = 1u << 14,

        AnyShapeModification    = ShapeEdited | ShapeAdded | ShapeRemoved,
        AnyMaterialModification = MaterialEdited | MaterialAdded | MaterialRemoved,
        AnyEntityModification   = EntityEdited | EntityAdded | EntityRemoved,
        AnyEmitterModification  = EmitterEdited | EmitterAdded | EmitterRemoved,
        AnyTextureModification  = TexturesEdited | TexturesAdded | TexturesRemoved
    };

    void reset() noexcept { m_bits = 0; }
    void addAction(EditActionType a) noexcept { m_bits |= (uint32_t)a; }
    bool hasAction(EditActionType a) const noexcept { return (m_bits & (uint32_t)a) != 0; }
    uint32_t bits() const noexcept { return m_bits; }

private:
    uint32_t m_bits = 0;
};

using Action = SceneEditAction::EditActionType;
using Transform = std::array<float, 12>; /* row-major 3x4, object -> world */

inline Transform identityTransform() { return {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0}; }

class Scene;

/* Scene nodes: every node has a scene-unique id (nodebase/node.h:11-78). */
class Node
{
public:
    explicit Node(uint64_t id) : m_id(id) {}
    uint64_t getID() const noexcept { return m_id; }

private:
    uint64_t m_id;
};

/* core::ImageTexture2D: an RGB or RGBA float32 image. */
class Texture : public Node
{
public:
    Texture(uint64_t id, std::vector<float> data, uint32_t w, uint32_t h, uint32_t channels)
        : Node(id), m_data(std::move(data)), m_w(w), m_h(h), m_channels(channels)
    {
    }
    std::pair<uint32_t, uint32_t> getTextureResolution() const noexcept { return {m_w, m_h}; }
    const std::vector<float>&     data() const noexcept { return m_data; }
    uint32_t                      channels() const noexcept { return m_channels; }
    int32_t                       deviceId() const noexcept { return m_deviceId; }

private:
    friend class RenderEngine;
    std::vector<float> m_data;
    uint32_t           m_w, m_h, m_channels;
    int32_t            m_deviceId = -1; /* integrator texture id once uploaded */
};

/* core::Material (nodebase/material.h) compiled to clhip_material: the
 * constants, and the textures (reflectance, roughness, normal map) by the
 * integrator ids they receive when the engine uploads them. */
class Material : public Node
{
public:
    Material(uint64_t id, const clhip_material& m, Scene* scene, std::shared_ptr<Texture> reflectance = nullptr,
             std::shared_ptr<Texture> roughness = nullptr)
        : Node(id), m_mtl(m), m_scene(scene), m_reflectanceTex(std::move(reflectance)),
          m_roughnessTex(std::move(roughness))
    {
    }
    clhip_material compile() const
    {
        clhip_material m  = m_mtl;
        m.reflectance_tex = texId(m_reflectanceTex);
        m.roughness_tex   = texId(m_roughnessTex);
        m.normal_tex      = texId(m_normalmapTex);
        return m;
    }
    /* MaterialEdited */
    void setReflectance(const vec3f& r);
    void setRoughness(float r);
    /* Material::setNormalmap / getNormalmap (nodebase/material.h:28-36) */
    void                            setNormalmap(std::shared_ptr<Texture> normalmap);
    const std::shared_ptr<Texture>& getNormalmap() const noexcept { return m_normalmapTex; }

private:
    static uint32_t texId(const std::shared_ptr<Texture>& t)
    {
        if (!t)
            return 0u;
        if (t->deviceId() < 0)
            throw Error(CLHIP_E_STATE, "material texture not uploaded: compile through the RenderEngine");
        return (uint32_t)t->deviceId() + 1u;
    }
    clhip_material           m_mtl;
    Scene*                   m_scene;
    std::shared_ptr<Texture> m_reflectanceTex, m_roughnessTex, m_normalmapTex;
};

/* core::Emitter: directional sun or HDRI dome. */
class Emitter : public Node
{
public:
    Emitter(uint64_t id, const clhip_emitter& e, std::shared_ptr<Texture> sky)
        : Node(id), m_emitter(e), m_sky(std::move(sky))
    {
    }
    uint32_t                        getEmitterType() const noexcept { return m_emitter.type; }
    const std::shared_ptr<Texture>& getEnvmap() const noexcept { return m_sky; }
    clhip_emitter                   compile() const
    {
        clhip_emitter e = m_emitter;
        if (m_sky)
            e.texture = m_sky->deviceId();
        return e;
    }

private:
    friend class RenderEngine;
    clhip_emitter            m_emitter;
    std::shared_ptr<Texture> m_sky;
};

/* core::TriangleMesh (nodes/shapes/trianglemesh.h): owns its arrays. */
class TriangleMesh : public Node
{
public:
    TriangleMesh(uint64_t id, Scene* scene, std::vector<float> verts, std::vector<uint32_t> tris,
                 std::vector<float> normals, std::vector<float> tangents, std::vector<float> uvs)
        : Node(id), m_scene(scene), m_verts(std::move(verts)), m_tris(std::move(tris)),
          m_normals(std::move(normals)), m_tangents(std::move(tangents)), m_uvs(std::move(uvs))
    {
    }
    TriangleMeshView view() const
    {
        TriangleMeshView v;
        v.vertices     = m_verts.data();
        v.indices      = m_tris.data();
        v.normals      = m_normals.empty() ? nullptr : m_normals.data();
        v.tangents     = m_tangents.empty() ? nullptr : m_tangents.data();
        v.uvs          = m_uvs.empty() ? nullptr : m_uvs.data();
        v.numVertices  = (uint32_t)(m_verts.size() / 3);
        v.numTriangles = (uint32_t)(m_tris.size() / 3);
        return v;
    }
    /* ShapeEdited: new vertex positions (same count) */
    void setVertices(std::vector<float> verts);
    /* TriMeshBLAS::needRebuildBLAS (asdataset.h:16-124) */
    bool needRebuildBLAS() const noexcept { return !m_blasValid; }

private:
    friend class Scene;
    Scene*                m_scene;
    bool                  m_blasValid = false;
    std::vector<float>    m_verts;
    std::vector<uint32_t> m_tris;
    std::vector<float>    m_normals, m_tangents, m_uvs;
};

/* core::Entity (librender/entity.h): a mesh instance with a material. */
class Entity : public Node
{
public:
    Entity(uint64_t id, Scene* scene, std::shared_ptr<TriangleMesh> mesh, std::shared_ptr<Material> mtl,
           const Transform& xf)
        : Node(id), m_scene(scene), m_mesh(std::move(mesh)), m_mtl(std::move(mtl)), m_xf(xf)
    {
    }
    const TriangleMesh*              getTrimesh() const noexcept { return m_mesh.get(); }
    const std::shared_ptr<Material>& getMaterial() const noexcept { return m_mtl; }
    const Transform&                 getTransform() const noexcept { return m_xf; }
    /* EntityEdited: material swap or new transform */
    void setMaterial(std::shared_ptr<Material> m);
    void setTransform(const Transform& xf);

private:
    Scene*                        m_scene;
    std::shared_ptr<TriangleMesh> m_mesh;
    std::shared_ptr<Material>     m_mtl;
    Transform                     m_xf;
};

struct CompiledEmitterResult
{
    std::vector<clhip_emitter> emitters;
    int32_t                    domeIndex = -1;
};

struct BLASBuildSource
{
    std::vector<uint32_t>         meshIndex; /* position in the scene's mesh list */
    std::vector<TriangleMeshView> meshes;
};

struct TLASBuildSource
{
    std::vector<uint32_t> meshOfInstance; /* index into the BLAS list (scene mesh order) */
    std::vector<uint32_t> instanceIDs;    /* entity index, as scene.cpp:232-236 */
    std::vector<float>    transforms;     /* 12 per instance */
};

class Scene
{
public:
    static std::unique_ptr<Scene> createScene() { return std::unique_ptr<Scene>(new Scene()); }

    /* createEmitter (scene.h:130-150) */
    std::shared_ptr<Emitter> createEmitter(const vec3f& colorMulIntensity, const vec3f& sunDirection,
                                           float sunAngularRadius)
    {
        clhip_emitter e{};
        e.type           = CLHIP_EMITTER_DIRECTIONAL;
        e.angular_radius = sunAngularRadius;
        e.texture        = -1;
        for (int k = 0; k < 3; ++k)
        {
            e.intensity[k] = colorMulIntensity[k];
            e.direction[k] = sunDirection[k];
        }
        return add(m_emitters, std::make_shared<Emitter>(nextId(), e, nullptr), Action::EmitterAdded);
    }
    std::shared_ptr<Emitter> createEmitter(const std::shared_ptr<Texture>& sky)
    {
        if (!sky)
            throw Error(CLHIP_E_INVALID, "createEmitter: an HDRI dome needs a texture");
        clhip_emitter e{};
        e.type    = CLHIP_EMITTER_HDRIDOME;
        e.texture = -1;
        return add(m_emitters, std::make_shared<Emitter>(nextId(), e, sky), Action::EmitterAdded);
    }
    /* createTexture (scene.h:152-160): RGB or RGBA float32 */
    std::shared_ptr<Texture> createTexture(std::vector<float> data, uint32_t w, uint32_t h, uint32_t channels)
    {
        if ((channels != 3 && channels != 4) || data.size() != (size_t)w * h * channels)
            throw Error(CLHIP_E_INVALID, "createTexture: data size does not match w * h * channels");
        return add(m_textures, std::make_shared<Texture>(nextId(), std::move(data), w, h, channels),
                   Action::TexturesAdded);
    }
    /* createMaterial / createMetalMaterial / createGlassMaterial (scene.h:162-200) */
    std::shared_ptr<Material> createDiffuseMaterial(const vec3f& reflectance)
    {
        clhip_material m{};
        m.type = CLHIP_MTL_DIFFUSE;
        for (int k = 0; k < 3; ++k)
            m.reflectance[k] = reflectance[k];
        return add(m_materials, std::make_shared<Material>(nextId(), m, this), Action::MaterialAdded);
    }
    std::shared_ptr<Material> createMetalMaterial(const vec3f& specularReflectance, float roughness, const vec3f& eta,
                                                  const vec3f& k)
    {
        clhip_material m{};
        m.type      = CLHIP_MTL_METAL;
        m.roughness = roughness;
        for (int c = 0; c < 3; ++c)
        {
            m.reflectance[c] = specularReflectance[c];
            m.eta[c]         = eta[c];
            m.k[c]           = k[c];
        }
        return add(m_materials, std::make_shared<Material>(nextId(), m, this), Action::MaterialAdded);
    }
    /* createMaterial(Diffuse, reflectanceTex) (scene.h:163) */
    std::shared_ptr<Material> createDiffuseMaterial(const std::shared_ptr<Texture>& reflectanceTex)
    {
        clhip_material m{};
        m.type = CLHIP_MTL_DIFFUSE;
        return add(m_materials, std::make_shared<Material>(nextId(), m, this, reflectanceTex),
                   Action::MaterialAdded);
    }
    /* createMetalMaterial with a roughness texture (scene.h:166) */
    std::shared_ptr<Material> createMetalMaterial(const vec3f& specularReflectance,
                                                  const std::shared_ptr<Texture>& roughness, const vec3f& eta,
                                                  const vec3f& k)
    {
        clhip_material m{};
        m.type = CLHIP_MTL_METAL;
        for (int c = 0; c < 3; ++c)
        {
            m.reflectance[c] = specularReflectance[c];
            m.eta[c]         = eta[c];
            m.k[c]           = k[c];
        }
        return add(m_materials, std::make_shared<Material>(nextId(), m, this, nullptr, roughness),
                   Action::MaterialAdded);
    }
    std::shared_ptr<Material> createGlassMaterial(float roughness, float interiorIOR)
    {
        clhip_material m{};
        m.type         = CLHIP_MTL_GLASS;
        m.roughness    = roughness;
        m.interior_ior = interiorIOR;
        return add(m_materials, std::make_shared<Material>(nextId(), m, this), Action::MaterialAdded);
    }
    /* createTriangleMesh (scene.h:202-226): xyz per vertex, 3 indices per triangle */
    std::shared_ptr<TriangleMesh> createTriangleMesh(std::vector<float> verts, std::vector<uint32_t> tris,
                                                     std::vector<float> normals = {}, std::vector<float> tangents = {},
                                                     std::vector<float> uvs = {})
    {
        if (verts.size() % 3 || tris.size() % 3)
            throw Error(CLHIP_E_INVALID, "createTriangleMesh: arrays must hold whole vertices / triangles");
        return add(m_trimeshes,
                   std::make_shared<TriangleMesh>(nextId(), this, std::move(verts), std::move(tris), std::move(normals),
                                                  std::move(tangents), std::move(uvs)),
                   Action::ShapeAdded);
    }
    /* createEntity (scene.h:228-236); the transform is this build's extension */
    std::shared_ptr<Entity> createEntity(const std::shared_ptr<TriangleMesh>& shape,
                                         const std::shared_ptr<Material>& material,
                                         const Transform&                 xf = identityTransform())
    {
        if (!shape || !material)
            throw Error(CLHIP_E_INVALID, "createEntity: shape and material are required");
        return add(m_entities, std::make_shared<Entity>(nextId(), this, shape, material, xf), Action::EntityAdded);
    }
    void setCamera(const Camera& c) { m_camera = c; }
    const Camera& collectCamera() const noexcept { return m_camera; }

    /* collectTriangleMeshForBLASBuilding (scene.cpp:186-210): when a shape
     * was added or edited, the meshes whose BLAS needs a rebuild */
    std::optional<BLASBuildSource> collectTriangleMeshForBLASBuilding() const
    {
        if (!m_actions.hasAction(Action::ShapeEdited) && !m_actions.hasAction(Action::ShapeAdded))
            return std::nullopt;
        BLASBuildSource b;
        for (size_t i = 0; i < m_trimeshes.size(); ++i)
            if (m_trimeshes[i]->needRebuildBLAS())
            {
                b.meshIndex.push_back((uint32_t)i);
                b.meshes.push_back(m_trimeshes[i]->view());
            }
        return b;
    }
    /* the BLASes of these meshes are built (TriMeshBLAS state) */
    void markBLASBuilt(const std::vector<uint32_t>& meshIndex)
    {
        for (uint32_t i : meshIndex)
            m_trimeshes.at(i)->m_blasValid = true;
    }
    /* collectTriangleMeshForTLASBuilding (scene.cpp:212-242): one instance
     * per entity, instance id = entity index; also on entity edits (new
     * transforms) and after a BLAS rebuild */
    std::optional<TLASBuildSource> collectTriangleMeshForTLASBuilding() const
    {
        if (!m_actions.hasAction(Action::AnyShapeModification) && !m_actions.hasAction(Action::AnyEntityModification))
            return std::nullopt;
        TLASBuildSource t;
        for (size_t i = 0; i < m_entities.size(); ++i)
        {
            const TriangleMesh* mesh = m_entities[i]->getTrimesh();
            const auto it = std::find_if(m_trimeshes.begin(), m_trimeshes.end(),
                                         [mesh](const auto& p) { return p.get() == mesh; });
            t.meshOfInstance.push_back((uint32_t)(it - m_trimeshes.begin()));
            t.instanceIDs.push_back((uint32_t)i);
            const Transform& xf = m_entities[i]->getTransform();
            t.transforms.insert(t.transforms.end(), xf.begin(), xf.end());
        }
        return t;
    }
    /* compileEntity (scene.cpp:244-280): material index per entity */
    std::optional<std::vector<uint32_t>> compileEntity() const
    {
        if (!m_actions.hasAction(Action::AnyEntityModification) &&
            !m_actions.hasAction(Action::AnyMaterialModification))
            return std::nullopt;
        std::vector<uint32_t> idx;
        for (const auto& e : m_entities)
        {
            const auto it = std::find_if(m_materials.begin(), m_materials.end(), [&e](const auto& m) {
                return m->getID() == e->getMaterial()->getID();
            });
            if (it == m_materials.end())
                throw Error(CLHIP_E_STATE, "entity refers to a material that is not in the scene");
            idx.push_back((uint32_t)(it - m_materials.begin()));
        }
        return idx;
    }
    /* compileMaterials (scene.cpp:282-300) */
    std::optional<std::vector<clhip_material>> compileMaterials() const
    {
        if (!m_actions.hasAction(Action::AnyMaterialModification))
            return std::nullopt;
        std::vector<clhip_material> v;
        for (const auto& m : m_materials)
            v.push_back(m->compile());
        return v;
    }
    /* compileEmitters (scene.cpp:302-345) */
    std::optional<CompiledEmitterResult> compileEmitters() const
    {
        if (!m_actions.hasAction(Action::AnyEmitterModification) && !m_actions.hasAction(Action::AnyTextureModification))
            return std::nullopt;
        CompiledEmitterResult r;
        for (size_t i = 0; i < m_emitters.size(); ++i)
        {
            r.emitters.push_back(m_emitters[i]->compile());
            if (m_emitters[i]->getEmitterType() == CLHIP_EMITTER_HDRIDOME)
                r.domeIndex = (int32_t)i;
        }
        return r;
    }

    const SceneEditAction& editActions() const noexcept { return m_actions; }
    void                   resetSceneEditActions() noexcept { m_actions.reset(); }
    void                   markEdited(Action a) noexcept { m_actions.addAction(a); }
    const std::vector<std::shared_ptr<Texture>>& textures() const noexcept { return m_textures; }

private:
    Scene() = default;
    uint64_t nextId() noexcept { return ++m_lastId; }
    template <typename T>
    std::shared_ptr<T> add(std::vector<std::shared_ptr<T>>& list, std::shared_ptr<T> node, Action a)
    {
        list.push_back(node);
        m_actions.addAction(a);
        return node;
    }

    uint64_t                                   m_lastId = 0;
    SceneEditAction                            m_actions;
    std::vector<std::shared_ptr<TriangleMesh>> m_trimeshes;
    std::vector<std::shared_ptr<Material>>     m_materials;
    std::vector<std::shared_ptr<Texture>>      m_textures;
    std::vector<std::shared_ptr<Emitter>>      m_emitters;
    std::vector<std::shared_ptr<Entity>>       m_entities;
    Camera                                     m_camera{};
};

inline void Material::setReflectance(const vec3f& r)
{
    for (int k = 0; k < 3; ++k)
        m_mtl.reflectance[k] = r[k];
    m_scene->markEdited(Action::MaterialEdited);
}
inline void Material::setRoughness(float r)
{
    m_mtl.roughness = r;
    m_scene->markEdited(Action::MaterialEdited);
}
inline void Material::setNormalmap(std::shared_ptr<Texture> normalmap)
{
    m_normalmapTex = std::move(normalmap);
    m_scene->markEdited(Action::MaterialEdited);
}
inline void TriangleMesh::setVertices(std::vector<float> verts)
{
    if (verts.size() != m_verts.size())
        throw Error(CLHIP_E_INVALID, "setVertices: vertex count changed");
    m_verts     = std::move(verts);
    m_blasValid = false;
    m_scene->markEdited(Action::ShapeEdited);
}
inline void Entity::setMaterial(std::shared_ptr<Material> m)
{
    m_mtl = std::move(m);
    m_scene->markEdited(Action::EntityEdited);
}
inline void Entity::setTransform(const Transform& xf)
{
    m_xf = xf;
    m_scene->markEdited(Action::EntityEdited);
}

/* RenderEngine (renderengine.cpp:27-121) */
class RenderEngine
{
public:
    static std::unique_ptr<RenderEngine> createRenderEngine(std::shared_ptr<HipWavefrontPathTracingIntegrator> integ,
                                                            std::shared_ptr<Scene>                             scene)
    {
        return std::unique_ptr<RenderEngine>(new RenderEngine(std::move(integ), std::move(scene)));
    }

    /* startRendering (renderengine.cpp:27-51) */
    void startRendering()
    {
        compileTextures();
        compileMaterials();
        compileEntities();
        compileAccelStructs();
        compileEmitters();
        m_integ->updateCamera(m_scene->collectCamera());
        m_scene->resetSceneEditActions();
        m_integ->render(1);
    }
    void endRendering() { m_integ->sync(); }
    void resize(uint32_t w, uint32_t h)
    {
        m_integ->resize(w, h);
        m_integ->resetIterationIndex();
    }
    std::pair<uint32_t, uint32_t>           getFilmSize() const { return m_integ->getFilmSize(); }
    std::unique_ptr<std::array<float, 4>[]> readbackFramebuffer() { return m_integ->readbackFramebuffer(); }
    HipWavefrontPathTracingIntegrator&      integrator() { return *m_integ; }

private:
    RenderEngine(std::shared_ptr<HipWavefrontPathTracingIntegrator> integ, std::shared_ptr<Scene> scene)
        : m_integ(std::move(integ)), m_scene(std::move(scene))
    {
    }
    void compileTextures()
    {
        if (!m_scene->editActions().hasAction(Action::TexturesAdded))
            return;
        for (const auto& t : m_scene->textures())
            if (t->m_deviceId < 0)
            {
                const auto res = t->getTextureResolution();
                t->m_deviceId  = m_integ->createTexture(t->data().data(), res.first, res.second, t->channels());
            }
    }
    void compileMaterials()
    {
        if (auto m = m_scene->compileMaterials())
        {
            m_integ->buildMaterials(*m);
            m_integ->resetIterationIndex();
        }
    }
    void compileEntities()
    {
        if (auto e = m_scene->compileEntity())
        {
            m_integ->buildGeometryEntities(*e);
            m_integ->resetIterationIndex();
        }
    }
    void compileAccelStructs()
    {
        if (auto b = m_scene->collectTriangleMeshForBLASBuilding())
        {
            if (!b->meshes.empty())
                m_integ->buildBLAS(b->meshIndex, b->meshes);
            m_scene->markBLASBuilt(b->meshIndex);
            m_integ->resetIterationIndex();
        }
        if (auto t = m_scene->collectTriangleMeshForTLASBuilding())
        {
            m_integ->buildTLAS(t->meshOfInstance, t->instanceIDs, t->transforms);
            m_integ->resetIterationIndex();
        }
    }
    void compileEmitters()
    {
        if (auto r = m_scene->compileEmitters())
        {
            m_integ->buildEmitters(r->emitters, r->domeIndex);
            m_integ->resetIterationIndex();
        }
    }

    std::shared_ptr<HipWavefrontPathTracingIntegrator> m_integ;
    std::shared_ptr<Scene>                             m_scene;
};

} // namespace clhip

#endif
