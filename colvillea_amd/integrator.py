"""Host-side mirror of Colvillea's rendering interface over the C ABI.

`Integrator` keeps the method names, argument meaning and call order of
core::Integrator (include/colvillea/librender/integrator.h:34-103) and of
Integrator::createIntegrator (src/core/librender/integrator.cpp:10-22), so a
RenderEngine-style driver (renderengine.cpp:27-121) can call it unchanged.
The value types mirror the reference's kernel records:
TriangleMesh (nodes/shapes/trianglemesh.h), DiffuseMtl / MetalMtl / GlassMtl
(libkernel/materials/*.h), DirectionalEmitter / HDRIDome
(libkernel/emitters/*.h), core::Camera (librender/nodebase/camera.h:8-27).

Errors raise ClhipError (the reference aborts the process instead,
src/core/devices/cudacommon.h:13-26).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from enum import IntEnum
from typing import Optional, Sequence

import numpy as np

from . import _capi as capi
from ._capi import check


class IntegratorType(IntEnum):
    """integrator.h:22-32"""
    None_ = 0
    WavefrontDirectLighting = 1
    WavefrontPathTracing = 2


@dataclass
class TriangleMesh:
    vertices: np.ndarray                 # (n, 3) float32
    triangles: np.ndarray                # (m, 3) uint32
    normals: Optional[np.ndarray] = None
    tangents: Optional[np.ndarray] = None
    uvs: Optional[np.ndarray] = None

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, np.float32).reshape(-1, 3)
        self.triangles = np.ascontiguousarray(self.triangles, np.uint32).reshape(-1, 3)
        for name in ("normals", "tangents"):
            v = getattr(self, name)
            if v is not None:
                setattr(self, name, np.ascontiguousarray(v, np.float32).reshape(-1, 3))
        if self.uvs is not None:
            self.uvs = np.ascontiguousarray(self.uvs, np.float32).reshape(-1, 2)


class _Normalmapped:
    """Material::setNormalmap / getNormalmap (nodebase/material.h:28-36):
    a tangent-space normal map, texel (r, g, b) -> normalize(2 rgb - 1)."""

    def setNormalmap(self, normalmap: Optional["ImageTexture2D"]):
        self.normalmap = normalmap

    def getNormalmap(self) -> Optional["ImageTexture2D"]:
        return self.normalmap


@dataclass
class DiffuseMtl(_Normalmapped):
    """reflectance: rgb, or an ImageTexture2D (Scene::createMaterial with a
    reflectance texture, scene.h:161-163)."""
    reflectance: object
    normalmap: Optional["ImageTexture2D"] = None


@dataclass
class MetalMtl(_Normalmapped):
    """specularReflectance: rgb or ImageTexture2D; roughness: float or an
    ImageTexture2D whose first channel is the roughness (metal.h:36-55)."""
    specularReflectance: object
    roughness: object
    eta: Sequence[float]
    k: Sequence[float]
    normalmap: Optional["ImageTexture2D"] = None


@dataclass
class GlassMtl(_Normalmapped):
    roughness: float
    interiorIOR: float
    normalmap: Optional["ImageTexture2D"] = None


@dataclass
class Entity:
    """kernel::Entity (libkernel/base/entity.h:24-31)"""
    materialIndex: int


@dataclass
class ImageTexture2D:
    """RGB(A) image, (h, w, c); row 0 is v = 0.  float32 data is used as
    is (RGBA32F); uint8 data is read as c / 255 (RGBAU8), and isSRGB marks
    an 8-bit sRGB image whose colour channels are decoded to linear
    (imagetex2d.cpp:13-15, 50-66)."""
    image: np.ndarray
    isSRGB: bool = False


@dataclass
class DirectionalEmitter:
    colorMulIntensity: Sequence[float]
    sunDirection: Sequence[float]
    sunAngularRadius: float


@dataclass
class HDRIDome:
    texture: ImageTexture2D


@dataclass
class Camera:
    """core::Camera: position and the pixel-plane basis d00, ddu, ddv."""
    m_camera_pos: Sequence[float]
    m_camera_d00: Sequence[float]
    m_camera_ddu: Sequence[float]
    m_camera_ddv: Sequence[float]


@dataclass
class View:
    """One view of a batched render: film size, samples, depth, RNG lane
    and camera."""
    width: int
    height: int
    spp: int
    max_depth: int
    camera: Camera
    rr_depth: int = 0
    lane: int = 0


def _f3(v) -> C.Array:
    return (C.c_float * 3)(*[float(x) for x in v])


class Integrator:
    """Wavefront path-tracing integrator on one MI355X (gfx950)."""

    @staticmethod
    def createIntegrator(type: IntegratorType, width: int, height: int) -> "Integrator":
        return Integrator(type, width, height)

    def __init__(self, type: IntegratorType, width: int, height: int):
        self._lib = capi.lib()
        ctx = C.c_void_p()
        check(self._lib.clhip_create(int(width), int(height), int(type), C.byref(ctx)))
        self._ctx = ctx
        self._type = IntegratorType(type)
        self._meshes: list = []
        self._keep: list = []  # host arrays referenced by the last BLAS build
        self._views: list = []
        self.spp_per_render = 1

    # -- lifetime
    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.clhip_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def getIntegratorType(self) -> IntegratorType:
        return self._type

    def _check(self, rc):
        check(rc, self._ctx)

    # -- build* (integrator.h:47-58)
    def setBVHWidth(self, width: int):
        """4 or 8; call before buildBLAS (no reference counterpart)."""
        self._check(self._lib.clhip_set_bvh_width(self._ctx, int(width)))

    @staticmethod
    def _mesh_array(trimeshes):
        arr = (capi.clhip_mesh * max(1, len(trimeshes)))()
        for i, m in enumerate(trimeshes):
            arr[i].vertices = capi.fptr(m.vertices)
            arr[i].indices = capi.uptr(m.triangles)
            arr[i].normals = capi.fptr(m.normals)
            arr[i].tangents = capi.fptr(m.tangents)
            arr[i].uvs = capi.fptr(m.uvs)
            arr[i].num_vertices = m.vertices.shape[0]
            arr[i].num_triangles = m.triangles.shape[0]
        return arr

    def buildBLAS(self, trimeshes: Sequence[TriangleMesh]):
        """Every mesh's tree (the mesh list of the scene, in order)."""
        self._check(self._lib.clhip_build_blas(self._ctx, self._mesh_array(trimeshes), len(trimeshes)))
        self._meshes = list(trimeshes)

    def rebuildBLAS(self, meshIndex: Sequence[int], trimeshes: Sequence[TriangleMesh]):
        """The trees of the meshes whose BLAS needs a rebuild, as RenderEngine
        passes them (scene.cpp:186-210): trimeshes[k] replaces mesh number
        meshIndex[k] (or appends it, one past the last); the other meshes keep
        their trees.  Build the TLAS again afterwards."""
        idx = np.ascontiguousarray(meshIndex, np.uint32)
        if len(idx) != len(trimeshes):
            raise ValueError("rebuildBLAS: one index per mesh")
        self._check(self._lib.clhip_build_blas_subset(self._ctx, self._mesh_array(trimeshes), capi.uptr(idx),
                                                      len(idx)))
        meshes = list(getattr(self, "_meshes", []))
        for i, m in zip(idx.tolist(), trimeshes):
            if i == len(meshes):
                meshes.append(m)
            else:
                meshes[i] = m
        self._meshes = meshes

    def buildTLAS(self, trimeshes: Sequence[TriangleMesh], instanceIDs: Sequence[int],
                  transforms: Optional[np.ndarray] = None):
        """One instance per (mesh, entity id); `transforms` (n, 3, 4)
        object-to-world matrices extend the reference's identity instances."""
        index = {id(m): i for i, m in enumerate(self._meshes)}
        try:
            mesh_of = np.array([index[id(m)] for m in trimeshes], np.uint32)
        except KeyError as e:
            raise ValueError("buildTLAS: mesh not passed to buildBLAS") from e
        ids = np.ascontiguousarray(instanceIDs, np.uint32)
        xf = None if transforms is None else np.ascontiguousarray(transforms, np.float32).reshape(-1, 12)
        self._check(self._lib.clhip_build_tlas(self._ctx, capi.uptr(mesh_of), capi.uptr(ids),
                                               capi.fptr(xf), len(mesh_of)))

    def buildMaterials(self, materials):
        """Materials in order; their textures (reflectance, roughness,
        normal map) are created on the device once per texture object."""
        arr = (capi.clhip_material * max(1, len(materials)))()
        made = {}

        def tex(t) -> int:  # texture id + 1 (0 = none)
            if t is None:
                return 0
            if id(t) not in made:
                made[id(t)] = (t, self._texture(t) + 1)
            return made[id(t)][1]

        for i, m in enumerate(materials):
            if isinstance(m, DiffuseMtl):
                arr[i].type = capi.MTL_DIFFUSE
                if isinstance(m.reflectance, ImageTexture2D):
                    arr[i].reflectance_tex = tex(m.reflectance)
                else:
                    arr[i].reflectance = _f3(m.reflectance)
            elif isinstance(m, MetalMtl):
                arr[i].type = capi.MTL_METAL
                if isinstance(m.specularReflectance, ImageTexture2D):
                    arr[i].reflectance_tex = tex(m.specularReflectance)
                else:
                    arr[i].reflectance = _f3(m.specularReflectance)
                if isinstance(m.roughness, ImageTexture2D):
                    arr[i].roughness_tex = tex(m.roughness)
                else:
                    arr[i].roughness = float(m.roughness)
                arr[i].eta = _f3(m.eta)
                arr[i].k = _f3(m.k)
            elif isinstance(m, GlassMtl):
                arr[i].type = capi.MTL_GLASS
                arr[i].roughness = float(m.roughness)
                arr[i].interior_ior = float(m.interiorIOR)
            else:
                raise TypeError(f"unknown material {type(m).__name__}")
            arr[i].normal_tex = tex(getattr(m, "normalmap", None))
        self._check(self._lib.clhip_build_materials(self._ctx, arr, len(materials)))

    def buildGeometryEntities(self, entities: Sequence[Entity]):
        idx = np.array([e.materialIndex for e in entities], np.uint32)
        self._check(self._lib.clhip_build_entities(self._ctx, capi.uptr(idx), len(idx)))

    def _texture(self, tex: ImageTexture2D) -> int:
        u8 = np.asarray(tex.image).dtype == np.uint8
        img = np.asarray(tex.image, np.uint8 if u8 else np.float32)
        if img.ndim == 2:  # one channel (e.g. a roughness map): replicated to rgb
            img = img[..., None]
        if img.shape[2] == 1:
            img = np.repeat(img, 3, axis=2)
        img = np.ascontiguousarray(img)
        h, w, ch = img.shape
        tid = C.c_int32()
        if u8:
            self._check(self._lib.clhip_create_texture_u8(self._ctx, img.ctypes.data_as(C.POINTER(C.c_uint8)), w, h,
                                                          ch, int(bool(tex.isSRGB)), C.byref(tid)))
        else:
            if tex.isSRGB:
                raise ValueError("sRGB decoding applies to 8-bit images only (float images are linear)")
            self._check(self._lib.clhip_create_texture(self._ctx, capi.fptr(img), w, h, ch, C.byref(tid)))
        return tid.value

    def buildEmitters(self, emitters, domeEmitter=None, domeEmitterTextureResolution=None):
        """`domeEmitter` is the HDRIDome object (also an element of
        `emitters`) or None; the resolution is taken from its texture."""
        arr = (capi.clhip_emitter * max(1, len(emitters)))()
        dome_index = -1
        for i, e in enumerate(emitters):
            if isinstance(e, DirectionalEmitter):
                arr[i].type = capi.EMITTER_DIRECTIONAL
                arr[i].intensity = _f3(e.colorMulIntensity)
                arr[i].direction = _f3(e.sunDirection)
                arr[i].angular_radius = float(e.sunAngularRadius)
                arr[i].texture = -1
            elif isinstance(e, HDRIDome):
                arr[i].type = capi.EMITTER_HDRIDOME
                arr[i].texture = self._texture(e.texture)
                if domeEmitter is not None and e is not domeEmitter:
                    raise ValueError("only one HDRI dome is supported")
                dome_index = i
            else:
                raise TypeError(f"unknown emitter {type(e).__name__}")
        self._check(self._lib.clhip_build_emitters(self._ctx, arr, len(emitters), dome_index))

    # -- camera / parameters
    def updateCamera(self, camera: Camera):
        self._check(self._lib.clhip_set_camera(self._ctx, _f3(camera.m_camera_pos), _f3(camera.m_camera_d00),
                                               _f3(camera.m_camera_ddu), _f3(camera.m_camera_ddv)))

    def setParams(self, max_depth: int = 5, rr_depth: int = 0, view: int = 0, samples_per_batch: int = 0):
        p = capi.clhip_params(max_depth, rr_depth, view, samples_per_batch)
        self._check(self._lib.clhip_set_params(self._ctx, C.byref(p)))

    # -- render (integrator.h:61): one progressive sample per pixel
    def render(self, spp: Optional[int] = None):
        self._check(self._lib.clhip_render(self._ctx, int(self.spp_per_render if spp is None else spp)))

    def sync(self):
        self._check(self._lib.clhip_sync(self._ctx))

    def resize(self, width: int, height: int):
        self._check(self._lib.clhip_resize(self._ctx, int(width), int(height)))

    def resetIterationIndex(self):
        self._check(self._lib.clhip_reset(self._ctx))

    def getFilmSize(self):
        w, h, s = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._check(self._lib.clhip_get_film_size(self._ctx, C.byref(w), C.byref(h), C.byref(s)))
        return w.value, h.value

    def iterationIndex(self) -> int:
        w, h, s = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._check(self._lib.clhip_get_film_size(self._ctx, C.byref(w), C.byref(h), C.byref(s)))
        return s.value

    def readbackFramebuffer(self) -> np.ndarray:
        """(h, w, 4) float32 as the reference returns it: sum / spp encoded by
        convertFromLinearTosRGB (math.h:114-122, raygeneration.cu:498-506),
        alpha 1.  The film itself stays linear."""
        w, h = self.getFilmSize()
        out = np.empty((h, w, 4), np.float32)
        self._check(self._lib.clhip_readback_srgb(self._ctx, capi.fptr(out), out.nbytes))
        return out

    def readbackFramebufferLinear(self) -> np.ndarray:
        """(h, w, 4) float32 linear HDR: sum / spp, alpha 1."""
        w, h = self.getFilmSize()
        out = np.empty((h, w, 4), np.float32)
        self._check(self._lib.clhip_readback(self._ctx, capi.fptr(out), out.nbytes))
        return out

    def readbackSums(self) -> np.ndarray:
        w, h = self.getFilmSize()
        out = np.empty((h, w, 3), np.float64)
        self._check(self._lib.clhip_readback_sums(self._ctx, out.ctypes.data_as(C.POINTER(C.c_double)), out.nbytes))
        return out

    def counters(self) -> dict:
        st = capi.clhip_stats()
        self._check(self._lib.clhip_counters(self._ctx, C.byref(st)))
        d = {k: np.array(getattr(st, k)[:], np.uint64) for k in ("rays", "hits", "shadow", "visible")}
        for k in ("trace_nodes", "trace_tris", "shadow_nodes", "shadow_tris", "trace_launches", "shadow_launches",
                  "trace_ms", "shade_ms", "shadow_ms", "render_ms", "batches", "graph_replays", "graph_captures",
                  "queue_paths", "scene_bytes", "blas_builds"):
            d[k] = getattr(st, k)
        return d

    def setTiming(self, on: bool = True):
        self._check(self._lib.clhip_set_timing(self._ctx, 1 if on else 0))

    # -- launch shape and batched multi-view rendering (no reference counterpart)
    LAUNCH_FIELDS = ("trace_blocks_per_cu", "shadow_blocks_per_cu", "shade_blocks_per_cu", "paths_per_batch",
                     "graph", "order", "shade_block", "tile")

    def setLaunch(self, **fields):
        """Sets the bounce loop's launch shape; omitted or zero fields take
        their defaults (see clhip_launch)."""
        bad = set(fields) - set(self.LAUNCH_FIELDS)
        if bad:
            raise TypeError(f"unknown launch fields {sorted(bad)}")
        l = capi.clhip_launch(*[int(fields.get(k, 0)) for k in self.LAUNCH_FIELDS])
        self._check(self._lib.clhip_set_launch(self._ctx, C.byref(l)))

    def getLaunch(self) -> dict:
        l = capi.clhip_launch()
        self._check(self._lib.clhip_get_launch(self._ctx, C.byref(l)))
        return {k: getattr(l, k) for k in self.LAUNCH_FIELDS}

    def renderViews(self, views: Sequence["View"]):
        """Renders every view at its own size, spp and depth into per-view
        films through one sequence of batched launches (asynchronous)."""
        arr = (capi.clhip_view * max(1, len(views)))()
        for i, v in enumerate(views):
            cam = v.camera
            arr[i] = capi.clhip_view(int(v.width), int(v.height), int(v.spp), int(v.max_depth), int(v.rr_depth),
                                     int(v.lane), _f3(cam.m_camera_pos), _f3(cam.m_camera_d00),
                                     _f3(cam.m_camera_ddu), _f3(cam.m_camera_ddv))
        self._views = list(views)
        self._check(self._lib.clhip_render_views(self._ctx, arr, len(views)))

    def readbackView(self, i: int) -> np.ndarray:
        v = self._views[i]
        out = np.empty((v.height, v.width, 4), np.float32)
        self._check(self._lib.clhip_view_readback(self._ctx, int(i), capi.fptr(out), out.nbytes))
        return out

    def readbackViewSums(self, i: int) -> np.ndarray:
        v = self._views[i]
        out = np.empty((v.height, v.width, 3), np.float64)
        self._check(self._lib.clhip_view_sums(self._ctx, int(i), out.ctypes.data_as(C.POINTER(C.c_double)),
                                              out.nbytes))
        return out

    def viewLatencies(self) -> np.ndarray:
        """Per view: ms from the start of the last renderViews until the
        batch holding its last path finished."""
        out = np.empty(len(self._views), np.float64)
        self._check(self._lib.clhip_view_latency(self._ctx, out.ctypes.data_as(C.POINTER(C.c_double)), len(out)))
        return out

    # -- learned directional proposal (no reference counterpart)
    GUIDE_FIELDS = ("enable", "bsdf_fraction", "train", "record_paths", "batch_size", "learning_rate", "max_steps",
                    "precision", "steps_per_batch", "train_kernel")

    def setGuide(self, enable: bool = True, **fields):
        """Turns the learned proposal on or off; other fields as in
        clhip_guide (zero = default)."""
        bad = set(fields) - set(self.GUIDE_FIELDS)
        if bad:
            raise TypeError(f"unknown guide fields {sorted(bad)}")
        g = capi.clhip_guide()
        g.enable = 1 if enable else 0
        for k, v in fields.items():
            setattr(g, k, v)
        self._check(self._lib.clhip_set_guide(self._ctx, C.byref(g)))

    def guideReset(self, seed: int = 1):
        self._check(self._lib.clhip_guide_reset(self._ctx, int(seed)))

    def guideWeights(self) -> np.ndarray:
        w = np.empty(capi.GUIDE_PARAMS, np.float32)
        self._check(self._lib.clhip_guide_get_weights(self._ctx, capi.fptr(w), len(w)))
        return w

    def setGuideWeights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._check(self._lib.clhip_guide_set_weights(self._ctx, capi.fptr(w), len(w)))

    def guideStats(self) -> dict:
        st = capi.clhip_guide_stats()
        self._check(self._lib.clhip_guide_stats_get(self._ctx, C.byref(st)))
        return {k: getattr(st, k) for k, _ in capi.clhip_guide_stats._fields_}

    def guideFp8Scales(self) -> np.ndarray:
        s = np.empty(10, np.float32)
        self._check(self._lib.clhip_guide_fp8_scales(self._ctx, capi.fptr(s)))
        return s

    def guideInfer(self, q) -> np.ndarray:
        """Network outputs (n x 40) for queries (n x 9)."""
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 9)
        out = np.empty((q.shape[0], capi.GUIDE_OUTPUTS), np.float32)
        self._check(self._lib.clhip_guide_infer(self._ctx, capi.fptr(q), q.shape[0], capi.fptr(out)))
        return out

    def guideTrain(self, samples, want_grad: bool = False):
        """One Adam step on samples (n x 13: query, direction, weight);
        returns (loss, gradient or None)."""
        s = np.ascontiguousarray(samples, np.float32).reshape(-1, 13)
        g = np.empty(capi.GUIDE_PARAMS, np.float32) if want_grad else None
        loss = C.c_float()
        self._check(self._lib.clhip_guide_train(self._ctx, capi.fptr(s), s.shape[0], capi.fptr(g), C.byref(loss)))
        return loss.value, g

    def guideRecords(self, max_samples: int = 1 << 20):
        """Training records of the last recorded batch and the samples
        gathered from them: (V [4, D, S, 4], E [D, S, 4], samples [n, 13])
        with n = min(max_samples, capacity)."""
        info = (C.c_uint32 * 3)()
        self._check(self._lib.clhip_guide_records(self._ctx, None, None, None, 0, info))
        S, D, capn = info[0], info[1], info[2]
        V = np.empty((4, D, S, 4), np.float32)
        E = np.empty((D, S, 4), np.float32)
        n = min(int(max_samples), int(capn))
        smp = np.empty((n, 13), np.float32)
        self._check(self._lib.clhip_guide_records(self._ctx, capi.fptr(V), capi.fptr(E), capi.fptr(smp), n, info))
        return V, E, smp

    # -- GL interop (integrator.h:66-71) has no counterpart without a display
    def registerFramebuffer(self, glTexture):
        raise NotImplementedError("GL interop is not available on MI355X; use readbackFramebuffer()")

    unregisterFramebuffer = registerFramebuffer
    mapFramebuffer = registerFramebuffer

    # -- batch parity surfaces (same kernels)
    @property
    def ctx(self):
        return self._ctx


def upload_scene(integ: Integrator, desc) -> None:
    """Feed a scenes.SceneDesc through the reference's build order
    (RenderEngine::startRendering, renderengine.cpp:27-51)."""
    integ.buildMaterials(desc.materials)
    integ.buildGeometryEntities([Entity(m) for m in desc.instance_material])
    integ.buildBLAS(desc.meshes)
    integ.buildTLAS([desc.meshes[m] for m in desc.instance_mesh], list(range(len(desc.instance_mesh))),
                    desc.instance_xform)
    dome = next((e for e in desc.emitters if isinstance(e, HDRIDome)), None)
    integ.buildEmitters(desc.emitters, dome)
    integ.updateCamera(desc.camera)
    integ.setParams(desc.max_depth, desc.rr_depth, desc.view, desc.samples_per_batch)
