"""ctypes loader of the oracle (oracle/_build/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The renderer never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liborc.so")
MAX_DEPTH = 64

f32p, u32p, u8p, dp = C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.POINTER(C.c_uint8), C.POINTER(C.c_double)


class orc_mesh(C.Structure):
    _fields_ = [("vertices", f32p), ("indices", u32p), ("normals", f32p), ("tangents", f32p), ("uvs", f32p),
                ("nv", C.c_uint32), ("nt", C.c_uint32)]


class orc_instance(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("material", C.c_uint32), ("xform", C.c_float * 12)]


class orc_material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("reflectance", C.c_float * 3), ("roughness", C.c_float),
                ("eta", C.c_float * 3), ("k", C.c_float * 3), ("ior", C.c_float),
                ("reflectance_tex", C.c_uint32), ("roughness_tex", C.c_uint32), ("normal_tex", C.c_uint32)]


class orc_texture(C.Structure):
    _fields_ = [("rgba", f32p), ("w", C.c_uint32), ("h", C.c_uint32)]


class orc_emitter(C.Structure):
    _fields_ = [("type", C.c_uint32), ("intensity", C.c_float * 3), ("direction", C.c_float * 3),
                ("angular_radius", C.c_float)]


class orc_dome(C.Structure):
    _fields_ = [("rgba", f32p), ("w", C.c_uint32), ("h", C.c_uint32), ("pcond", f32p), ("cdfcond", f32p),
                ("pv", f32p), ("cdfv", f32p)]


class orc_scene_desc(C.Structure):
    _fields_ = [("meshes", C.POINTER(orc_mesh)), ("nmeshes", C.c_uint32),
                ("instances", C.POINTER(orc_instance)), ("ninstances", C.c_uint32),
                ("materials", C.POINTER(orc_material)), ("nmaterials", C.c_uint32),
                ("emitters", C.POINTER(orc_emitter)), ("nemitters", C.c_uint32),
                ("dome_index", C.c_int32), ("dome", orc_dome),
                ("textures", C.POINTER(orc_texture)), ("ntextures", C.c_uint32)]


class orc_render_params(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("cam_pos", C.c_float * 3),
                ("cam_d00", C.c_float * 3), ("cam_ddu", C.c_float * 3), ("cam_ddv", C.c_float * 3),
                ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32), ("view", C.c_uint32)]


class orc_counts(C.Structure):
    _fields_ = [("rays", C.c_uint64 * (MAX_DEPTH + 1)), ("hits", C.c_uint64 * (MAX_DEPTH + 1)),
                ("shadow", C.c_uint64 * (MAX_DEPTH + 1)), ("visible", C.c_uint64 * (MAX_DEPTH + 1))]


def build():
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def use_native():
    """Build (for this host's CPU: -O3 -march=native, the same
    floating-point flags) and use oracle/_build/liborc_native.so instead of
    the portable build.  Must precede the first oracle call in the process."""
    global LIB_PATH
    path = os.path.join(HERE, "_build", "liborc_native.so")
    if _LIB is not None and LIB_PATH != path:
        raise RuntimeError("the portable oracle library is already loaded")
    subprocess.run(["make", "-s", "-C", HERE, "native"], check=True)
    LIB_PATH = path
    return path


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_pcg4d.argtypes = [u32p, u32p]
        L.orc_uint_to_float.argtypes = [C.c_uint32]
        L.orc_uint_to_float.restype = C.c_float
        L.orc_draw.argtypes = [C.c_uint32] * 4
        L.orc_draw.restype = C.c_float
        L.orc_rng_batch.argtypes = [u32p, C.c_uint32, C.c_uint32, u32p, f32p]
        L.orc_inverse3x4.argtypes = [f32p, f32p, dp]
        L.orc_elem_batch.argtypes = [C.c_int, f32p, f32p, C.c_uint32, f32p]
        L.orc_dome_tables_ref32.argtypes = [f32p, C.c_uint32, C.c_uint32, f32p, f32p, f32p, f32p]
        L.orc_dome_tables_f64.argtypes = [f32p, C.c_uint32, C.c_uint32, f32p, f32p, f32p, f32p]
        L.orc_scene_create.argtypes = [C.POINTER(orc_scene_desc)]
        L.orc_scene_create.restype = P
        L.orc_scene_destroy.argtypes = [P]
        for sfx in ("f32", "f64"):
            g = lambda n: getattr(L, f"orc_{n}_{sfx}")
            g("bsdf_eval").argtypes = [C.POINTER(orc_material), f32p, f32p, C.c_uint32, f32p, f32p]
            g("bsdf_sample").argtypes = [C.POINTER(orc_material), f32p, f32p, C.c_uint32, f32p, f32p, f32p]
            g("warp_batch").argtypes = [f32p, C.c_uint32, f32p, f32p]
            g("ggx_batch").argtypes = [C.c_float, f32p, f32p, f32p, C.c_uint32, f32p, f32p, f32p, f32p]
            g("fresnel_batch").argtypes = [f32p, C.c_uint32, f32p, f32p, C.c_float, f32p, f32p, f32p]
            g("tex_eval").argtypes = [f32p, C.c_uint32, C.c_uint32, f32p, C.c_uint32, f32p]
            g("srgb").argtypes = [f32p, C.c_uint32, C.c_int, f32p]
            g("pull_up").argtypes = [f32p, f32p, f32p, C.c_uint32, f32p]
            g("emitter_sample").argtypes = [P, f32p, C.c_uint32, f32p, f32p, f32p, u32p]
            g("dome_eval").argtypes = [P, f32p, C.c_uint32, f32p, f32p]
            g("trace").argtypes = [P, f32p, C.c_uint32, C.c_int, f32p, f32p, u32p, u32p]
            g("occluded").argtypes = [P, f32p, C.c_uint32, C.c_int, u8p]
            g("render").argtypes = [P, C.POINTER(orc_render_params), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, dp, C.POINTER(orc_counts)]
            g("render_rows").argtypes = [P, C.POINTER(orc_render_params), u32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.c_uint32, dp, C.POINTER(orc_counts)]
        _LIB = L
    return _LIB


def F(a):
    return None if a is None else np.ascontiguousarray(a, np.float32).ctypes.data_as(f32p)


def _fp(a):
    return None if a is None else a.ctypes.data_as(f32p)


def _up(a):
    return None if a is None else a.ctypes.data_as(u32p)


def texture_rgba(t) -> np.ndarray:
    """An ImageTexture2D as contiguous RGBA float32 (1 or 3 channels are
    widened as the renderer widens them: grey -> rgb, alpha 1).  8-bit
    images read as c / 255, sRGB ones decoded by the oracle's decoder."""
    if np.asarray(t.image).dtype == np.uint8:
        img = np.asarray(t.image).astype(np.float32) / np.float32(255.0)
        if getattr(t, "isSRGB", False):
            if img.ndim == 2:
                img = srgb(img, decode=True)
            else:
                img = np.concatenate([srgb(img[..., :3], decode=True), img[..., 3:]], -1)
    else:
        img = np.asarray(t.image, np.float32)
    if img.ndim == 2:
        img = img[..., None]
    if img.shape[2] == 1:
        img = np.repeat(img, 3, axis=2)
    if img.shape[2] == 3:
        img = np.concatenate([img, np.ones(img.shape[:2] + (1,), np.float32)], 2)
    return np.ascontiguousarray(img, np.float32)


def material_struct(m, tex=None) -> orc_material:
    """Convert an integrator material (DiffuseMtl / MetalMtl / GlassMtl).
    tex(texture) -> index + 1 in the scene's texture table."""
    from colvillea_amd.integrator import DiffuseMtl, GlassMtl, ImageTexture2D, MetalMtl

    def tid(t):
        if t is None:
            return 0
        if tex is None:
            raise ValueError("a textured material needs a scene texture table")
        return tex(t)

    s = orc_material()
    if isinstance(m, DiffuseMtl):
        s.type = 0
        if isinstance(m.reflectance, ImageTexture2D):
            s.reflectance_tex = tid(m.reflectance)
        else:
            s.reflectance = (C.c_float * 3)(*m.reflectance)
    elif isinstance(m, MetalMtl):
        s.type = 1
        if isinstance(m.specularReflectance, ImageTexture2D):
            s.reflectance_tex = tid(m.specularReflectance)
        else:
            s.reflectance = (C.c_float * 3)(*m.specularReflectance)
        if isinstance(m.roughness, ImageTexture2D):
            s.roughness_tex = tid(m.roughness)
        else:
            s.roughness = m.roughness
        s.eta = (C.c_float * 3)(*m.eta)
        s.k = (C.c_float * 3)(*m.k)
    elif isinstance(m, GlassMtl):
        s.type = 2
        s.roughness = m.roughness
        s.ior = m.interiorIOR
    else:
        raise TypeError(type(m))
    s.normal_tex = tid(getattr(m, "normalmap", None))
    return s


def dome_tables(rgba: np.ndarray, variant: str = "f64"):
    """The oracle's HDRI tables: 'ref32' = the reference's float32 serial
    algorithm, 'f64' = the renderer's float64-accumulated definition."""
    rgba = np.ascontiguousarray(rgba, np.float32)
    h, w = rgba.shape[:2]
    pc = np.zeros((h, w + 1), np.float32)
    cc = np.zeros((h, w + 1), np.float32)
    pv = np.zeros(h + 1, np.float32)
    cv = np.zeros(h + 1, np.float32)
    fn = lib().orc_dome_tables_ref32 if variant == "ref32" else lib().orc_dome_tables_f64
    fn(_fp(rgba), w, h, _fp(pc), _fp(cc), _fp(pv), _fp(cv))
    return pc, cc, pv, cv


class Scene:
    """An oracle scene built from a colvillea_amd.scenes.SceneDesc.  Dome
    tables default to the oracle's own float64 build; pass `tables` (for
    example the renderer's, read back) to share them exactly."""

    def __init__(self, desc, tables=None):
        from colvillea_amd.integrator import DirectionalEmitter, HDRIDome
        L = lib()
        self._keep = []
        self.desc = desc
        meshes = (orc_mesh * max(1, len(desc.meshes)))()
        for i, m in enumerate(desc.meshes):
            arrs = [np.ascontiguousarray(x, np.float32) if x is not None else None
                    for x in (m.vertices, m.normals, m.tangents, m.uvs)]
            idx = np.ascontiguousarray(m.triangles, np.uint32)
            self._keep += arrs + [idx]
            meshes[i] = orc_mesh(_fp(arrs[0]), _up(idx), _fp(arrs[1]), _fp(arrs[2]), _fp(arrs[3]),
                                 m.vertices.shape[0], m.triangles.shape[0])
        n = len(desc.instance_mesh)
        inst = (orc_instance * max(1, n))()
        for i in range(n):
            inst[i].mesh = desc.instance_mesh[i]
            inst[i].material = desc.instance_material[i]
            inst[i].xform = (C.c_float * 12)(*np.asarray(desc.instance_xform[i], np.float32).reshape(-1))
        texs, tex_index = [], {}

        def tex(t):
            if id(t) not in tex_index:
                rgba = texture_rgba(t)
                self._keep.append(rgba)
                texs.append(orc_texture(_fp(rgba), rgba.shape[1], rgba.shape[0]))
                tex_index[id(t)] = len(texs)
            return tex_index[id(t)]

        mats = (orc_material * max(1, len(desc.materials)))()
        for i, m in enumerate(desc.materials):
            mats[i] = material_struct(m, tex)
        tarr = (orc_texture * max(1, len(texs)))(*texs)
        ems = (orc_emitter * max(1, len(desc.emitters)))()
        dome_index = -1
        dome = orc_dome()
        for i, e in enumerate(desc.emitters):
            if isinstance(e, DirectionalEmitter):
                ems[i].type = 0
                ems[i].intensity = (C.c_float * 3)(*e.colorMulIntensity)
                ems[i].direction = (C.c_float * 3)(*e.sunDirection)
                ems[i].angular_radius = e.sunAngularRadius
            elif isinstance(e, HDRIDome):
                ems[i].type = 1
                dome_index = i
                rgba = np.ascontiguousarray(e.texture.image, np.float32)
                if rgba.shape[2] == 3:
                    rgba = np.concatenate([rgba, np.ones(rgba.shape[:2] + (1,), np.float32)], 2)
                tb = tables if tables is not None else dome_tables(rgba)
                tb = [np.ascontiguousarray(t, np.float32) for t in tb]
                self._keep += [rgba] + tb
                self.tables = tb
                dome = orc_dome(_fp(rgba), rgba.shape[1], rgba.shape[0], _fp(tb[0]), _fp(tb[1]), _fp(tb[2]),
                                _fp(tb[3]))
        self._keep += [meshes, inst, mats, ems, tarr]
        d = orc_scene_desc(meshes, len(desc.meshes), inst, n, mats, len(desc.materials), ems, len(desc.emitters),
                           dome_index, dome, tarr, len(texs))
        self._desc_struct = d
        self.handle = L.orc_scene_create(C.byref(d))

    def close(self):
        if self.handle:
            lib().orc_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self, width=None, height=None, max_depth=None, rr_depth=None, view=None) -> orc_render_params:
        d = self.desc
        p = orc_render_params()
        p.width = d.width if width is None else width
        p.height = d.height if height is None else height
        p.cam_pos = (C.c_float * 3)(*d.camera.m_camera_pos)
        p.cam_d00 = (C.c_float * 3)(*d.camera.m_camera_d00)
        p.cam_ddu = (C.c_float * 3)(*d.camera.m_camera_ddu)
        p.cam_ddv = (C.c_float * 3)(*d.camera.m_camera_ddv)
        p.max_depth = d.max_depth if max_depth is None else max_depth
        p.rr_depth = d.rr_depth if rr_depth is None else rr_depth
        p.view = d.view if view is None else view
        return p

    def render(self, variant="f32", rows=None, samples=(0, 1), nthreads=None, params=None):
        """Float64 RGB sums of rows [y0, y1) over samples [s0, s1) and the
        per-depth counts."""
        p = params or self.params()
        y0, y1 = rows if rows is not None else (0, p.height)
        out = np.zeros((y1 - y0, p.width, 3), np.float64)
        cnt = orc_counts()
        nthreads = nthreads or min(16, os.cpu_count() or 1)
        getattr(lib(), f"orc_render_{variant}")(self.handle, C.byref(p), y0, y1, samples[0], samples[1], nthreads,
                                                 out.ctypes.data_as(dp), C.byref(cnt))
        counts = {k: np.array(getattr(cnt, k)[:], np.uint64) for k in ("rays", "hits", "shadow", "visible")}
        return out, counts

    def render_rows(self, rows, variant="f32", samples=(0, 1), nthreads=None, params=None):
        """Float64 RGB sums of an explicit row list (nrows, width, 3)."""
        p = params or self.params()
        rows = np.ascontiguousarray(rows, np.uint32)
        out = np.zeros((len(rows), p.width, 3), np.float64)
        cnt = orc_counts()
        nthreads = nthreads or min(16, os.cpu_count() or 1)
        getattr(lib(), f"orc_render_rows_{variant}")(self.handle, C.byref(p), _up(rows), len(rows), samples[0],
                                                      samples[1], nthreads, out.ctypes.data_as(dp), C.byref(cnt))
        counts = {k: np.array(getattr(cnt, k)[:], np.uint64) for k in ("rays", "hits", "shadow", "visible")}
        return out, counts

    def trace(self, rays8: np.ndarray, brute=False, variant="f32"):
        rays8 = np.ascontiguousarray(rays8, np.float32)
        n = rays8.shape[0]
        t = np.zeros(n, np.float32)
        b = np.zeros((n, 2), np.float32)
        inst = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.uint32)
        getattr(lib(), f"orc_trace_{variant}")(self.handle, _fp(rays8), n, 1 if brute else 0, _fp(t), _fp(b),
                                                _up(inst), _up(prim))
        return t, b, inst, prim

    def occluded(self, rays8: np.ndarray, brute=False, variant="f32"):
        rays8 = np.ascontiguousarray(rays8, np.float32)
        n = rays8.shape[0]
        occ = np.zeros(n, np.uint8)
        getattr(lib(), f"orc_occluded_{variant}")(self.handle, _fp(rays8), n, 1 if brute else 0,
                                                   occ.ctypes.data_as(u8p))
        return occ

    def emitter_sample(self, u2: np.ndarray, variant="f32"):
        u2 = np.ascontiguousarray(u2, np.float32)
        n = u2.shape[0]
        d = np.zeros((n, 3), np.float32)
        v = np.zeros((n, 3), np.float32)
        p = np.zeros(n, np.float32)
        w = np.zeros(n, np.uint32)
        getattr(lib(), f"orc_emitter_sample_{variant}")(self.handle, _fp(u2), n, _fp(d), _fp(v), _fp(p), _up(w))
        return d, v, p, w

    def dome_eval(self, dirs: np.ndarray, variant="f32"):
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = dirs.shape[0]
        le = np.zeros((n, 3), np.float32)
        p = np.zeros(n, np.float32)
        getattr(lib(), f"orc_dome_eval_{variant}")(self.handle, _fp(dirs), n, _fp(le), _fp(p))
        return le, p


def bsdf_eval(materials, wi, wo, variant="f32"):
    n = len(wi)
    ms = (orc_material * n)(*[material_struct(m) for m in materials])
    wi = np.ascontiguousarray(wi, np.float32)
    wo = np.ascontiguousarray(wo, np.float32)
    f = np.zeros((n, 3), np.float32)
    p = np.zeros(n, np.float32)
    getattr(lib(), f"orc_bsdf_eval_{variant}")(ms, _fp(wi), _fp(wo), n, _fp(f), _fp(p))
    return f, p


def bsdf_sample(materials, wo, u3, variant="f32"):
    n = len(wo)
    ms = (orc_material * n)(*[material_struct(m) for m in materials])
    wo = np.ascontiguousarray(wo, np.float32)
    u3 = np.ascontiguousarray(u3, np.float32)
    wi = np.zeros((n, 3), np.float32)
    f = np.zeros((n, 3), np.float32)
    p = np.zeros(n, np.float32)
    getattr(lib(), f"orc_bsdf_sample_{variant}")(ms, _fp(wo), _fp(u3), n, _fp(wi), _fp(f), _fp(p))
    return wi, f, p


def rng_batch(seeds4: np.ndarray, ndraws: int):
    seeds4 = np.ascontiguousarray(seeds4, np.uint32)
    n = seeds4.shape[0]
    u = np.zeros((n, ndraws), np.uint32)
    f = np.zeros((n, ndraws), np.float32)
    lib().orc_rng_batch(_up(seeds4), n, ndraws, _up(u), _fp(f))
    return u, f


def pcg4d(v) -> np.ndarray:
    a = np.ascontiguousarray(v, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_pcg4d(_up(a), _up(o))
    return o


def warp_batch(u2, variant="f32"):
    u2 = np.ascontiguousarray(u2, np.float32)
    n = u2.shape[0]
    d = np.zeros((n, 2), np.float32)
    h = np.zeros((n, 3), np.float32)
    getattr(lib(), f"orc_warp_batch_{variant}")(_fp(u2), n, _fp(d), _fp(h))
    return d, h


def ggx_batch(alpha, m3, v3, u2, variant="f32"):
    m3, v3, u2 = (np.ascontiguousarray(a, np.float32) for a in (m3, v3, u2))
    n = m3.shape[0]
    D = np.zeros(n, np.float32)
    G1 = np.zeros(n, np.float32)
    ms = np.zeros((n, 3), np.float32)
    mp = np.zeros(n, np.float32)
    getattr(lib(), f"orc_ggx_batch_{variant}")(alpha, _fp(m3), _fp(v3), _fp(u2), n, _fp(D), _fp(G1), _fp(ms), _fp(mp))
    return D, G1, ms, mp


def fresnel_batch(cos_i, eta3, k3, ior, variant="f32"):
    cos_i = np.ascontiguousarray(cos_i, np.float32)
    n = cos_i.shape[0]
    fc = np.zeros((n, 3), np.float32)
    fd = np.zeros(n, np.float32)
    ct = np.zeros(n, np.float32)
    e = np.ascontiguousarray(eta3, np.float32)
    k = np.ascontiguousarray(k3, np.float32)
    getattr(lib(), f"orc_fresnel_batch_{variant}")(_fp(cos_i), n, _fp(e), _fp(k), ior, _fp(fc), _fp(fd), _fp(ct))
    return fc, fd, ct


def tex_eval(rgba, uv, variant="f32"):
    rgba = np.ascontiguousarray(rgba, np.float32)
    uv = np.ascontiguousarray(uv, np.float32)
    n = uv.shape[0]
    out = np.zeros((n, 4), np.float32)
    getattr(lib(), f"orc_tex_eval_{variant}")(_fp(rgba), rgba.shape[1], rgba.shape[0], _fp(uv), n, _fp(out))
    return out


def pull_up(s, t, n, ng, wo, variant="f32"):
    """Material::fixNormalPerturbingShadingFrame (material.h:236-285) on
    frames (s, t, n), geometric normals and wo: the fixed (s, t, n)."""
    fr = np.ascontiguousarray(np.concatenate([s, t, n], 1), np.float32)
    ng = np.ascontiguousarray(ng, np.float32)
    wo = np.ascontiguousarray(wo, np.float32)
    out = np.zeros_like(fr)
    getattr(lib(), f"orc_pull_up_{variant}")(_fp(fr), _fp(ng), _fp(wo), len(fr), _fp(out))
    return out[:, 0:3], out[:, 3:6], out[:, 6:9]


def srgb(x, decode=False, variant="f32"):
    """convertFromLinearTosRGB (decode=False) or convertsRGBToLinear
    (math.h:114-140) per value."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    getattr(lib(), f"orc_srgb_{variant}")(_fp(x), x.size, int(bool(decode)), _fp(out))
    return out


ELEM = {"sin": 0, "cos": 1, "atan2": 2, "acos": 3, "asin": 4, "exp": 5, "log": 6, "tan": 7}


def elem(fn: str, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(x if y is None else y, np.float32)
    out = np.zeros_like(x)
    lib().orc_elem_batch(ELEM[fn], _fp(x), _fp(y), x.size, _fp(out))
    return out


def inverse3x4(m):
    m = np.ascontiguousarray(m, np.float32).reshape(12)
    o = np.zeros(12, np.float32)
    od = np.zeros(12, np.float64)
    lib().orc_inverse3x4(_fp(m), _fp(o), od.ctypes.data_as(dp))
    return o, od
