"""ctypes binding of libclpt_hip.so (C ABI declared in include/clhip.h).

The renderer has no CPU fallback: if the shared library is missing this
module raises ImportError, and every call into it on a machine without a
HIP device returns an error status that is raised as ClhipError.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libclpt_hip.so")

CLHIP_MAX_DEPTH = 32

INTEGRATOR_DIRECT_LIGHTING = 1
INTEGRATOR_PATH_TRACING = 2

MTL_DIFFUSE, MTL_METAL, MTL_GLASS = 0, 1, 2
EMITTER_DIRECTIONAL, EMITTER_HDRIDOME = 0, 1


class ClhipError(RuntimeError):
    """A clhip_* call returned a nonzero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"clhip error {code}: {msg}")
        self.code = code


class clhip_mesh(C.Structure):
    _fields_ = [
        ("vertices", C.POINTER(C.c_float)),
        ("indices", C.POINTER(C.c_uint32)),
        ("normals", C.POINTER(C.c_float)),
        ("tangents", C.POINTER(C.c_float)),
        ("uvs", C.POINTER(C.c_float)),
        ("num_vertices", C.c_uint32),
        ("num_triangles", C.c_uint32),
    ]


class clhip_material(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("reflectance", C.c_float * 3),
        ("roughness", C.c_float),
        ("eta", C.c_float * 3),
        ("k", C.c_float * 3),
        ("interior_ior", C.c_float),
        ("reflectance_tex", C.c_uint32),
        ("roughness_tex", C.c_uint32),
        ("normal_tex", C.c_uint32),
    ]


class clhip_emitter(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("intensity", C.c_float * 3),
        ("direction", C.c_float * 3),
        ("angular_radius", C.c_float),
        ("texture", C.c_int32),
    ]


class clhip_params(C.Structure):
    _fields_ = [
        ("max_depth", C.c_uint32),
        ("rr_depth", C.c_uint32),
        ("view", C.c_uint32),
        ("samples_per_batch", C.c_uint32),
    ]


_D = CLHIP_MAX_DEPTH + 1


class clhip_stats(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64 * _D),
        ("hits", C.c_uint64 * _D),
        ("shadow", C.c_uint64 * _D),
        ("visible", C.c_uint64 * _D),
        ("trace_nodes", C.c_uint64),
        ("trace_tris", C.c_uint64),
        ("shadow_nodes", C.c_uint64),
        ("shadow_tris", C.c_uint64),
        ("trace_launches", C.c_uint64),
        ("shadow_launches", C.c_uint64),
        ("trace_ms", C.c_double),
        ("shade_ms", C.c_double),
        ("shadow_ms", C.c_double),
        ("render_ms", C.c_double),
        ("batches", C.c_uint64),
        ("graph_replays", C.c_uint64),
        ("graph_captures", C.c_uint64),
        ("queue_paths", C.c_uint64),
        ("scene_bytes", C.c_uint64),
        ("blas_builds", C.c_uint64),
    ]


class clhip_view(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("spp", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("rr_depth", C.c_uint32),
        ("lane", C.c_uint32),
        ("pos", C.c_float * 3),
        ("d00", C.c_float * 3),
        ("ddu", C.c_float * 3),
        ("ddv", C.c_float * 3),
    ]


class clhip_launch(C.Structure):
    _fields_ = [
        ("trace_blocks_per_cu", C.c_uint32),
        ("shadow_blocks_per_cu", C.c_uint32),
        ("shade_blocks_per_cu", C.c_uint32),
        ("paths_per_batch", C.c_uint32),
        ("graph", C.c_uint32),
        ("order", C.c_uint32),
        ("shade_block", C.c_uint32),
        ("tile", C.c_uint32),
    ]


class clhip_guide(C.Structure):
    _fields_ = [
        ("enable", C.c_uint32),
        ("bsdf_fraction", C.c_float),
        ("train", C.c_uint32),
        ("record_paths", C.c_uint32),
        ("batch_size", C.c_uint32),
        ("learning_rate", C.c_float),
        ("max_steps", C.c_uint32),
        ("precision", C.c_uint32),
        ("steps_per_batch", C.c_uint32),
        ("train_kernel", C.c_uint32),
    ]


class clhip_guide_stats(C.Structure):
    _fields_ = [
        ("steps", C.c_uint64),
        ("samples", C.c_uint64),
        ("last_count", C.c_uint32),
        ("last_loss", C.c_float),
        ("train_ms", C.c_double),
        ("issued", C.c_uint64),
        ("training", C.c_uint32),
    ]


GUIDE_PARAMS = 32768
GUIDE_OUTPUTS = 40

P = C.c_void_p
u32, i32, f32p, u32p = C.c_uint32, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_uint32)

# name -> (restype, argtypes); this list is also what the CPU test checks
# against include/clhip.h.
SIGNATURES = {
    "clhip_create": (C.c_int, [u32, u32, u32, C.POINTER(P)]),
    "clhip_destroy": (C.c_int, [P]),
    "clhip_last_error": (C.c_char_p, [P]),
    "clhip_create_texture": (C.c_int, [P, f32p, u32, u32, u32, C.POINTER(i32)]),
    "clhip_create_texture_u8": (C.c_int, [P, C.POINTER(C.c_uint8), u32, u32, u32, C.c_int, C.POINTER(i32)]),
    "clhip_set_bvh_width": (C.c_int, [P, u32]),
    "clhip_build_blas": (C.c_int, [P, C.POINTER(clhip_mesh), u32]),
    "clhip_build_blas_subset": (C.c_int, [P, C.POINTER(clhip_mesh), u32p, u32]),
    "clhip_build_tlas": (C.c_int, [P, u32p, u32p, f32p, u32]),
    "clhip_build_materials": (C.c_int, [P, C.POINTER(clhip_material), u32]),
    "clhip_build_entities": (C.c_int, [P, u32p, u32]),
    "clhip_build_emitters": (C.c_int, [P, C.POINTER(clhip_emitter), u32, i32]),
    "clhip_set_camera": (C.c_int, [P, f32p, f32p, f32p, f32p]),
    "clhip_set_params": (C.c_int, [P, C.POINTER(clhip_params)]),
    "clhip_render": (C.c_int, [P, u32]),
    "clhip_sync": (C.c_int, [P]),
    "clhip_resize": (C.c_int, [P, u32, u32]),
    "clhip_reset": (C.c_int, [P]),
    "clhip_get_film_size": (C.c_int, [P, u32p, u32p, u32p]),
    "clhip_readback": (C.c_int, [P, f32p, C.c_size_t]),
    "clhip_readback_srgb": (C.c_int, [P, f32p, C.c_size_t]),
    "clhip_srgb_batch": (C.c_int, [P, f32p, u32, C.c_int, f32p]),
    "clhip_readback_sums": (C.c_int, [P, C.POINTER(C.c_double), C.c_size_t]),
    "clhip_counters": (C.c_int, [P, C.POINTER(clhip_stats)]),
    "clhip_set_timing": (C.c_int, [P, C.c_int]),
    "clhip_rng_batch": (C.c_int, [P, u32p, u32, u32, u32p, f32p]),
    "clhip_camera_rays": (C.c_int, [P, u32p, u32, f32p]),
    "clhip_bsdf_batch": (C.c_int, [P, u32p, f32p, f32p, f32p, u32, f32p, f32p, f32p, f32p, f32p]),
    "clhip_emitter_batch": (C.c_int, [P, f32p, u32, f32p, f32p, f32p, u32p]),
    "clhip_dome_eval_batch": (C.c_int, [P, f32p, u32, f32p, f32p]),
    "clhip_dome_tables": (C.c_int, [P, f32p, f32p, f32p, f32p, u32p, u32p]),
    "clhip_trace_batch": (C.c_int, [P, f32p, u32, f32p, f32p, u32p, u32p]),
    "clhip_occlusion_batch": (C.c_int, [P, f32p, u32, C.POINTER(C.c_uint8)]),
    "clhip_instance_inverses": (C.c_int, [P, f32p, u32]),
    "clhip_copy_probe": (C.c_int, [P, C.c_size_t, u32, C.POINTER(C.c_double)]),
    "clhip_set_launch": (C.c_int, [P, C.POINTER(clhip_launch)]),
    "clhip_get_launch": (C.c_int, [P, C.POINTER(clhip_launch)]),
    "clhip_render_views": (C.c_int, [P, C.POINTER(clhip_view), u32]),
    "clhip_view_readback": (C.c_int, [P, u32, f32p, C.c_size_t]),
    "clhip_view_sums": (C.c_int, [P, u32, C.POINTER(C.c_double), C.c_size_t]),
    "clhip_view_latency": (C.c_int, [P, C.POINTER(C.c_double), u32]),
    "clhip_image_last_error": (C.c_char_p, []),
    "clhip_image_load_hdr": (C.c_int, [C.c_char_p, f32p, C.c_size_t, u32p, u32p]),
    "clhip_image_save_hdr": (C.c_int, [C.c_char_p, f32p, u32, u32]),
    "clhip_image_load_pfm": (C.c_int, [C.c_char_p, f32p, C.c_size_t, u32p, u32p]),
    "clhip_image_save_pfm": (C.c_int, [C.c_char_p, f32p, u32, u32]),
    "clhip_image_save_exr": (C.c_int, [C.c_char_p, f32p, u32, u32]),
    "clhip_mesh_last_error": (C.c_char_p, []),
    "clhip_mesh_load_obj": (C.c_int, [C.c_char_p, f32p, f32p, f32p, f32p, u32p, u32p, u32p, u32p]),
    "clhip_set_guide": (C.c_int, [P, C.POINTER(clhip_guide)]),
    "clhip_guide_reset": (C.c_int, [P, C.c_uint64]),
    "clhip_guide_get_weights": (C.c_int, [P, f32p, u32]),
    "clhip_guide_set_weights": (C.c_int, [P, f32p, u32]),
    "clhip_guide_stats_get": (C.c_int, [P, C.POINTER(clhip_guide_stats)]),
    "clhip_guide_infer": (C.c_int, [P, f32p, u32, f32p]),
    "clhip_guide_fp8_scales": (C.c_int, [P, f32p]),
    "clhip_guide_train": (C.c_int, [P, f32p, u32, f32p, f32p]),
    "clhip_trace_rays_device": (C.c_int, [P, C.c_void_p, u32, C.c_int, C.c_void_p, C.c_void_p]),
    "clhip_guide_mixture_batch": (C.c_int, [P, f32p, f32p, f32p, u32, f32p, f32p, f32p]),
    "clhip_guide_query_batch": (C.c_int, [P, f32p, f32p, f32p, u32, f32p, f32p]),
    "clhip_guide_records": (C.c_int, [P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), u32,
                                       C.POINTER(C.c_uint32)]),
}


def load_library(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build the HIP extension first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


def last_error(ctx=None) -> str:
    msg = lib().clhip_last_error(ctx)
    return msg.decode() if msg else ""


def check(rc: int, ctx=None) -> None:
    if rc != 0:
        raise ClhipError(rc, last_error(ctx))


def fptr(a):
    return a.ctypes.data_as(f32p) if a is not None else None


def uptr(a):
    return a.ctypes.data_as(u32p) if a is not None else None
