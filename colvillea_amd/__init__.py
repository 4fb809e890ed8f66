"""colvillea_amd -- MI355X (gfx950) wavefront path tracer behind Colvillea's
Integrator / Scene / Material-BSDF / Emitter / Sampler interface.

The compute path is libclpt_hip.so (hand-written HIP kernels, C ABI in
include/clhip.h).  Importing the package loads it and fails loudly if it has
not been built; there is no CPU rendering fallback.
"""
from ._capi import ClhipError, LIB_PATH, lib as _load_lib
from .integrator import (Camera, DiffuseMtl, DirectionalEmitter, Entity, GlassMtl, HDRIDome, ImageTexture2D,
                         Integrator, IntegratorType, MetalMtl, TriangleMesh, upload_scene)

_load_lib()

__all__ = ["ClhipError", "LIB_PATH", "Camera", "DiffuseMtl", "DirectionalEmitter", "Entity", "GlassMtl", "HDRIDome",
           "ImageTexture2D", "Integrator", "IntegratorType", "MetalMtl", "TriangleMesh", "upload_scene"]
