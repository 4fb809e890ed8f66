"""Files through the library (no stb / FreeImage / assimp): Radiance .hdr
read/write, PFM read/write, OpenEXR write, Wavefront OBJ import.  Image
arrays are (h, w, 4) float32 RGBA with row 0 at the top; RGB input gets
alpha 1."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi


def _err() -> str:
    m = capi.lib().clhip_image_last_error()
    return m.decode() if m else ""


def _rgba(img) -> np.ndarray:
    a = np.asarray(img, np.float32)
    if a.ndim != 3 or a.shape[2] not in (3, 4):
        raise ValueError("expected an (h, w, 3|4) image")
    if a.shape[2] == 3:
        a = np.concatenate([a, np.ones(a.shape[:2] + (1,), np.float32)], 2)
    return np.ascontiguousarray(a)


def _load(fn, path: str) -> np.ndarray:
    w, h = C.c_uint32(), C.c_uint32()
    p = str(path).encode()
    if fn(p, None, 0, C.byref(w), C.byref(h)) != 0:
        raise OSError(_err())
    out = np.empty((h.value, w.value, 4), np.float32)
    if fn(p, capi.fptr(out), out.nbytes, C.byref(w), C.byref(h)) != 0:
        raise OSError(_err())
    return out


def _save(fn, path: str, img) -> None:
    a = _rgba(img)
    if fn(str(path).encode(), capi.fptr(a), a.shape[1], a.shape[0]) != 0:
        raise OSError(_err())


def load_hdr(path) -> np.ndarray:
    return _load(capi.lib().clhip_image_load_hdr, path)


def save_hdr(path, img) -> None:
    _save(capi.lib().clhip_image_save_hdr, path, img)


def load_pfm(path) -> np.ndarray:
    return _load(capi.lib().clhip_image_load_pfm, path)


def save_pfm(path, img) -> None:
    _save(capi.lib().clhip_image_save_pfm, path, img)


def save_exr(path, img) -> None:
    _save(capi.lib().clhip_image_save_exr, path, img)


def load_obj(path):
    """A TriangleMesh from a Wavefront OBJ file (triangulated, smooth
    normals when absent, flipped texture v, tangents when textured)."""
    from .integrator import TriangleMesh

    lib = capi.lib()
    nv, nt, huv = C.c_uint32(), C.c_uint32(), C.c_uint32()
    p = str(path).encode()
    if lib.clhip_mesh_load_obj(p, None, None, None, None, None, C.byref(nv), C.byref(nt), C.byref(huv)) != 0:
        raise OSError(lib.clhip_mesh_last_error().decode())
    v = np.empty((nv.value, 3), np.float32)
    n = np.empty((nv.value, 3), np.float32)
    t = np.empty((nv.value, 3), np.float32)
    uv = np.empty((nv.value, 2), np.float32)
    idx = np.empty((nt.value, 3), np.uint32)
    if lib.clhip_mesh_load_obj(p, capi.fptr(v), capi.fptr(n), capi.fptr(t), capi.fptr(uv), capi.uptr(idx),
                               C.byref(nv), C.byref(nt), C.byref(huv)) != 0:
        raise OSError(lib.clhip_mesh_last_error().decode())
    if huv.value:
        return TriangleMesh(v, idx, normals=n, tangents=t, uvs=uv)
    return TriangleMesh(v, idx, normals=n)
