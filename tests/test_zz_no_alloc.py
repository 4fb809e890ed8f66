"""clhip_trace_rays_device -- closest hit and any hit on caller-owned
device memory, with the traversal stacks' overflow in the library's fixed
scratch array -- against clhip_trace_batch / clhip_occlusion_batch on the
same rays, bit for bit, on two calls in a row (the second smaller), and on
both tree widths."""
import ctypes as C
import sys

import numpy as np
import pytest

from colvillea_amd import Integrator, IntegratorType, scenes, upload_scene
from colvillea_amd import _capi as capi

pytestmark = pytest.mark.gpu

f32 = np.float32


class DeviceArray:
    """Device memory of the caller (the test), through the HIP runtime the
    library itself is linked against: hipMalloc / hipMemcpy / hipFree."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            capi.lib()  # the library's runtime is loaded first; this finds the same one
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            cls._hip = h
        return cls._hip

    def __init__(self, host: np.ndarray):
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 16)) == 0
        h = np.ascontiguousarray(host)
        assert self.hip().hipMemcpy(self.ptr, h.ctypes.data, self.nbytes, 1) == 0  # host to device

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return out

    def __del__(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()


def rays(desc, n, seed, tmin):
    rng = np.random.default_rng(seed)
    o = np.stack([rng.uniform(-5, 5, n), rng.uniform(0.2, 4.0, n), rng.uniform(-5, 5, n)], 1)
    d = rng.normal(size=(n, 3))
    k = n // 3
    tgt = np.asarray(desc.instance_xform, np.float64)[1:, :, 3]
    d[:k] = tgt[rng.integers(0, len(tgt), k)] - o[:k]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, np.full((n, 1), tmin), d, np.full((n, 1), 3.4e38)], 1).astype(f32)


def trace_device(integ, r, any_hit):
    n = len(r)
    dr = DeviceArray(r)
    hit = DeviceArray(np.full((n, 4), 7.0, f32))
    prim = DeviceArray(np.full(n, 7, np.uint32))
    capi.check(capi.lib().clhip_trace_rays_device(integ.ctx, dr.ptr, n, int(any_hit), hit.ptr, prim.ptr), integ.ctx)
    integ.sync()
    return hit.numpy(), prim.numpy()


@pytest.mark.parametrize("width", [4, 8], ids=["bvh4", "bvh8"])
def test_trace_rays_device_equals_trace_batch(width):
    desc = scenes.make_c0(width=16, height=16, spp=1, hdri_w=32, hdri_h=16)
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, 16, 16) as integ:
        integ.setBVHWidth(width)
        upload_scene(integ, desc)
        for call, n in enumerate((200000, 30011)):
            print(f"allocation-free trace, BVH{width}, call {call + 1}: {n} rays", file=sys.stderr, flush=True)
            r = rays(desc, n, 100 + call, 0.0)
            t, b, inst, prim = np.zeros(n, f32), np.zeros((n, 2), f32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
            capi.check(capi.lib().clhip_trace_batch(integ.ctx, capi.fptr(r), n, capi.fptr(t), capi.fptr(b),
                                                    capi.uptr(inst), capi.uptr(prim)), integ.ctx)
            hit, p = trace_device(integ, r, False)
            assert (t >= 0).mean() > 0.3
            assert np.array_equal(hit[:, 0].view(np.uint32), t.view(np.uint32))
            assert np.array_equal(hit[:, 1:3].view(np.uint32), b.view(np.uint32))
            assert np.array_equal(hit[:, 3].view(np.uint32), inst)
            assert np.array_equal(p, prim)
            rs = rays(desc, n, 200 + call, 1e-3)
            occ = np.zeros(n, np.uint8)
            capi.check(capi.lib().clhip_occlusion_batch(integ.ctx, capi.fptr(rs), n,
                                                        occ.ctypes.data_as(C.POINTER(C.c_uint8))), integ.ctx)
            hit, p = trace_device(integ, rs, True)
            assert 0.2 < occ.mean() < 0.95
            assert np.array_equal((hit[:, 0] != 0).astype(np.uint8), occ)
            assert np.all(hit[:, 1:] == 0) and np.all(p == 0)


def test_trace_rays_device_rejects_bad_calls():
    from colvillea_amd._capi import ClhipError
    desc = scenes.make_c0(width=8, height=8, spp=1, hdri_w=32, hdri_h=16)
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, 8, 8) as integ:
        dummy = DeviceArray(np.zeros(8, f32))
        ptr = dummy.ptr
        with pytest.raises(ClhipError):  # no TLAS yet
            capi.check(capi.lib().clhip_trace_rays_device(integ.ctx, ptr, 1, 0, ptr, ptr), integ.ctx)
        upload_scene(integ, desc)
        with pytest.raises(ClhipError):  # beyond the maximum
            capi.check(capi.lib().clhip_trace_rays_device(integ.ctx, ptr, 0x80000001, 0, ptr, ptr), integ.ctx)
        with pytest.raises(ClhipError):  # missing output
            capi.check(capi.lib().clhip_trace_rays_device(integ.ctx, ptr, 1, 0, None, ptr), integ.ctx)
        # zero rays: nothing to do, and no error
        capi.check(capi.lib().clhip_trace_rays_device(integ.ctx, ptr, 0, 0, ptr, ptr), integ.ctx)
