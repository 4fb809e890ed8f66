"""GPU parity: the gfx950 kernels (through the C ABI) against the oracle.

Bit-exact: RNG words and floats, camera rays, BSDF eval/pdf/sample, emitter
samples (indices, directions, values, pdfs), HDRI lookups, hit instance /
primitive ids and t, occlusion, per-depth queue counts.  The HDRI tables are
float64 prefix sums in a different summation order than the oracle's serial
loop, so they are compared to <= 1 ulp; every downstream comparison feeds the
GPU's tables to the oracle so both sample from identical data.
"""
import ctypes as C

import numpy as np
import pytest

import orc
from colvillea_amd import GlassMtl, Integrator, IntegratorType, MetalMtl, DiffuseMtl, scenes, upload_scene
from colvillea_amd.integrator import ImageTexture2D
from colvillea_amd import _capi as capi

pytestmark = pytest.mark.gpu

f32 = np.float32


def ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture(scope="module", params=[4, 8], ids=["bvh4", "bvh8"])
def c0_small(request):
    """The small C0 scene, built once per BVH width (4: float boxes, 8:
    byte-quantised boxes); every test on it must pass for both trees."""
    desc = scenes.make_c0(width=64, height=48, spp=4, hdri_w=128, hdri_h=64)
    integ = Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height)
    integ.setBVHWidth(request.param)
    upload_scene(integ, desc)
    yield desc, integ
    integ.close()


def gpu_tables(integ, desc):
    h, w = desc.hdri.shape[:2]
    pc = np.zeros((h, w + 1), f32)
    cc = np.zeros((h, w + 1), f32)
    pv = np.zeros(h + 1, f32)
    cv = np.zeros(h + 1, f32)
    capi.check(capi.lib().clhip_dome_tables(integ.ctx, capi.fptr(pc), capi.fptr(cc), capi.fptr(pv), capi.fptr(cv),
                                            None, None), integ.ctx)
    return pc, cc, pv, cv


def test_rng_bit_exact(c0_small):
    _, integ = c0_small
    rng = np.random.default_rng(11)
    seeds = rng.integers(0, 2**32, size=(4096, 4), dtype=np.uint64).astype(np.uint32)
    seeds[:16] = [[x, y, 0, 0] for x in range(4) for y in range(4)]
    nd = 9
    gu = np.zeros((4096, nd), np.uint32)
    gf = np.zeros((4096, nd), f32)
    capi.check(capi.lib().clhip_rng_batch(integ.ctx, capi.uptr(seeds), 4096, nd, capi.uptr(gu), capi.fptr(gf)),
               integ.ctx)
    ou, of = orc.rng_batch(seeds, nd)
    assert np.array_equal(gu, ou)
    assert bits_equal(gf, of)


def test_camera_rays_bit_exact(c0_small):
    desc, integ = c0_small
    npix = desc.width * desc.height
    ps = np.stack([np.arange(npix, dtype=np.uint32), np.full(npix, 3, np.uint32)], 1)
    out = np.zeros((npix, 8), f32)
    capi.check(capi.lib().clhip_camera_rays(integ.ctx, capi.uptr(ps), npix, capi.fptr(out)), integ.ctx)
    # oracle: same formula via orc_draw + the reference's normalize form
    px, py = ps[:, 0] % desc.width, ps[:, 0] // desc.width
    L = orc.lib()
    u0 = np.array([L.orc_draw(int(x), int(y), 3, 1) for x, y in zip(px, py)], f32)
    u1 = np.array([L.orc_draw(int(x), int(y), 3, 2) for x, y in zip(px, py)], f32)
    cam = desc.camera
    sx = (px.astype(f32) + u0) / f32(desc.width)
    sy = (py.astype(f32) + u1) / f32(desc.height)
    d00, ddu, ddv = (np.asarray(v, f32) for v in (cam.m_camera_d00, cam.m_camera_ddu, cam.m_camera_ddv))
    d = (d00[None] + sx[:, None] * ddu[None]) + sy[:, None] * ddv[None]
    dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    inv = f32(1.0) / np.sqrt(dd)
    d = d * inv[:, None]
    assert bits_equal(out[:, 4:7], d)
    assert bits_equal(out[:, 0:3], np.broadcast_to(np.asarray(cam.m_camera_pos, f32), (npix, 3)))


def _bsdf_gpu(integ, mtl_idx, wi, wo, u3):
    n = len(mtl_idx)
    out = [np.zeros((n, 3), f32), np.zeros(n, f32), np.zeros((n, 3), f32), np.zeros((n, 3), f32), np.zeros(n, f32)]
    capi.check(capi.lib().clhip_bsdf_batch(integ.ctx, capi.uptr(np.ascontiguousarray(mtl_idx, np.uint32)),
                                           capi.fptr(wi), capi.fptr(wo), capi.fptr(u3), n,
                                           *[capi.fptr(o) for o in out]), integ.ctx)
    return out


def random_dirs(rng, n, upper=None):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    if upper is True:
        d[:, 2] = np.abs(d[:, 2])
    return d.astype(f32)


def test_bsdf_batch_bit_exact():
    mats = [DiffuseMtl((0.7, 0.5, 0.3)),
            MetalMtl((1.0, 0.9, 0.8), 0.3, (0.143, 0.374, 1.442), (3.983, 2.385, 1.603)),
            MetalMtl((1.0, 1.0, 1.0), 0.05, (0.2, 0.92, 1.1), (3.9, 2.45, 2.14)),
            GlassMtl(0.4, 1.5), GlassMtl(0.1, 1.33), GlassMtl(0.7, 2.4)]
    integ = Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, 8, 8)
    integ.buildMaterials(mats)
    rng = np.random.default_rng(5)
    n = 60000
    idx = rng.integers(0, len(mats), n).astype(np.uint32)
    wo = random_dirs(rng, n)
    wi = random_dirs(rng, n)
    # half the wo in the upper hemisphere so reflection lobes are exercised
    wo[: n // 2, 2] = np.abs(wo[: n // 2, 2])
    u3 = rng.random((n, 3), dtype=np.float64).astype(f32)
    ef, ep, sw, sf, sp = _bsdf_gpu(integ, idx, wi, wo, u3)
    integ.close()
    mlist = [mats[i] for i in idx]
    of, op = orc.bsdf_eval(mlist, wi, wo)
    owi, osf, osp = orc.bsdf_sample(mlist, wo, u3)
    assert bits_equal(ef, of), f"eval f: max ulp {ulps(ef, of).max()}"
    assert bits_equal(ep, op), f"eval pdf: max ulp {ulps(ep, op).max()}"
    assert bits_equal(sf, osf), f"sample f: max ulp {ulps(sf, osf).max()}"
    assert bits_equal(sp, osp), f"sample pdf: max ulp {ulps(sp, osp).max()}"
    ok = (osf > 0).any(axis=1)
    assert bits_equal(sw[ok], owi[ok])
    assert ok.mean() > 0.3


def test_dome_tables_vs_oracle(c0_small):
    desc, integ = c0_small
    g = gpu_tables(integ, desc)
    o = orc.dome_tables(desc.hdri, "f64")
    h, w = desc.hdri.shape[:2]
    # prefiltered row values are exact (same float32 expression)
    assert bits_equal(g[0][:, :w], o[0][:, :w])
    for a, b in zip(g, o):
        assert ulps(a, b).max() <= 1


def test_emitter_sampling_bit_exact(c0_small):
    desc, integ = c0_small
    ref = orc.Scene(desc, tables=gpu_tables(integ, desc))
    rng = np.random.default_rng(3)
    n = 50000
    u2 = rng.random((n, 2)).astype(f32)
    gd, gv, gp, gw = (np.zeros((n, 3), f32), np.zeros((n, 3), f32), np.zeros(n, f32), np.zeros(n, np.uint32))
    capi.check(capi.lib().clhip_emitter_batch(integ.ctx, capi.fptr(u2), n, capi.fptr(gd), capi.fptr(gv),
                                              capi.fptr(gp), capi.uptr(gw)), integ.ctx)
    od, ov, op, ow = ref.emitter_sample(u2)
    assert np.array_equal(gw, ow)
    assert bits_equal(gp, op)
    assert bits_equal(gv, ov)
    assert bits_equal(gd, od)
    # both emitters are chosen, pdfs positive where values are
    assert set(np.unique(gw)) == {0, 1}
    assert (gp[(gv > 0).any(1)] > 0).all()


def test_dome_eval_bit_exact(c0_small):
    desc, integ = c0_small
    ref = orc.Scene(desc, tables=gpu_tables(integ, desc))
    rng = np.random.default_rng(4)
    n = 50000
    d = random_dirs(rng, n)
    d[:4] = [[0, 1, 0], [0, -1, 0], [1, 0, 0], [0, 0, -1]]
    # |y| just above 1, as a float32 frame rotation can produce: clamped
    d[4:6] = [[4.27260995e-4, 1.00000012, 2.73570418e-4], [1e-4, -1.0000001, -3e-4]]
    le, p = np.zeros((n, 3), f32), np.zeros(n, f32)
    capi.check(capi.lib().clhip_dome_eval_batch(integ.ctx, capi.fptr(d), n, capi.fptr(le), capi.fptr(p)), integ.ctx)
    ole, op = ref.dome_eval(d)
    assert np.isfinite(le).all() and np.isfinite(p).all()
    assert bits_equal(le, ole)
    assert bits_equal(p, op)


def _trace_rays(desc, n, seed, tmin=0.0):
    rng = np.random.default_rng(seed)
    o = np.stack([rng.uniform(-5, 5, n), rng.uniform(0.05, 3.0, n), rng.uniform(-5, 5, n)], 1)
    d = random_dirs(rng, n)
    # aim a third of the rays at sphere centres so most of them hit
    k = n // 3
    tgt = np.asarray(desc.instance_xform, np.float32)[1:, :, 3]
    pick = tgt[rng.integers(0, len(tgt), k)]
    v = pick - o[:k]
    d[:k] = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(f32)
    return np.concatenate([o, np.full((n, 1), tmin), d, np.full((n, 1), 3.4e38)], 1).astype(f32)


def test_closest_hit_ids_vs_brute_force(c0_small):
    desc, integ = c0_small
    ref = orc.Scene(desc)
    n = 20000
    rays = _trace_rays(desc, n, 7)
    t, b, inst, prim = np.zeros(n, f32), np.zeros((n, 2), f32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    capi.check(capi.lib().clhip_trace_batch(integ.ctx, capi.fptr(rays), n, capi.fptr(t), capi.fptr(b),
                                            capi.uptr(inst), capi.uptr(prim)), integ.ctx)
    ot, ob, oi, op = ref.trace(rays, brute=True)
    assert (ot >= 0).mean() > 0.3
    assert np.array_equal(inst, oi)
    assert np.array_equal(prim, op)
    assert bits_equal(t, ot)
    assert bits_equal(b, ob)


def test_occlusion_vs_brute_force(c0_small):
    desc, integ = c0_small
    ref = orc.Scene(desc)
    n = 20000
    rays = _trace_rays(desc, n, 8, tmin=1e-3)
    occ = np.zeros(n, np.uint8)
    capi.check(capi.lib().clhip_occlusion_batch(integ.ctx, capi.fptr(rays), n, occ.ctypes.data_as(
        C.POINTER(C.c_uint8))), integ.ctx)
    oo = ref.occluded(rays, brute=True)
    assert np.array_equal(occ, oo)
    assert 0.2 < occ.mean() < 0.95


def test_instance_inverse_matches_oracle(c0_small):
    desc, integ = c0_small
    n = len(desc.instance_mesh)
    out = np.zeros((n, 12), f32)
    capi.check(capi.lib().clhip_instance_inverses(integ.ctx, capi.fptr(out), n), integ.ctx)
    for i in range(n):
        o, _ = orc.inverse3x4(desc.instance_xform[i])
        assert bits_equal(out[i], o)


def test_render_counts_and_image_match_oracle(c0_small):
    desc, integ = c0_small
    integ.resetIterationIndex()
    integ.render(desc.spp)
    sums = integ.readbackSums()
    cnt = integ.counters()
    ref = orc.Scene(desc, tables=gpu_tables(integ, desc))
    osums, ocnt = ref.render("f32", samples=(0, desc.spp))
    D = desc.max_depth + 1
    for k in ("rays", "hits", "shadow", "visible"):
        assert np.array_equal(cnt[k][:D], ocnt[k][:D]), (k, cnt[k][:D], ocnt[k][:D])
    assert cnt["rays"][1] == desc.width * desc.height * desc.spp
    # float64 sums of identical float32 contributions, different order
    rel = np.mean((sums - osums) ** 2 / (osums ** 2 + 1e-2))
    assert rel < 1e-12, rel
    img = integ.readbackFramebufferLinear()
    assert np.allclose(img[..., :3], (sums / desc.spp).astype(f32), rtol=1e-6, atol=1e-7)
    # the reference's readbackFramebuffer: the same resolve, sRGB-encoded
    # (raygeneration.cu:498-506), against the oracle's encoder (<= 4 ulp)
    disp = integ.readbackFramebuffer()
    assert np.all(disp[..., 3] == 1.0)
    assert ulps(disp[..., :3], orc.srgb(img[..., :3], decode=False)).max() <= 4


def test_render_matches_f64_oracle():
    """Same seeds, FP32 GPU vs the FP64 oracle: relMSE far below 1e-4."""
    desc = scenes.make_c0(width=48, height=48, spp=8, hdri_w=128, hdri_h=64)
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height) as integ:
        upload_scene(integ, desc)
        integ.render(desc.spp)
        img = integ.readbackSums() / desc.spp
        tb = gpu_tables(integ, desc)
    ref = orc.Scene(desc, tables=tb)
    o64, _ = ref.render("f64", samples=(0, desc.spp))
    o64 /= desc.spp
    rel = float(np.mean((img - o64) ** 2 / (o64 ** 2 + 1e-2)))
    assert rel < 1e-4, rel


def test_progressive_render_accumulates(c0_small):
    """Two render(2) calls equal one render(4) (counter-based RNG)."""
    desc, integ = c0_small
    integ.resetIterationIndex()
    integ.render(4)
    a = integ.readbackSums()
    integ.resetIterationIndex()
    integ.render(2)
    integ.render(2)
    b = integ.readbackSums()
    assert integ.iterationIndex() == 4
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12)


def test_direct_lighting_integrator_matches_oracle():
    """WavefrontDirectLighting (direct.cpp:84-201, raygeneration.cu:139-282):
    light sample + BSDF sample at the first hit, MIS-weighted; the oracle
    restates it as the path tracer stopped after two segments without
    Russian roulette.  Counts exact, sums to float64 order."""
    desc = scenes.make_c0(width=48, height=40, spp=4, hdri_w=128, hdri_h=64)
    with Integrator.createIntegrator(IntegratorType.WavefrontDirectLighting, desc.width, desc.height) as it:
        upload_scene(it, desc)  # scene params ask for depth 5 + RR: the DL integrator ignores both
        it.render(desc.spp)
        sums = it.readbackSums()
        cnt = it.counters()
        tb = gpu_tables(it, desc)
    ref = orc.Scene(desc, tables=tb)
    osums, ocnt = ref.render("f32", samples=(0, desc.spp), params=ref.params(max_depth=2, rr_depth=0))
    for k in ("rays", "hits", "shadow", "visible"):
        assert np.array_equal(cnt[k][:3], ocnt[k][:3]), k
        assert cnt[k][3:].sum() == 0, k
    rel = np.mean((sums - osums) ** 2 / (osums ** 2 + 1e-2))
    assert rel < 1e-12, rel


def test_textured_materials_match_oracle():
    """Reflectance, roughness and normal-map textures (Material::getBSDF,
    getShadingFrame with the pull-up fix) on uv-mapped and fallback-uv
    meshes, with and without tangents: counts exact, sums to float64
    order."""
    desc = scenes.make_textured(width=96, height=72, spp=4)
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height) as integ:
        upload_scene(integ, desc)
        integ.render(desc.spp)
        sums = integ.readbackSums()
        cnt = integ.counters()
        tb = gpu_tables(integ, desc)
    ref = orc.Scene(desc, tables=tb)
    osums, ocnt = ref.render("f32", samples=(0, desc.spp))
    D = desc.max_depth + 1
    for k in ("rays", "hits", "shadow", "visible"):
        assert np.array_equal(cnt[k][:D], ocnt[k][:D]), (k, cnt[k][:D], ocnt[k][:D])
    rel = np.mean((sums - osums) ** 2 / (osums ** 2 + 1e-2))
    assert rel < 1e-12, rel
    assert sums.mean() > 0



def test_srgb_display_functions_within_4_ulp(c0_small):
    """convertFromLinearTosRGB / convertsRGBToLinear (math.h:114-140) on the
    GPU against the oracle: HDR values up to 1e4, the linear segment, the
    threshold, negatives and the decoder's clamps."""
    _, integ = c0_small
    rng = np.random.default_rng(31)
    x = np.concatenate([np.geomspace(1e-8, 1e4, 200001), rng.uniform(0, 1, 100000), -np.geomspace(1e-6, 1, 101),
                        [0.0, 0.0031308, np.nextafter(f32(0.0031308), f32(0)), 1.0]]).astype(f32)
    s = np.concatenate([np.linspace(-0.5, 1.5, 200001), [0.04045, np.nextafter(f32(0.04045), f32(1)), 0.0, 1.0]]
                       ).astype(f32)
    for vals, dec in ((x, 0), (s, 1)):
        out = np.zeros_like(vals)
        capi.check(capi.lib().clhip_srgb_batch(integ.ctx, capi.fptr(vals), len(vals), dec, capi.fptr(out)), integ.ctx)
        ref = orc.srgb(vals, decode=bool(dec))
        assert ulps(out, ref).max() <= 4, dec
    # 1.055f * 1 - 0.055f: the reference's float32 result at 1
    out = np.zeros(1, f32)
    capi.check(capi.lib().clhip_srgb_batch(integ.ctx, capi.fptr(np.ones(1, f32)), 1, 0, capi.fptr(out)), integ.ctx)
    assert out[0] == f32(0.99999994)


def test_u8_srgb_texture_equals_decoded_float_texture():
    """An 8-bit sRGB reflectance texture renders exactly as the float
    texture of its texels decoded (c / 255, then convertsRGBToLinear, per
    texel, before filtering) by the renderer's own decoder."""
    import dataclasses

    base = scenes.make_textured(width=96, height=64, spp=4)
    rng = np.random.default_rng(32)
    c8 = rng.integers(0, 256, size=(32, 32, 4), dtype=np.uint8)
    renders = []
    dec = None
    for kind in ("u8", "float"):
        with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, base.width, base.height) as integ:
            if kind == "u8":
                tex = ImageTexture2D(c8, isSRGB=True)
            else:
                v = (c8[..., :3].astype(f32) / f32(255)).ravel()
                d = np.zeros_like(v)
                capi.check(capi.lib().clhip_srgb_batch(integ.ctx, capi.fptr(v), len(v), 1, capi.fptr(d)), integ.ctx)
                dec = np.concatenate([d.reshape(32, 32, 3), c8[..., 3:].astype(f32) / f32(255)], -1)
                tex = ImageTexture2D(dec)
            mats = list(base.materials)
            mats[1] = DiffuseMtl(tex)
            desc = dataclasses.replace(base, materials=mats)
            upload_scene(integ, desc)
            integ.render(desc.spp)
            renders.append((integ.readbackSums(), integ.counters()))
    (sa, ca), (sb, cb) = renders
    for k in ("rays", "hits", "shadow", "visible"):
        assert np.array_equal(ca[k], cb[k]), k
    assert np.array_equal(sa, sb) or np.allclose(sa, sb, rtol=1e-12, atol=1e-12)
    assert np.allclose(dec[..., :3], orc.srgb(c8[..., :3].astype(f32) / f32(255), decode=True), rtol=5e-7, atol=1e-9)
