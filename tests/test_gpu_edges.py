"""Edge cases of the render path on the GPU, each against the oracle on the
same seeded scene: ragged film sizes and a last batch that is only partly
full, a one-pixel film, depth 1, a scene without geometry, without
emitters, with only the sun or only the dome, and an empty mesh among
instanced ones.  Every case runs on both BVH widths (4: float boxes, 8:
byte-quantised boxes, whose empty slots are decoded differently).  Counts
are exact and float64 sums agree to summation order, as in the main parity
tests."""
import dataclasses

import numpy as np
import pytest

import orc
from colvillea_amd import Integrator, IntegratorType, scenes, upload_scene
from colvillea_amd import _capi as capi
from colvillea_amd.integrator import DirectionalEmitter, HDRIDome, TriangleMesh

pytestmark = pytest.mark.gpu

f32 = np.float32


def _c0(**kw):
    return scenes.make_c0(hdri_w=64, hdri_h=32, **kw)


def _cases():
    out = {}
    d = _c0(width=37, height=23, spp=5)
    d.samples_per_batch = 2  # batches of 2, 2 and 1 samples
    out["ragged_film_and_batches"] = d
    out["one_pixel"] = _c0(width=1, height=1, spp=7)
    d = _c0(width=24, height=16, spp=3)
    d.max_depth, d.rr_depth = 1, 0
    out["depth_1"] = d
    d = _c0(width=24, height=16, spp=3)
    out["no_geometry"] = dataclasses.replace(d, meshes=[], instance_mesh=[], instance_material=[],
                                             instance_xform=np.zeros((0, 3, 4), f32))
    out["no_emitters"] = dataclasses.replace(d, emitters=[])
    out["sun_only"] = dataclasses.replace(d, emitters=[e for e in d.emitters if isinstance(e, DirectionalEmitter)])
    out["dome_only"] = dataclasses.replace(d, emitters=[e for e in d.emitters if isinstance(e, HDRIDome)])
    empty = TriangleMesh(np.zeros((0, 3)), np.zeros((0, 3)))
    out["empty_mesh_instance"] = dataclasses.replace(
        d, meshes=d.meshes + [empty], instance_mesh=d.instance_mesh + [2],
        instance_material=d.instance_material + [0],
        instance_xform=np.concatenate([d.instance_xform, d.instance_xform[:1]]))
    return out


CASES = _cases()


def _gpu_tables(integ, desc):
    if desc.hdri is None:
        return None
    h, w = desc.hdri.shape[:2]
    pc = np.zeros((h, w + 1), f32)
    cc = np.zeros((h, w + 1), f32)
    pv = np.zeros(h + 1, f32)
    cv = np.zeros(h + 1, f32)
    capi.check(capi.lib().clhip_dome_tables(integ.ctx, capi.fptr(pc), capi.fptr(cc), capi.fptr(pv), capi.fptr(cv),
                                            None, None), integ.ctx)
    return pc, cc, pv, cv


def _render(desc, width):
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height) as integ:
        integ.setBVHWidth(width)
        upload_scene(integ, desc)
        integ.render(desc.spp)
        return integ.readbackSums(), integ.counters(), integ.readbackFramebufferLinear(), _gpu_tables(integ, desc)


@pytest.mark.parametrize("width", [4, 8], ids=["bvh4", "bvh8"])
@pytest.mark.parametrize("case", list(CASES))
def test_edge_case_matches_oracle(case, width):
    desc = CASES[case]
    print(f"edge case {case}, BVH{width}, {desc.width}x{desc.height}, {desc.spp} spp", flush=True)
    sums, cnt, img, tables = _render(desc, width)
    osums, ocnt = orc.Scene(desc, tables=tables).render("f32", samples=(0, desc.spp))
    D = desc.max_depth + 1
    for k in ("rays", "hits", "shadow", "visible"):
        assert np.array_equal(cnt[k][:D], ocnt[k][:D]), (k, cnt[k][:D], ocnt[k][:D])
        assert cnt[k][D:].sum() == 0, k
    assert cnt["rays"][1] == desc.width * desc.height * desc.spp
    assert np.all(np.isfinite(sums))
    if not desc.emitters:
        assert not sums.any()
    if not desc.instance_mesh:
        assert cnt["hits"].sum() == 0
    rel = np.mean((sums - osums) ** 2 / (osums ** 2 + 1e-2))
    assert rel < 1e-12, rel
    assert np.allclose(img[..., :3], (sums / desc.spp).astype(f32), rtol=1e-6, atol=1e-7)


def test_batching_does_not_change_the_image():
    """One sample per batch, two per batch (last one partial) and the
    default batching render the same samples: counts exact, sums to
    float64 summation order."""
    base = CASES["ragged_film_and_batches"]
    ref = None
    for spb in (1, 2, 0):
        desc = dataclasses.replace(base, samples_per_batch=spb)
        sums, cnt, _, _ = _render(desc, 4)
        if ref is None:
            ref = (sums, cnt)
            continue
        for k in ("rays", "hits", "shadow", "visible"):
            assert np.array_equal(cnt[k], ref[1][k]), (spb, k)
        assert np.allclose(sums, ref[0], rtol=1e-12, atol=1e-12), spb


def test_resize_then_render_equals_fresh_context():
    """resize() (integrator.h:83) drops the film; after the camera is set for
    the new size, the next render equals one from a new context."""
    big = _c0(width=40, height=30, spp=2)
    small = _c0(width=29, height=17, spp=3)
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, big.width, big.height) as integ:
        upload_scene(integ, big)
        integ.render(big.spp)
        integ.resize(small.width, small.height)
        assert integ.getFilmSize() == (small.width, small.height)
        assert integ.iterationIndex() == 0
        integ.updateCamera(small.camera)
        integ.render(small.spp)
        sums, cnt = integ.readbackSums(), integ.counters()
    fsums, fcnt, _, _ = _render(small, 4)
    assert sums.shape == (small.height, small.width, 3)
    for k in ("rays", "hits", "shadow", "visible"):
        assert np.array_equal(cnt[k], fcnt[k]), k
    assert np.allclose(sums, fsums, rtol=1e-12, atol=1e-12)


def test_render_before_scene_is_built_reports_state_error():
    """The reference renders whatever was built; the library names the
    missing build step instead, and launches nothing."""
    desc = _c0(width=8, height=8, spp=1)
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, 8, 8) as integ:
        with pytest.raises(capi.ClhipError, match="camera"):
            integ.render(1)
        integ.updateCamera(desc.camera)
        with pytest.raises(capi.ClhipError, match="TLAS"):
            integ.render(1)
        assert integ.iterationIndex() == 0


@pytest.mark.parametrize("width", [4, 8], ids=["bvh4", "bvh8"])
def test_rebuilding_one_blas_equals_a_full_build(width):
    """An edit to one mesh rebuilds that mesh's tree only (per-mesh BLAS
    validity, scene.cpp:186-210 / asdataset.h:16-124); the render then equals
    a fresh context given the edited scene, and the other trees were not
    built again."""
    import dataclasses

    from colvillea_amd.integrator import TriangleMesh

    desc = scenes.make_c0(width=48, height=32, spp=2, hdri_w=64, hdri_h=32)
    s = desc.meshes[1]
    squashed = TriangleMesh((s.vertices * np.array([1.0, 0.4, 1.0], np.float32)).astype(np.float32), s.triangles,
                            normals=s.normals)
    edited = dataclasses.replace(desc, meshes=[desc.meshes[0], squashed])
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height) as a:
        a.setBVHWidth(width)
        upload_scene(a, desc)
        before = a.counters()["blas_builds"]
        a.rebuildBLAS([1], [squashed])
        a.buildTLAS([edited.meshes[m] for m in edited.instance_mesh], list(range(len(edited.instance_mesh))),
                    edited.instance_xform)
        a.resetIterationIndex()
        a.render(desc.spp)
        ca, sa = a.counters(), a.readbackSums()
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height) as b:
        b.setBVHWidth(width)
        upload_scene(b, edited)
        b.render(desc.spp)
        cb, sb = b.counters(), b.readbackSums()
    assert before == 2 and ca["blas_builds"] == 3
    for k in ("rays", "hits", "shadow", "visible"):
        assert np.array_equal(ca[k], cb[k]), k
    assert np.allclose(sa, sb, rtol=1e-12, atol=1e-12)
