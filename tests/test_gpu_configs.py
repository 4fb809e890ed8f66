"""GPU parity on the workloads the benchmark measures, at their full scene
sizes (the per-function and small-scene parity lives in test_gpu_parity.py).

- 512 x 512 ground + 64 icospheres (81,922 triangles), 16 spp, depth 5, RR
  from 3: the whole frame against the oracle.
- 1920 x 1080 terrain scene (1024 x 512 fBm grid + 200 instanced
  subdivision-4 icospheres = 2,069,506 triangles, 2048 x 1024 HDRI with a
  10^4 x hot spot), depth 8: 256 x 256 crops at 4 spp against the oracle on
  both tree widths (per-depth counts exact, float64 sums to summation
  order), and the whole 64-spp frame in the default single batch of 2^27
  paths, checked for its camera-ray count, a finite image and the same
  per-depth counts as a render of 16 spp per batch.
- 256 views of that scene (square sides log-uniform 64..512, spp U{4..32},
  depth U{3..10}) in one batched render: per-view films finite and
  non-negative, total camera rays, and four views against the oracle.
- The window-lit 10 x 10 x 4 room (247,808 triangles), depth 10: a 256 x 256
  crop at 4 spp against the oracle, and the learned proposal's estimator
  on the full room geometry (untrained, and trained then frozen) against
  the plain tracer's mean.

The reference bounce loop these cover is WavefrontPathTracingIntegrator::
render (src/core/integrators/path.cpp:84-231).
"""
import dataclasses
import sys
import time

import numpy as np
import pytest

import orc
from colvillea_amd import Integrator, IntegratorType, scenes, upload_scene

pytestmark = pytest.mark.gpu

KEYS = ("rays", "hits", "shadow", "visible")


def note(msg):
    print(msg, file=sys.stderr, flush=True)


def gpu_tables(integ, desc):
    from colvillea_amd import _capi as capi

    h, w = desc.hdri.shape[:2]
    pc = np.zeros((h, w + 1), np.float32)
    cc = np.zeros((h, w + 1), np.float32)
    pv = np.zeros(h + 1, np.float32)
    cv = np.zeros(h + 1, np.float32)
    capi.check(capi.lib().clhip_dome_tables(integ.ctx, capi.fptr(pc), capi.fptr(cc), capi.fptr(pv), capi.fptr(cv),
                                            None, None), integ.ctx)
    return pc, cc, pv, cv


def relmse(a, b):
    return float(np.mean((a - b) ** 2 / (b ** 2 + 1e-2)))


def assert_counts_equal(got, ref, depth):
    D = depth + 1
    for k in KEYS:
        assert np.array_equal(got[k][:D], ref[k][:D]), (k, got[k][:D], ref[k][:D])


def render_vs_oracle(desc, spp, width=4):
    """GPU render of desc at `spp` and the float32 oracle's render of the same
    samples on the GPU's own dome tables."""
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height) as integ:
        integ.setBVHWidth(width)
        upload_scene(integ, desc)
        integ.render(spp)
        sums = integ.readbackSums()
        cnt = integ.counters()
        tables = gpu_tables(integ, desc)
    t = time.perf_counter()
    osums, ocnt = orc.Scene(desc, tables=tables).render("f32", samples=(0, spp))
    note(f"{desc.name} {desc.width}x{desc.height}x{spp}: oracle {time.perf_counter() - t:.1f} s, "
         f"rays {cnt['rays'][1:desc.max_depth + 1].tolist()}")
    return sums, cnt, osums, ocnt


# ------------------------------------------------------------ 512 x 512

def test_c0_full_frame_matches_oracle():
    desc = scenes.make_c0()
    note(f"C0 full frame {desc.width}x{desc.height}, {desc.spp} spp, depth {desc.max_depth}, RR from {desc.rr_depth}")
    sums, cnt, osums, ocnt = render_vs_oracle(desc, desc.spp)
    assert cnt["rays"][1] == desc.width * desc.height * desc.spp == 4194304
    assert_counts_equal(cnt, ocnt, desc.max_depth)
    assert relmse(sums, osums) < 1e-12


# ------------------------------------------------------------ 1920 x 1080 terrain

@pytest.fixture(scope="module")
def c1():
    t = time.perf_counter()
    desc = scenes.make_c1()
    assert desc.num_triangles() == 2069506
    assert desc.hdri.shape[:2] == (1024, 2048)
    note(f"C1: {desc.num_triangles()} triangles generated in {time.perf_counter() - t:.1f} s")
    return desc


# the film centre (terrain, spheres of all three materials) and the upper
# left (horizon and sky: escaped rays and the hot spot's MIS)
C1_CROPS = {"centre": (832, 412, 256, 256), "horizon": (160, 40, 256, 256)}


@pytest.mark.parametrize("width", [4, 8], ids=["bvh4", "bvh8"])
@pytest.mark.parametrize("where", sorted(C1_CROPS))
def test_c1_crop_matches_oracle(c1, where, width):
    desc = scenes.crop(c1, *C1_CROPS[where], spp=4)
    note(f"C1 crop {where} {C1_CROPS[where]} at 4 spp, BVH{width}")
    sums, cnt, osums, ocnt = render_vs_oracle(desc, 4, width)
    assert_counts_equal(cnt, ocnt, desc.max_depth)
    assert cnt["rays"][desc.max_depth] > 0  # paths reach the last depth
    assert relmse(sums, osums) < 1e-12


def test_c1_full_frame_single_batch(c1):
    """The benchmark's own render: 1920 x 1080 x 64 spp in the default
    batch (2^27 paths, so the whole frame is one batch), against the same
    samples rendered 16 spp per batch."""
    note("C1 full frame 1920x1080x64 in the default batch")
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, c1.width, c1.height) as integ:
        upload_scene(integ, c1)
        assert integ.getLaunch()["paths_per_batch"] == 1 << 27
        integ.render(c1.spp)
        a = integ.counters()
        img = integ.readbackFramebufferLinear()
        sa = integ.readbackSums()
        note(f"default batch: {a['batches']} batch(es), rays {a['rays'][1:9].tolist()}")
        integ.setParams(max_depth=c1.max_depth, rr_depth=c1.rr_depth, view=0, samples_per_batch=16)
        integ.resetIterationIndex()
        integ.render(c1.spp)
        b = integ.counters()
        sb = integ.readbackSums()
    assert a["batches"] == 1 and b["batches"] == 4
    assert a["rays"][1] == 1920 * 1080 * 64 == 132710400
    assert_counts_equal(a, b, c1.max_depth)
    assert a["rays"][c1.max_depth] > 0
    assert np.isfinite(img).all() and (img[..., :3] >= 0).all()
    assert img[..., :3].mean() > 0
    assert np.allclose(sa, sb, rtol=1e-9, atol=1e-9)


# ------------------------------------------------------------ 256 views

# the three cheapest views (depths 10, 9 and 3) and one of the largest
# (482 x 482, 31 spp, depth 9)
C2_ORACLE_VIEWS = (59, 42, 116, 213)


def test_c2_views_match_oracle(c1):
    views = scenes.make_c2_views()
    assert len(views) == 256
    nsamp = sum(v.width * v.height * v.spp for v in views)
    note(f"C2: 256 views, {nsamp} samples, one batched render at the default launch")
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, c1.width, c1.height) as integ:
        upload_scene(integ, c1)
        integ.renderViews(views)
        mc = integ.counters()
        films = [integ.readbackViewSums(i) for i in range(len(views))]
        lat = integ.viewLatencies()
        tables = gpu_tables(integ, c1)
        single = {}
        for i in C2_ORACLE_VIEWS:
            v = views[i]
            integ.resize(v.width, v.height)
            integ.updateCamera(v.camera)
            integ.setParams(max_depth=v.max_depth, rr_depth=v.rr_depth, view=v.lane)
            integ.render(v.spp)
            single[i] = (integ.readbackSums(), integ.counters())
    assert mc["rays"][1] == nsamp
    assert np.isfinite(lat).all() and (lat > 0).all()
    for i, (v, f) in enumerate(zip(views, films)):
        assert f.shape == (v.height, v.width, 3), i
        assert np.isfinite(f).all() and (f >= 0).all(), i
        assert f.mean() > 0, i
    for i in C2_ORACLE_VIEWS:
        v = views[i]
        vd = dataclasses.replace(c1, width=v.width, height=v.height, camera=v.camera, max_depth=v.max_depth,
                                 rr_depth=v.rr_depth, view=v.lane, spp=v.spp)
        t = time.perf_counter()
        osums, ocnt = orc.Scene(vd, tables=tables).render("f32", samples=(0, v.spp))
        note(f"view {i} ({v.width}x{v.height}x{v.spp}, depth {v.max_depth}): oracle {time.perf_counter() - t:.1f} s")
        ssum, scnt = single[i]
        assert_counts_equal(scnt, ocnt, v.max_depth)
        assert relmse(films[i], osums) < 1e-12, i
        assert relmse(ssum, osums) < 1e-12, i


# ------------------------------------------------------------ window-lit room

@pytest.fixture(scope="module")
def c3():
    desc = scenes.make_c3()
    assert desc.num_triangles() == 247808
    return desc


def test_c3_crop_matches_oracle(c3):
    desc = scenes.crop(c3, 832, 412, 256, 256, spp=4)
    note("C3 crop (832, 412, 256, 256) at 4 spp, depth 10")
    sums, cnt, osums, ocnt = render_vs_oracle(desc, 4)
    assert_counts_equal(cnt, ocnt, desc.max_depth)
    assert cnt["rays"][desc.max_depth] > 0
    assert relmse(sums, osums) < 1e-12


def chunk_means(it, chunks, spp):
    it.resetIterationIndex()
    means, prev = [], None
    for _ in range(chunks):
        it.render(spp)
        s = it.readbackSums()
        cur = s if prev is None else s - prev
        prev = s
        means.append(cur.mean() / spp)
    return np.array(means)


@pytest.mark.parametrize("trained", [False, True], ids=["untrained", "trained"])
def test_c3_guided_estimator_is_unbiased(c3, trained):
    """The full room (every wall cell, 50 props, the 2048 x 1024 HDRI) on a
    160 x 90 film: the guided estimator's frame mean over 16 independent
    chunks equals the plain tracer's mean over 16 chunks of 256 spp within
    4 standard errors of the difference, with a random network and with one
    trained online and then frozen."""
    desc = dataclasses.replace(c3, width=160, height=90)
    note(f"C3 guided unbiasedness, {'trained' if trained else 'untrained'} network")
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height) as it:
        upload_scene(it, desc)
        it.setParams(max_depth=desc.max_depth, rr_depth=desc.rr_depth, view=2)
        pm = chunk_means(it, 16, 256)
        it.guideReset(21)
        if trained:
            it.setParams(max_depth=desc.max_depth, rr_depth=desc.rr_depth, view=0, samples_per_batch=1)
            it.setGuide(True, bsdf_fraction=0.5, train=1, record_paths=desc.width * desc.height, batch_size=8192,
                        learning_rate=3e-3, steps_per_batch=2)
            it.resetIterationIndex()
            it.render(128)
            st = it.guideStats()
            assert st["steps"] >= 128 and np.isfinite(st["last_loss"])
        it.setParams(max_depth=desc.max_depth, rr_depth=desc.rr_depth, view=1)
        it.setGuide(True, bsdf_fraction=0.5, train=0)
        try:
            gm = chunk_means(it, 16, 64)
        finally:
            it.setGuide(False)
    se = np.sqrt(gm.var(ddof=1) / len(gm) + pm.var(ddof=1) / len(pm))
    note(f"guided mean {gm.mean():.6g}, plain mean {pm.mean():.6g}, standard error of the difference {se:.3g}")
    assert abs(gm.mean() - pm.mean()) < 4 * se, (gm.mean(), pm.mean(), se)


def test_c3_gpu_reference_consistent_with_f64_oracle(c3):
    """The error reference of the guided-vs-plain measurement is the plain
    GPU tracer at high spp.  Its expectation equals the float64 oracle's
    (libm elementary functions, float64 HDRI tables of its own): the full
    room on a 96 x 54 film, depth 10, the oracle at 8 chunks of 32 spp
    against the GPU at 8 chunks of 2048 spp (independent RNG lanes).  The
    frame means agree within 4 standard errors (about 0.8% of the mean per
    standard error), and the 72 block means (8 x 9 pixels) have a mean
    squared t-statistic below 2.5 (about 1.4 expected for 7 degrees of
    freedom per side)."""
    desc = dataclasses.replace(c3, width=96, height=54)
    t = time.perf_counter()
    sc = orc.Scene(desc)
    ref = np.array([sc.render("f64", samples=(32 * k, 32 * k + 32))[0] / 32 for k in range(8)])
    note(f"float64 oracle: 96x54 x 256 spp in {time.perf_counter() - t:.1f} s")
    gpu = []
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height) as it:
        upload_scene(it, desc)
        for lane in range(2, 10):
            it.setParams(max_depth=desc.max_depth, rr_depth=desc.rr_depth, view=lane)
            it.render(2048)
            gpu.append(it.readbackSums() / 2048)
    gpu = np.array(gpu)

    def stats(ch):
        frame = ch.sum(-1).mean(axis=(1, 2))
        blk = ch.sum(-1).reshape(len(ch), 6, 9, 12, 8).mean(axis=(2, 4))
        return frame.mean(), frame.std(ddof=1) / np.sqrt(len(ch)), blk.mean(0), blk.std(0, ddof=1) / np.sqrt(len(ch))

    fr, fse, br, bse = stats(ref)
    fg, gse, bg, bgse = stats(gpu)
    z = (fr - fg) / np.hypot(fse, gse)
    tb = (br - bg) / np.sqrt(bse ** 2 + bgse ** 2 + 1e-30)
    note(f"frame mean: oracle {fr:.5g} +- {fse:.2g}, GPU {fg:.5g} +- {gse:.2g} (z = {z:.2f}); "
         f"blocks: mean t^2 {np.mean(tb ** 2):.2f}, max |t| {np.abs(tb).max():.2f}")
    assert abs(z) < 4.0
    assert np.mean(tb ** 2) < 2.5 and np.abs(tb).max() < 8.0
