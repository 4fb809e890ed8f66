"""Batched multi-view rendering, captured batch graphs and launch shapes.

Per-view films of one clhip_render_views call equal single-view renders of
the same views (same RNG lane, counter-based samples): queue counts exactly,
float64 sums up to summation order.  Graph replays equal plain launches call
by call.  No launch shape changes a result.
"""
import dataclasses

import numpy as np
import pytest

import orc
from colvillea_amd import Integrator, IntegratorType, scenes, upload_scene
from colvillea_amd.integrator import View

pytestmark = pytest.mark.gpu

DEPTH_KEYS = ("rays", "hits", "shadow", "visible")


@pytest.fixture(scope="module")
def c1_small():
    desc = scenes.make_c1(width=64, height=48, spp=4, grid=(96, 48), n_spheres=30, sphere_subdiv=2, hdri_w=128,
                          hdri_h=64)
    integ = Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height)
    upload_scene(integ, desc)
    yield desc, integ
    integ.close()


def small_views(n=10, seed=5):
    vs = scenes.make_c2_views(seed=seed, n_views=n, side=(16, 72), spp=(1, 5), depth=(2, 7))
    # a view with Russian roulette and one off the default lanes
    vs[1] = dataclasses.replace(vs[1], rr_depth=2)
    vs[2] = dataclasses.replace(vs[2], lane=4095)
    return vs


def single_view(integ, desc, v):
    integ.resize(v.width, v.height)
    integ.updateCamera(v.camera)
    integ.setParams(max_depth=v.max_depth, rr_depth=v.rr_depth, view=v.lane)
    integ.render(v.spp)
    return integ.readbackSums(), integ.counters()


def restore(integ, desc):
    integ.setLaunch()
    integ.resize(desc.width, desc.height)
    integ.updateCamera(desc.camera)
    integ.setParams(max_depth=desc.max_depth, rr_depth=desc.rr_depth, view=desc.view)


def close_sums(a, b):
    return np.allclose(a, b, rtol=1e-12, atol=1e-12)


def test_views_equal_single_view_renders(c1_small):
    desc, integ = c1_small
    views = small_views()
    # 4096-path batches: views share batches and large views span several
    integ.setLaunch(paths_per_batch=4096)
    integ.renderViews(views)
    multi = [integ.readbackViewSums(i) for i in range(len(views))]
    mc = integ.counters()
    assert mc["batches"] > 3
    D = max(v.max_depth for v in views) + 1
    tot = {k: np.zeros(D, np.uint64) for k in DEPTH_KEYS}
    try:
        for i, v in enumerate(views):
            s, c = single_view(integ, desc, v)
            assert close_sums(multi[i], s), i
            for k in DEPTH_KEYS:
                tot[k] += c[k][:D]
    finally:
        restore(integ, desc)
    for k in DEPTH_KEYS:
        assert np.array_equal(mc[k][:D], tot[k]), (k, mc[k][:D], tot[k])
    assert mc["rays"][1] == sum(v.width * v.height * v.spp for v in views)


def test_views_match_oracle_counts(c1_small):
    """Two views against the oracle's own render of each view: per-depth
    counts bit-exact, sums to float64 summation order."""
    desc, integ = c1_small
    views = small_views()[:3]
    integ.setLaunch(paths_per_batch=4096, order=2)
    try:
        integ.renderViews(views)
        got = [(integ.readbackViewSums(i)) for i in range(len(views))]
    finally:
        restore(integ, desc)
    from test_gpu_parity import gpu_tables
    tables = gpu_tables(integ, desc)
    for i in (1, 2):
        v = views[i]
        vd = dataclasses.replace(desc, width=v.width, height=v.height, camera=v.camera, max_depth=v.max_depth,
                                 rr_depth=v.rr_depth, view=v.lane, spp=v.spp)
        osums, _ = orc.Scene(vd, tables=tables).render("f32", samples=(0, v.spp))
        rel = np.mean((got[i] - osums) ** 2 / (osums ** 2 + 1e-2))
        assert rel < 1e-12, (i, rel)


def test_graph_replays_equal_plain_launches(c1_small):
    """Each render(1) call, replayed from the cached batch graph with
    readbacks and syncs between the calls, equals the plain launch of the
    same call; a multi-batch render replays the graph once per batch."""
    desc, integ = c1_small
    runs = {}
    try:
        for mode in (2, 1):
            integ.setLaunch(graph=mode)
            integ.resetIterationIndex()
            calls = []
            for _ in range(4):
                integ.render(1)
                c = integ.counters()
                calls.append((integ.readbackSums(), c))
            runs[mode] = calls
        plain, graph = runs[2], runs[1]
        for k in range(4):
            for key in DEPTH_KEYS:
                assert np.array_equal(plain[k][1][key], graph[k][1][key]), (k, key)
            assert close_sums(plain[k][0], graph[k][0]), k
            assert plain[k][1]["graph_replays"] == 0
            assert graph[k][1]["graph_replays"] == 1
        # captured at most once (an identical shape may already be cached),
        # then replayed from the cache
        assert graph[0][1]["graph_captures"] <= 1
        assert all(g[1]["graph_captures"] == 0 for g in graph[1:])
        # one call of 4 batches (samples_per_batch = 1)
        integ.setParams(max_depth=desc.max_depth, rr_depth=desc.rr_depth, view=desc.view, samples_per_batch=1)
        out = {}
        for mode in (2, 1):
            integ.setLaunch(graph=mode)
            integ.resetIterationIndex()
            integ.render(4)
            out[mode] = (integ.readbackSums(), integ.counters())
        assert out[1][1]["graph_replays"] == 4 and out[1][1]["batches"] == 4
        for key in DEPTH_KEYS:
            assert np.array_equal(out[1][1][key], out[2][1][key]), key
        assert close_sums(out[1][0], out[2][0])
        # the same 4 samples as the four progressive render(1) calls
        assert close_sums(out[1][0], plain[-1][0])
    finally:
        restore(integ, desc)


def test_launch_shape_does_not_change_results(c1_small):
    desc, integ = c1_small
    views = small_views(6, seed=9)
    shapes = [dict(), dict(trace_blocks_per_cu=1, shadow_blocks_per_cu=1, shade_blocks_per_cu=2),
              dict(paths_per_batch=5000, order=2, graph=2, tile=1), dict(shade_block=64, tile=8),
              dict(shade_block=128, tile=32, paths_per_batch=7000)]
    res = []
    try:
        for sh in shapes:
            integ.setLaunch(**sh)
            integ.renderViews(views)
            res.append(([integ.readbackViewSums(i) for i in range(len(views))], integ.counters()))
    finally:
        restore(integ, desc)
    for sums, cnt in res[1:]:
        for key in DEPTH_KEYS:
            assert np.array_equal(cnt[key], res[0][1][key]), key
        for a, b in zip(sums, res[0][0]):
            assert close_sums(a, b)


def test_view_latencies(c1_small):
    desc, integ = c1_small
    views = small_views(8, seed=13)
    try:
        integ.setLaunch(paths_per_batch=4096)
        integ.renderViews(views)
        lat = integ.viewLatencies()
        c = integ.counters()
    finally:
        restore(integ, desc)
    assert lat.shape == (8,)
    assert np.all(lat > 0) and np.all(lat <= c["render_ms"] + 1e-3)
    # views are laid out in the given order: completion times never decrease
    assert np.all(np.diff(lat) >= 0)


def test_launch_and_view_arguments_are_checked(c1_small):
    from colvillea_amd._capi import ClhipError
    desc, integ = c1_small
    with pytest.raises(ClhipError):
        integ.setLaunch(trace_blocks_per_cu=10_000)
    with pytest.raises(ClhipError):
        integ.setLaunch(paths_per_batch=100)
    with pytest.raises(ClhipError):
        integ.setLaunch(shade_block=96)
    with pytest.raises(ClhipError):
        integ.setLaunch(tile=12)
    with pytest.raises(ClhipError):
        integ.setLaunch(tile=64)
    with pytest.raises(ClhipError):
        integ.renderViews([View(8, 8, 1, 40, desc.camera)])  # depth above 32
    with pytest.raises(ClhipError):
        integ.renderViews([View(8, 8, 1, 4, desc.camera, lane=5000)])
    integ.setLaunch()
    got = integ.getLaunch()
    assert got["graph"] == 1 and got["trace_blocks_per_cu"] >= 1 and got["paths_per_batch"] >= 4096
    assert got["shade_block"] == 256 and got["tile"] == 16


def test_autotuner_logs_and_reapplies(c1_small, tmp_path):
    from colvillea_amd import autotune
    desc, integ = c1_small
    views = small_views(6, seed=21)
    space = {"trace_blocks_per_cu": [1, 2], "shade_blocks_per_cu": [2, 8], "shade_block": [256, 64],
             "paths_per_batch": [4096, 1 << 16], "tile": [1, 8], "order": [1, 2]}
    try:
        integ.setLaunch()
        base = integ.getLaunch()
        res = autotune.tune(integ, views, space=space, reps=1)
        for k, vals in space.items():
            assert res["choice"][k] in set(vals) | {base[k]}, k
        assert len(res["trials"]) >= 1 + sum(len(v) for v in space.values()) - len(space)
        assert all(t["ms"] is None or t["ms"] > 0 for t in res["trials"])
        path = tmp_path / "tune.json"
        autotune.save(res, str(path), workload="test views")
        integ.setLaunch()
        assert autotune.apply(integ, str(path)) == res["choice"]
        got = integ.getLaunch()
        for k in ("trace_blocks_per_cu", "shade_blocks_per_cu", "shade_block", "paths_per_batch", "tile", "order"):
            assert got[k] == res["choice"][k], k
    finally:
        restore(integ, desc)


def test_small_views_allocate_small_queues():
    """At the default launch (2^27 paths per batch) a call with a few small
    views sizes its queues by the batch it built, not by the launch's
    ceiling."""
    desc = scenes.make_c1(width=64, height=48, spp=1, grid=(96, 48), n_spheres=30, sphere_subdiv=2, hdri_w=128,
                          hdri_h=64)
    views = small_views(4, seed=9)
    with Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, 8, 8) as integ:
        upload_scene(integ, desc)
        assert integ.getLaunch()["paths_per_batch"] == 1 << 27
        integ.renderViews(views)
        c = integ.counters()
    need = sum(v.width * v.height * v.spp for v in views)
    assert c["batches"] == 1 and c["rays"][1] == need
    assert c["queue_paths"] == need
