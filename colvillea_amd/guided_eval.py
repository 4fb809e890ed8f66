"""Time to equal error: the guided tracer (online training included) against
the plain tracer on one scene.

Both tracers render the same view in chunks; after each chunk the
accumulated image is compared with a reference.  The error is the relative
MSE  mean over pixels and channels of (I - R)^2 / (R^2 + 1e-2), corrected
for the reference's own noise: the reference is the mean of two independent
halves R1, R2 (different RNG view lanes), whose per-pixel variance of the
mean is estimated by ((R1 - R2) / 2)^2 and subtracted in the same relative
form.  The reference comes from the plain tracer (its agreement with the
float64 oracle is tested separately at small sizes).

Time is wall-clock around render + synchronise of each chunk (readbacks
excluded), so the guided tracer's time includes its network inference and
every training step.  The plain tracer runs at the library's default batch
size (its fastest); the guided tracer at the schedule's `samples_per_batch`
(it trains after every batch).  With `warm_spp`, the guided tracer first renders and
trains on that many spp (with `warm_steps_per_batch` Adam steps per batch,
if given) and discards their film (their time is counted).  T_plain(eps) is the plain tracer's time to its final
error eps; T_guided(eps) is where the guided tracer's error curve crosses
eps (log-log interpolation between chunks).
"""
from __future__ import annotations

import math
import time
from typing import List, Optional

import numpy as np

from .integrator import Integrator, IntegratorType
from .scenes import SceneDesc
from .integrator import upload_scene

EPS_DEN = 1e-2


def _render_curve(it: Integrator, spp_points: List[int], ref, ref_var, warm_spp: int = 0,
                  warm_guide: Optional[dict] = None, guide: Optional[dict] = None, desc=None) -> list:
    """Error after each chunk.  warm_spp > 0: first render that many spp
    (the guided tracer trains on them, with the settings `warm_guide` if
    given, then `guide`), then discard that film; the time is counted.
    With `guide` and `desc`, once training has stopped (max_steps issued)
    the remaining chunks use the library's default batch size, as the plain
    tracer does.  Guided points also carry the device time of the Adam
    steps up to them (train_seconds)."""
    it.resetIterationIndex()
    it.sync()
    done, t_acc, train_acc, curve = 0, 0.0, 0.0, []
    if warm_spp:
        t = time.perf_counter()
        if warm_guide:
            it.setGuide(True, **warm_guide)
        it.render(warm_spp)
        if guide is not None:
            train_acc += it.guideStats()["train_ms"] * 1e-3
        if warm_guide:
            it.setGuide(True, **guide)
        it.resetIterationIndex()
        it.sync()
        t_acc += time.perf_counter() - t
    fast = guide is None
    for s in spp_points:
        if not fast and desc is not None and not it.guideStats()["training"]:
            it.setParams(max_depth=desc.max_depth, rr_depth=desc.rr_depth, view=0, samples_per_batch=0)
            fast = True
        t = time.perf_counter()
        it.render(s - done)
        it.sync()
        t_acc += time.perf_counter() - t
        if guide is not None:
            train_acc += it.guideStats()["train_ms"] * 1e-3
        done = s
        img = it.readbackSums() / s
        den = ref ** 2 + EPS_DEN
        e = float(np.mean((img - ref) ** 2 / den) - np.mean(ref_var / den))
        point = {"spp": s, "seconds": t_acc, "relmse": e}
        if guide is not None:
            point["train_seconds"] = train_acc  # device time of the Adam steps so far
        curve.append(point)
    return curve


def _crossing(curve: list, eps: float) -> Optional[float]:
    prev = None
    for p in curve:
        if p["relmse"] <= eps:
            if prev is None or prev["relmse"] <= 0 or p["relmse"] <= 0:
                return p["seconds"]
            # log-log interpolation of error against time
            x0, x1 = math.log(prev["seconds"]), math.log(p["seconds"])
            y0, y1 = math.log(prev["relmse"]), math.log(p["relmse"])
            y = math.log(eps)
            return math.exp(x0 + (y - y0) * (x1 - x0) / (y1 - y0)) if y1 != y0 else p["seconds"]
        prev = p
    return None


def reference(it: Integrator, desc: SceneDesc, ref_spp: int, log=print):
    """Plain-tracer reference: the mean of two halves of ref_spp (RNG lanes
    2 and 3) and the per-pixel variance of that mean."""
    t0 = time.perf_counter()
    halves = []
    for lane in (2, 3):
        it.setParams(max_depth=desc.max_depth, rr_depth=desc.rr_depth, view=lane)
        it.render(ref_spp // 2)
        halves.append(it.readbackSums() / (ref_spp // 2))
    log(f"reference: {ref_spp} spp in {time.perf_counter() - t0:.1f} s")
    return 0.5 * (halves[0] + halves[1]), ((halves[0] - halves[1]) * 0.5) ** 2


def time_to_equal_error(desc: SceneDesc, ref_spp: int, spp_points: List[int], guide: dict,
                        it: Optional[Integrator] = None, log=print, ref=None, seeds=None, plain=None) -> dict:
    """ref: a (reference, variance) pair from reference() to reuse.
    seeds: network initialisations to run the guided tracer from (default
    guide["seed"] or 1); the ratio is reported per seed and as their mean.
    plain: a plain curve from an earlier call on the same scene and points
    (the plain tracer is deterministic), else it is rendered here."""
    own = it is None
    if own:
        it = Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height)
        upload_scene(it, desc)
    try:
        ref, ref_var = ref if ref is not None else reference(it, desc, ref_spp, log)
        if plain is None:
            # the plain tracer at its own fastest setting: the library's batch size
            it.setParams(max_depth=desc.max_depth, rr_depth=desc.rr_depth, view=0, samples_per_batch=0)
            it.setGuide(False)
            plain = _render_curve(it, spp_points, ref, ref_var)
            log(f"plain: {plain[-1]}")
        eps = plain[-1]["relmse"]
        tp = plain[-1]["seconds"]
        # the guided tracer trains once per batch: its batch size is part of the schedule
        it.setParams(max_depth=desc.max_depth, rr_depth=desc.rr_depth, view=0,
                     samples_per_batch=int(guide.get("samples_per_batch", 1)))
        skip = ("seed", "warm_spp", "warm_steps_per_batch", "samples_per_batch")
        g = {k: v for k, v in guide.items() if k not in skip}
        wg = dict(g, steps_per_batch=guide["warm_steps_per_batch"]) if guide.get("warm_steps_per_batch") else None
        runs = []
        for seed in (seeds or [guide.get("seed", 1)]):
            it.guideReset(seed)
            it.setGuide(True, **g)
            it.setParams(max_depth=desc.max_depth, rr_depth=desc.rr_depth, view=0,
                         samples_per_batch=int(guide.get("samples_per_batch", 1)))
            guided = _render_curve(it, spp_points, ref, ref_var, guide.get("warm_spp", 0), wg, g, desc)
            gst = it.guideStats()
            log(f"guided (seed {seed}): {guided[-1]}, {gst['steps']} training steps")
            it.setGuide(False)
            tg = _crossing(guided, eps)
            runs.append({"seed": seed, "guided": guided, "t_guided": tg, "ratio": (tg / tp) if tg is not None else None,
                         "train_steps": int(gst["steps"]), "samples_trained": int(gst["samples"])})
        ratios = [r["ratio"] for r in runs]
        ok = all(r is not None for r in ratios)
        first = runs[0]
        return {"plain": plain, "guided": first["guided"], "eps": eps, "t_plain": tp, "t_guided": first["t_guided"],
                "ratio": sum(ratios) / len(ratios) if ok else None, "ratios": ratios, "runs": runs,
                "ref_spp": ref_spp, "guide": dict(guide), "train_steps": first["train_steps"],
                "samples_trained": first["samples_trained"]}
    finally:
        it.setParams(max_depth=desc.max_depth, rr_depth=desc.rr_depth, view=desc.view)
        if own:
            it.close()
