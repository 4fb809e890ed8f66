"""Launch-shape autotuner for the batched bounce loop.

The tuned fields are those of clhip_launch (include/clhip.h): persistent
grid sizes of the closest-hit and any-hit kernels, the grid and the block
size of the camera and shading kernels, the queue chunk (paths per batch),
the pixel tile that orders a sample's paths (rows or square tiles) and the
order in which views fill the batches.  No field changes a rendered value, only the
time, so the search may use a bounded sample of the workload.

The search is coordinate descent in a fixed field order starting from the
library defaults: every candidate value of one field is timed with the
others held, the fastest is kept (a candidate must beat the incumbent by
`margin` to replace it, so ties keep the earlier value), then the next
field.  Times are device-event milliseconds of whole renders (the minimum
of `reps` runs after one warm-up).  Every trial is logged; `save` writes
the log as JSON and `apply` sets a logged choice again, so a tuned shape is
reproduced without re-timing.
"""
from __future__ import annotations

import dataclasses
import json
from typing import Dict, List, Optional, Sequence

from .integrator import Integrator, View

FIELDS = ("trace_blocks_per_cu", "shadow_blocks_per_cu", "shade_blocks_per_cu", "shade_block", "paths_per_batch", "tile",
          "order")


def default_space(integ: Integrator) -> Dict[str, List[int]]:
    d = integ.getLaunch()
    t = d["trace_blocks_per_cu"]
    return {
        "trace_blocks_per_cu": sorted({max(1, t), max(1, t - 1), max(1, t - 2), max(1, t // 2)}, reverse=True),
        "shadow_blocks_per_cu": sorted({x for x in (max(1, d["shadow_blocks_per_cu"]), 3, 4, 6) if x <= max(1, d["shadow_blocks_per_cu"])}, reverse=True),
        "shade_blocks_per_cu": [d["shade_blocks_per_cu"], 4, 16],
        "shade_block": [d["shade_block"], 128, 64],
        "tile": [d["tile"], 1, 8, 32],
        "paths_per_batch": [d["paths_per_batch"], 16 << 20, 32 << 20, 64 << 20],
        "order": [1, 2],
    }


def sample_views(views: Sequence[View], spp_div: int = 4) -> List[View]:
    """The workload's views at 1/spp_div of their samples (at least one)."""
    return [dataclasses.replace(v, spp=max(1, v.spp // spp_div)) for v in views]


def _time(integ: Integrator, views: Sequence[View], reps: int) -> float:
    best = float("inf")
    for _ in range(reps):
        integ.renderViews(views)
        best = min(best, float(integ.counters()["render_ms"]))
    return best


def tune(integ: Integrator, views: Sequence[View], space: Optional[Dict[str, List[int]]] = None, reps: int = 2,
         margin: float = 0.01) -> dict:
    """Returns {"choice": {field: value}, "trials": [...], "ms": best}.
    Candidates the library rejects (e.g. more blocks than fit on a CU) are
    logged as skipped.  Leaves `choice` set on the integrator."""
    space = space or default_space(integ)
    base = integ.getLaunch()
    cur = {k: base[k] for k in FIELDS}
    cur["graph"] = 1
    trials = []
    integ.setLaunch(**cur)
    integ.renderViews(views)  # warm-up: buffers, graph capture
    best_ms = _time(integ, views, reps)
    trials.append({"launch": dict(cur), "ms": best_ms, "field": None})
    for field in FIELDS:
        for value in space.get(field, []):
            if value == cur[field]:
                continue
            cand = dict(cur, **{field: value})
            try:
                integ.setLaunch(**cand)
            except Exception as e:  # rejected shape
                trials.append({"launch": cand, "ms": None, "field": field, "skipped": str(e)})
                continue
            integ.renderViews(views)  # warm-up for this shape
            ms = _time(integ, views, reps)
            trials.append({"launch": cand, "ms": ms, "field": field})
            if ms < best_ms * (1.0 - margin):
                best_ms, cur = ms, cand
        integ.setLaunch(**cur)
    integ.setLaunch(**cur)
    return {"choice": dict(cur), "ms": best_ms, "trials": trials, "reps": reps, "margin": margin}


def save(result: dict, path: str, **extra) -> None:
    with open(path, "w") as f:
        json.dump(dict(result, **extra), f, indent=1)


def apply(integ: Integrator, path_or_result) -> dict:
    """Sets the choice of a saved (or returned) tuning result."""
    res = path_or_result
    if isinstance(res, str):
        with open(res) as f:
            res = json.load(f)
    integ.setLaunch(**res["choice"])
    return res["choice"]
