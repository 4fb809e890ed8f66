"""Benchmark: plain wavefront path tracing of the 1080p configuration on one
MI355X, reported in Msamples/s (one sample = one camera path traced to
termination), with the closest-hit traversal kernel's roofline and the CPU
oracle timed on the host cores beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one render of the full workload: 1920x1080, 64 spp, depth 8
(132.7 Msamples), inputs resident in HBM.  Prints one JSON line on stdout.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
QUEUE_BYTES_PER_RAY = 32 + 16 + 4  # ray o/d float4 x2 read, hit float4 + prim write
NODE_BYTES = 128
TRI_BYTES = 48


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(desc, seconds_target: float, threads: int):
    """Time the float32 oracle on a bounded sample of the same workload:
    full-width rows spread evenly over the frame (a stratified sample of the
    image, so sky and geometry are in proportion), 1 spp, sized to take about
    `seconds_target` seconds."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np
    import orc

    sc = orc.Scene(desc)
    w, h = desc.width, desc.height

    def rows_for(n):
        return np.floor((np.arange(n) + 0.5) * h / n).astype(np.uint32)

    n = 32
    t = time.perf_counter()
    sc.render_rows(rows_for(n), "f32", samples=(0, 1), nthreads=threads)
    dt = time.perf_counter() - t
    n = int(max(32, min(h, seconds_target * n / max(dt, 1e-9))))
    rows = rows_for(n)
    t = time.perf_counter()
    _, cnt = sc.render_rows(rows, "f32", samples=(0, 1), nthreads=threads)
    dt = time.perf_counter() - t
    samples = n * w
    return {"value": samples / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"{n} of {h} rows spread evenly over the {w}x{h} frame x 1 spp ({samples} camera paths, "
                      f"{int(cnt['rays'].sum())} closest-hit rays), depth {desc.max_depth}, float32 oracle, "
                      f"{threads} threads, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=1024, help="size of the streaming-copy probe")
    ap.add_argument("--copy-iters", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU sharding is described in "
              "DESIGN.md and not run.", file=sys.stderr)
        sys.exit(2)

    from colvillea_amd import Integrator, IntegratorType, scenes, upload_scene

    t0 = time.perf_counter()
    desc = scenes.make_c1(width=args.width, height=args.height, spp=args.spp)
    log(f"scene: {desc.num_triangles()} triangles, generated in {time.perf_counter() - t0:.1f} s")
    integ = Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height)
    t0 = time.perf_counter()
    upload_scene(integ, desc)
    integ.sync()
    log(f"upload + BVH + HDRI tables: {time.perf_counter() - t0:.1f} s")

    for i in range(args.warmup):
        integ.resetIterationIndex()
        integ.render(desc.spp)
        integ.sync()
        log(f"warmup {i + 1}/{args.warmup} done")

    integ.setTiming(True)
    trace_ms = shade_ms = shadow_ms = 0.0
    nodes = tris = launches = rays_traced = 0
    t_total = 0.0
    for i in range(args.steps):
        integ.resetIterationIndex()
        integ.sync()
        t = time.perf_counter()
        integ.render(desc.spp)
        integ.sync()
        dt = time.perf_counter() - t
        t_total += dt
        c = integ.counters()
        trace_ms += c["trace_ms"]
        shade_ms += c["shade_ms"]
        shadow_ms += c["shadow_ms"]
        nodes += c["trace_nodes"]
        tris += c["trace_tris"]
        launches += c["trace_launches"]
        rays_traced += int(c["rays"][1:desc.max_depth + 1].sum())
        log(f"step {i + 1}/{args.steps}: {dt * 1e3:.1f} ms, {desc.width * desc.height * desc.spp / dt / 1e6:.1f} "
            f"Msamples/s (trace {c['trace_ms']:.1f} ms, shade {c['shade_ms']:.1f} ms, shadow {c['shadow_ms']:.1f} ms)")
    samples = desc.width * desc.height * desc.spp * args.steps
    value = samples / t_total / 1e6
    algo_bytes = rays_traced * QUEUE_BYTES_PER_RAY + nodes * NODE_BYTES + tris * TRI_BYTES
    achieved = algo_bytes / (trace_ms * 1e-3) / 1e9 if trace_ms > 0 else 0.0
    per_launch = algo_bytes / max(1, launches)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_trace", "avg_launch_ms": round(trace_ms / max(1, launches), 4),
                "algorithmic_bytes_per_launch": int(per_launch),
                "nodes_per_ray": round(nodes / max(1, rays_traced), 2),
                "tris_per_ray": round(tris / max(1, rays_traced), 2)}
    import ctypes as C

    from colvillea_amd import _capi as capi

    copy_ms = C.c_double()
    copy_bytes = args.copy_mib << 20
    capi.check(capi.lib().clhip_copy_probe(integ.ctx, copy_bytes, args.copy_iters, C.byref(copy_ms)), integ.ctx)
    copy_gbs = 2 * copy_bytes / (copy_ms.value * 1e-3) / 1e9
    roofline["copy_ceiling_GBs"] = round(copy_gbs, 1)
    roofline["frac_of_copy_ceiling"] = round(achieved / copy_gbs, 4)
    log(f"copy probe: {args.copy_mib} MiB in {copy_ms.value:.3f} ms -> {copy_gbs:.0f} GB/s (read + write)")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(desc, args.cpu_seconds, args.cpu_threads)
        log(f"cpu baseline: {cpu}")
    out = {"metric": "Msamples/s at 1080p (plain wavefront PT)", "value": round(value, 2), "unit": "Msamples/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_total / args.steps * 1e3, 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"C1: {desc.width}x{desc.height}, {desc.spp} spp, depth {desc.max_depth}, "
                                  f"{desc.num_triangles()} triangles (fBm grid + {len(desc.instance_mesh) - 1} "
                                  f"instanced icospheres), 2048x1024 log-normal HDRI, seed {desc.info['seed']}",
                      "width": desc.width, "height": desc.height, "spp": desc.spp, "max_depth": desc.max_depth},
           "roofline": roofline, "cpu_baseline": cpu}
    integ.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
