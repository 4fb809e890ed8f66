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
# MI355X L2: 8 XCDs x 16 channels x 128 B per clock at 2.4 GHz
L2_PEAK_GBS = 8 * 16 * 128 * 2.4
QUEUE_BYTES_PER_RAY = 32 + 16  # ray o/d float4 x2 read, hit float4 (b1, b2, instance, primitive) written
ORIGIN_BYTES = 16  # a camera ray of a one-view render reads no origin: it is the camera position
NODE_BYTES = {4: 128, 8: 96}  # Node4: float child boxes; Node8: byte-quantised child boxes
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

    t = time.perf_counter()
    orc.use_native()  # built here, for this host's CPU: -O3 -march=native
    log(f"cpu baseline: oracle built for this host in {time.perf_counter() - t:.1f} s")
    sc = orc.Scene(desc)
    w, h = desc.width, desc.height

    def rows_for(n):
        return np.floor((np.arange(n) + 0.5) * h / n).astype(np.uint32)

    n = 32
    t = time.perf_counter()
    sc.render_rows(rows_for(n), "f32", samples=(0, 1), nthreads=threads)
    dt = time.perf_counter() - t
    full = dt * h / n  # estimated seconds for the whole frame at 1 spp
    if full >= seconds_target:
        n, k = int(max(32, seconds_target / full * h)), 1
    else:  # whole frame, the first k samples of every pixel
        n, k = h, int(max(1, min(desc.spp, round(seconds_target / full))))
    rows = rows_for(n)
    t = time.perf_counter()
    _, cnt = sc.render_rows(rows, "f32", samples=(0, k), nthreads=threads)
    dt = time.perf_counter() - t
    samples = n * w * k
    return {"value": samples / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "host": host_cpu(), "build": "gcc -O3 -march=native -fno-fast-math -ffp-contract=off",
            "sample": f"{n} of {h} rows spread evenly over the {w}x{h} frame x {k} spp ({samples} camera paths, "
                      f"{int(cnt['rays'].sum())} closest-hit rays), depth {desc.max_depth}, float32 oracle "
                      f"(BVH2 traversal), {threads} threads, {dt:.1f} s"}


def host_cpu() -> dict:
    """The host CPU the baseline ran on: lscpu's model, sockets, cores per
    socket, threads per core and logical CPUs, and the CPUs this process may
    run on."""
    import subprocess

    out = {"cpus_available": len(os.sched_getaffinity(0))}
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "logical_cpus"}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        text = ""
    for line in text.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys and keys[k.strip()] not in out:
            v = v.strip()
            out[keys[k.strip()]] = int(v) if v.isdigit() else v
    return out


def other_width(desc, args) -> dict:
    """The same 1080p render on the other tree width (BVH8 when the main
    line ran BVH4): one warm-up, the best of two timed renders, and the
    closest-hit kernel's event-timed launches."""
    from colvillea_amd import Integrator, IntegratorType, upload_scene

    width = 8 if args.bvh_width == 4 else 4
    it = Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height)
    try:
        it.setBVHWidth(width)
        upload_scene(it, desc)
        it.render(desc.spp)
        it.sync()
        best = None
        for _ in range(2):
            it.resetIterationIndex()
            it.sync()
            t = time.perf_counter()
            it.render(desc.spp)
            it.sync()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        it.setTiming(True)
        it.resetIterationIndex()
        it.render(desc.spp)
        it.sync()
        c = it.counters()
        it.setTiming(False)
        rays = int(c["rays"][1:desc.max_depth + 1].sum())
        out = {"bvh_width": width, "value": round(desc.width * desc.height * desc.spp / best / 1e6, 2),
               "unit": "Msamples/s", "k_trace_ms_per_frame": round(c["trace_ms"], 2),
               "nodes_per_ray": round(c["trace_nodes"] / max(1, rays), 2),
               "tris_per_ray": round(c["trace_tris"] / max(1, rays), 2)}
        log(f"BVH{width}: {out}")
        return out
    finally:
        it.close()


def batched_views(integ, args) -> dict:
    """256 views of the same scene (square films, side log-uniform in
    [64, 512], spp U{4..32}, depth U{3..10}) rendered by one
    clhip_render_views per step, launch shape from the autotuner."""
    import numpy as np

    from colvillea_amd import autotune, scenes

    views = scenes.make_c2_views()
    nsamp = sum(v.width * v.height * v.spp for v in views)
    tune = None
    if not args.no_autotune:
        t = time.perf_counter()
        tune = autotune.tune(integ, autotune.sample_views(views), reps=2)
        log(f"autotune: {len(tune['trials'])} trials in {time.perf_counter() - t:.1f} s -> {tune['choice']}")
        if args.autotune_log:
            autotune.save(tune, args.autotune_log, workload="256 views, sample = spp // 4")
    integ.renderViews(views)
    integ.sync()
    log("batched warm-up done")
    t_total, dev_ms = 0.0, 0.0
    for i in range(args.c2_steps):
        t = time.perf_counter()
        integ.renderViews(views)
        integ.sync()
        dt = time.perf_counter() - t
        t_total += dt
        c = integ.counters()
        dev_ms += c["render_ms"]
        log(f"batched step {i + 1}/{args.c2_steps}: {dt * 1e3:.1f} ms, {nsamp / dt / 1e6:.1f} Msamples/s, "
            f"{c['batches']} batches, {c['graph_replays']} graph replays")
    lat = integ.viewLatencies()
    edges = [0, 25, 50, 100, 200, 400, 800, 1600, 3200, 1e9]
    hist = np.histogram(lat, bins=edges)[0].tolist()
    return {"workload": f"256 views of the 1080p scene, {nsamp} samples, square sides log-uniform 64..512, "
                        f"spp U{{4..32}}, depth U{{3..10}}, seed 3",
            "value": round(nsamp * args.c2_steps / t_total / 1e6, 2), "unit": "Msamples/s",
            "ms_per_step": round(t_total / args.c2_steps * 1e3, 2),
            "device_ms_per_step": round(dev_ms / args.c2_steps, 2), "steps": args.c2_steps,
            "batches": int(c["batches"]), "graph_replays": int(c["graph_replays"]),
            "launch": integ.getLaunch(),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                           "p95": round(float(np.percentile(lat, 95)), 2),
                           "p99": round(float(np.percentile(lat, 99)), 2), "max": round(float(lat.max()), 2)},
            "latency_histogram": {"edges_ms": edges[:-1] + ["inf"], "counts": hist},
            "autotune": None if tune is None else {
                "trials": len(tune["trials"]), "best_sample_ms": tune["ms"],
                "searched": [[t["field"], None if t["field"] is None else t["launch"][t["field"]],
                              None if t["ms"] is None else round(t["ms"], 2)] for t in tune["trials"]]}}


# The configuration's learned-proposal hyperparameters (Adam lr 1e-3, 2^16
# samples per step) with the schedule around them; this is the headline
# guided line.  The schedule -- recorded paths, steps per batch, a discarded
# warm-up, the freeze point -- was picked by GPU sweeps over three network
# initialisations at these fixed hyperparameters.
GUIDE_SPEC = {"bsdf_fraction": 0.5, "train": 1, "record_paths": 1 << 21, "batch_size": 1 << 16,
              "learning_rate": 1e-3, "steps_per_batch": 16, "max_steps": 1024, "warm_spp": 8,
              "warm_steps_per_batch": 128}
# A faster-learning schedule found by the same sweeps (lr 3e-3, 2^15 samples
# per step), reported beside the headline.
GUIDE_TUNED = {"bsdf_fraction": 0.5, "train": 1, "record_paths": 1 << 20, "batch_size": 1 << 15,
               "learning_rate": 3e-3, "steps_per_batch": 8, "max_steps": 512, "warm_spp": 8,
               "warm_steps_per_batch": 32}
GUIDE_SEEDS = (1, 2, 3)


def guided_room(args) -> dict:
    """Time to equal error, guided (online training included) vs plain, on
    the closed room lit through a window (1920x1080, depth 10), for the
    configuration's hyperparameters (GUIDE_SPEC) and for the faster-learning
    schedule (GUIDE_TUNED), each the mean over three network
    initialisations.  Errors are relative MSE against the plain tracer at
    --guided-ref-spp, whose agreement with the float64 oracle is tested
    (tests/test_gpu_configs.py)."""
    from colvillea_amd import Integrator, IntegratorType, guided_eval, scenes, upload_scene

    t = time.perf_counter()
    desc = scenes.make_c3(spp=args.guided_spp)
    log(f"room scene: {desc.num_triangles()} triangles in {time.perf_counter() - t:.1f} s")
    pts = [p for p in (8, 16, 32, 64, 128, 256, 512, 1024) if p < args.guided_spp] + [args.guided_spp]
    it = Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height)
    upload_scene(it, desc)
    ref = guided_eval.reference(it, desc, args.guided_ref_spp, log)
    r = guided_eval.time_to_equal_error(desc, args.guided_ref_spp, pts, GUIDE_SPEC, it=it, log=log, ref=ref,
                                        seeds=GUIDE_SEEDS)
    r2 = guided_eval.time_to_equal_error(desc, args.guided_ref_spp, pts, GUIDE_TUNED, it=it, log=log, ref=ref,
                                         seeds=GUIDE_SEEDS, plain=r["plain"])
    it.close()
    nsamp = desc.width * desc.height * args.guided_spp

    def rnd(x, k=3):
        return None if x is None else round(x, k)

    def line(res):
        return {"guide": res["guide"], "t_guided_s": rnd(res["t_guided"]), "time_ratio": rnd(res["ratio"]),
                "time_ratio_per_seed": [rnd(x) for x in res["ratios"]],
                "relmse_guided_final": res["guided"][-1]["relmse"],
                "guided_Msamples_per_s": round(nsamp / res["guided"][-1]["seconds"] / 1e6, 2),
                "train_steps": res["train_steps"], "samples_trained": res["samples_trained"]}

    out = {"workload": f"closed 10x10x4 room lit through a 1x1 window by the 1080p scene's HDRI, "
                       f"{desc.num_triangles()} triangles, {desc.width}x{desc.height}, {args.guided_spp} spp, "
                       f"depth {desc.max_depth}, seed 4",
           "hyperparameters": "configuration's: Adam lr 1e-3, 2^16 samples per step", "seeds": list(GUIDE_SEEDS),
           "t_plain_s": rnd(r["t_plain"]), "relmse_plain_final": r["eps"],
           "plain_Msamples_per_s": round(nsamp / r["t_plain"] / 1e6, 2), "reference_spp": r["ref_spp"],
           "reference": "plain GPU tracer, two halves on independent RNG lanes, its own noise subtracted; its "
                        "expectation is tested against the float64 oracle"}
    out.update(line(r))
    out["tuned"] = line(r2)
    out["curves"] = {"plain": r["plain"], "guided": r["guided"], "guided_tuned": r2["guided"]}
    return out


def guide_train_step(args) -> dict:
    """Device time of one Adam step of the learned proposal (fused step and
    the per-GEMM comparison path) at its full size, as the guided tracer runs
    it: inside a render of the open sun-and-sky scene (512x512, 16 spp in one
    batch, 2^22 recorded paths), whose paths nearly all return light, so
    that every step holds its full batch of samples; 16 steps after the
    batch, timed by the library's HIP events around the training launches."""
    from colvillea_amd import Integrator, IntegratorType, scenes, upload_scene

    desc = scenes.make_c0(width=512, height=512, spp=args.train_spp)
    it = Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height)
    upload_scene(it, desc)
    res, live, nsteps = {}, {}, 0
    sizes = [int(x) for x in args.train_sizes.split(",")]
    for n in sizes:
        for kern, name in ((0, "fused"), (1, "per_gemm")):
            it.guideReset(1)
            it.setGuide(True, train=1, batch_size=n, record_paths=1 << 22, steps_per_batch=16, learning_rate=1e-3,
                        train_kernel=kern)
            it.resetIterationIndex()
            it.render(args.train_spp)
            i0 = it.guideStats()["issued"]
            it.render(args.train_spp)
            st = it.guideStats()
            nsteps = st["issued"] - i0
            res.setdefault(name, {})[str(n)] = round(1e3 * st["train_ms"] / max(1, nsteps), 2)
            live.setdefault(name, {})[str(n)] = int(st["last_count"])
    it.setGuide(False)
    it.close()
    # 163840 flop per sample: forward 2 x 32768, backward data 32768, weight gradients 65536
    flop = {str(n): 163840 * min(n, live["fused"][str(n)]) for n in sizes}
    return {"us_per_step": res, "live_samples_last_step": live, "steps_timed_per_case": nsteps,
            "tflops_fused": {k: round(flop[k] / (v * 1e-6) / 1e12, 1) for k, v in res["fused"].items()},
            "gflop_per_step": {k: round(v / 1e9, 2) for k, v in flop.items()}, "peak_bf16_tflops": 2500.0,
            "note": "device time per Adam step (training launches only) inside a guided render of the sun-and-sky "
                    "scene, 512x512, full steps"}


def stress_c4(args) -> dict:
    """The stress configuration (3840x2160, 128 spp, depth 12, 16.8 M-triangle
    fBm grid + 2000 instanced icospheres, half dielectric, 8192x4096 HDRI):
    plain throughput over the full frame, then the FP8 ablation of the
    learned proposal -- the same guided schedule with the BF16 and the FP8
    network, each against plain at equal error (reference: plain at
    --c4-ref-spp)."""
    from colvillea_amd import Integrator, IntegratorType, guided_eval, scenes, upload_scene

    t = time.perf_counter()
    desc = scenes.make_c4()
    log(f"C4: {desc.num_triangles()} triangles generated in {time.perf_counter() - t:.1f} s")
    it = Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height)
    t = time.perf_counter()
    upload_scene(it, desc)
    it.sync()
    log(f"C4 upload + BVH + HDRI tables: {time.perf_counter() - t:.1f} s")
    it.render(1)  # warm-up
    it.resetIterationIndex()
    it.sync()
    t = time.perf_counter()
    it.render(desc.spp)
    it.sync()
    dt = time.perf_counter() - t
    plain = desc.width * desc.height * desc.spp / dt / 1e6
    log(f"C4 plain: {desc.spp} spp in {dt:.2f} s -> {plain:.1f} Msamples/s")
    pts = [p for p in (4, 8, 16, 32, 64) if p < args.c4_guided_spp] + [args.c4_guided_spp]
    ref = guided_eval.reference(it, desc, args.c4_ref_spp, log)
    ablation, plain_curve = {}, None
    for name, prec in (("bf16", 0), ("fp8", 1)):
        r = guided_eval.time_to_equal_error(desc, args.c4_ref_spp, pts, dict(GUIDE_SPEC, precision=prec), it=it,
                                            log=log, ref=ref, plain=plain_curve)
        plain_curve = r["plain"]
        ablation[name] = {"time_ratio": None if r["ratio"] is None else round(r["ratio"], 3),
                          "t_guided_s": None if r["t_guided"] is None else round(r["t_guided"], 3),
                          "t_plain_s": round(r["t_plain"], 3), "relmse_plain_final": r["eps"],
                          "relmse_guided_final": r["guided"][-1]["relmse"],
                          "relmse_ratio_equal_spp": round(r["guided"][-1]["relmse"] / r["eps"], 3),
                          "guided_s_final": round(r["guided"][-1]["seconds"], 3),
                          "train_s_final": round(r["guided"][-1].get("train_seconds", 0.0), 3),
                          "curve": r["guided"]}
    it.close()
    return {"workload": f"C4: {desc.width}x{desc.height}, {desc.spp} spp, depth {desc.max_depth}, "
                        f"{desc.num_triangles()} triangles, 8192x4096 HDRI, seed {desc.info['seed']}",
            "plain_Msamples_per_s": round(plain, 2), "ablation_spp": args.c4_guided_spp,
            "reference_spp": args.c4_ref_spp, "guide": dict(GUIDE_SPEC), "ablation": ablation,
            "plain_curve": plain_curve}


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
    ap.add_argument("--bvh-width", type=int, default=4, choices=(4, 8), help="tree width of the C1 scene's BVH")
    ap.add_argument("--samples-per-batch", type=int, default=0,
                    help="spp of every pixel per batch for the 1080p render (0 = library default)")
    ap.add_argument("--no-other-width", action="store_true", help="skip the 1080p render on the other BVH width")
    ap.add_argument("--no-c2", action="store_true", help="skip the batched 256-view render")
    ap.add_argument("--c2-steps", type=int, default=2)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-guided", action="store_true", help="skip the guided-vs-plain room evaluation")
    ap.add_argument("--guided-ref-spp", type=int, default=1024)
    ap.add_argument("--guided-spp", type=int, default=256)
    ap.add_argument("--no-c4", action="store_true", help="skip the 4K stress scene and the FP8 ablation")
    ap.add_argument("--c4-ref-spp", type=int, default=256)
    ap.add_argument("--c4-guided-spp", type=int, default=32)
    ap.add_argument("--autotune-log", default=None, help="write the autotuner's trials here (JSON)")
    ap.add_argument("--train-spp", type=int, default=16, help="spp of the renders that time the Adam step")
    ap.add_argument("--train-sizes", default="32768,65536", help="samples per Adam step to time (comma list)")
    ap.add_argument("--only-guide-train", action="store_true",
                    help="print only the Adam-step timing of the learned proposal (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); sharding over GPUs (sample indices "
              "split across GPUs, then one float64 film reduction) is not run.", file=sys.stderr)
        sys.exit(2)

    if args.only_guide_train:
        print(json.dumps(guide_train_step(args)), flush=True)
        return
    from colvillea_amd import Integrator, IntegratorType, scenes, upload_scene

    t0 = time.perf_counter()
    desc = scenes.make_c1(width=args.width, height=args.height, spp=args.spp)
    log(f"scene: {desc.num_triangles()} triangles, generated in {time.perf_counter() - t0:.1f} s")
    integ = Integrator.createIntegrator(IntegratorType.WavefrontPathTracing, desc.width, desc.height)
    integ.setBVHWidth(args.bvh_width)
    desc.samples_per_batch = args.samples_per_batch
    t0 = time.perf_counter()
    upload_scene(integ, desc)
    integ.sync()
    log(f"BVH{args.bvh_width} upload + BVH + HDRI tables: {time.perf_counter() - t0:.1f} s")

    for i in range(args.warmup):
        integ.resetIterationIndex()
        integ.render(desc.spp)
        integ.sync()
        log(f"warmup {i + 1}/{args.warmup} done")

    # timed steps: plain renders (batch graphs replayed, no per-kernel events)
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
        log(f"step {i + 1}/{args.steps}: {dt * 1e3:.1f} ms, {desc.width * desc.height * desc.spp / dt / 1e6:.1f} "
            f"Msamples/s (device {c['render_ms']:.1f} ms, {c['batches']} batches, {c['graph_replays']} graph replays)")
    # one more step with HIP events around every launch, for the per-kernel
    # times of the roofline (not part of `value`)
    integ.setTiming(True)
    integ.resetIterationIndex()
    integ.render(desc.spp)
    integ.sync()
    c = integ.counters()
    integ.setTiming(False)
    trace_ms, shade_ms, shadow_ms = c["trace_ms"], c["shade_ms"], c["shadow_ms"]
    nodes, tris, launches = c["trace_nodes"], c["trace_tris"], c["trace_launches"]
    rays_traced = int(c["rays"][1:desc.max_depth + 1].sum())
    log(f"timed step: trace {trace_ms:.1f} ms, shade {shade_ms:.1f} ms, shadow {shadow_ms:.1f} ms over "
        f"{launches} closest-hit launches")
    samples = desc.width * desc.height * desc.spp * args.steps
    value = samples / t_total / 1e6
    batched = None if args.no_c2 else batched_views(integ, args)
    other = None if args.no_other_width else other_width(desc, args)
    # HBM roofline of the closest-hit kernel: its compulsory bytes per launch
    # -- every ray read and its hit written once, every node and triangle
    # record of the tree fetched once -- over its event-timed launches
    per_launch_ms = trace_ms / max(1, launches)
    queue_bytes = rays_traced * QUEUE_BYTES_PER_RAY - int(c["rays"][1]) * ORIGIN_BYTES
    compulsory = queue_bytes + launches * c["scene_bytes"]
    achieved = compulsory / (trace_ms * 1e-3) / 1e9 if trace_ms > 0 else 0.0
    # the bytes the kernel requests from the cache hierarchy: each node visit
    # and triangle test fetches its whole record (mostly L2/MALL hits)
    requested = queue_bytes + nodes * NODE_BYTES[args.bvh_width] + tris * TRI_BYTES
    req_gbs = requested / (trace_ms * 1e-3) / 1e9 if trace_ms > 0 else 0.0
    # measured per-launch counters of the same kernel and workload (rocprofv3
    # --pmc runs, summarised with their commands under profiles/)
    meas = None
    cpath = os.path.join(REPO, "profiles", "k_trace_c1_counters.json")
    if os.path.exists(cpath) and desc.width == 1920 and desc.height == 1080 and args.bvh_width == 4:
        with open(cpath) as fh:
            meas = json.load(fh)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": meas["hbm_bytes_per_launch"] if meas else None,
                "kernel": "k_trace", "bvh_width": args.bvh_width, "avg_launch_ms": round(per_launch_ms, 4),
                "algorithmic_bytes_per_launch": int(compulsory / max(1, launches)),
                "algorithmic_bytes": f"{QUEUE_BYTES_PER_RAY} B per closest-hit ray (ray read, hit written; camera rays "
                                     f"{QUEUE_BYTES_PER_RAY - ORIGIN_BYTES} B, their origin is the camera's) + the "
                                     f"tree's node and triangle records once per launch ({c['scene_bytes']} B)",
                "nodes_per_ray": round(nodes / max(1, rays_traced), 2),
                "tris_per_ray": round(tris / max(1, rays_traced), 2),
                "l2": {"requested_bytes_per_launch": int(requested / max(1, launches)), "achieved": round(req_gbs, 1),
                       "peak": L2_PEAK_GBS, "unit": "GB/s", "frac": round(req_gbs / L2_PEAK_GBS, 4),
                       "measured": meas.get("l2") if meas else None},
                "limiter": meas.get("limiter") if meas else None,
                "counters_source": "profiles/k_trace_c1_1080p.md" if meas else None}
    import ctypes as C

    from colvillea_amd import _capi as capi

    copy_ms = C.c_double()
    copy_bytes = args.copy_mib << 20
    capi.check(capi.lib().clhip_copy_probe(integ.ctx, copy_bytes, args.copy_iters, C.byref(copy_ms)), integ.ctx)
    copy_gbs = 2 * copy_bytes / (copy_ms.value * 1e-3) / 1e9
    roofline["copy_ceiling_GBs"] = round(copy_gbs, 1)
    log(f"copy probe: {args.copy_mib} MiB in {copy_ms.value:.3f} ms -> {copy_gbs:.0f} GB/s (read + write)")
    integ.close()
    integ = None
    guided = None if args.no_guided else guided_room(args)
    if guided is not None:
        guided["train_step"] = guide_train_step(args)
        log(f"Adam step: {guided['train_step']['us_per_step']}")
    stress = None if args.no_c4 else stress_c4(args)
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
           "roofline": roofline, "cpu_baseline": cpu, "other_bvh_width": other, "batched": batched, "guided": guided}
    if stress is not None:
        out["stress"] = stress
    if integ is not None:
        integ.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
