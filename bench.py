"""Benchmark: incoherent closest-hit throughput on the 1M-triangle scene.

Workload (the configuration the headline metric is quoted on): scene
configuration 2 -- 1,040,384 triangles (512x512-quad fBm terrain, 400
icospheres, 4096 emissive triangles) -- at 1920x1080. The ray set is R_inc:
one cosine-weighted bounce ray from each primary hit (2,073,600 rays),
generated on the GPU by the camera and bounce-ray kernels. One step = one
mipt_trace_closest launch over all 2,073,600 rays, inputs resident in HBM,
over the binned-SAH tree (--bvh lbvh: the LBVH).

Prints ONE JSON line: Mrays/s over the timed steps, the roofline of the
traversal kernel (algorithmic bytes per launch / mean launch time, HIP events
on the launch stream) and the CPU baseline (the oracle's restatement of the
reference CPU renderer's octree traversal, OpenMP over all host cores, on the
same rays). A second object, "plain_render", times the plain path tracer on
the 512x512 Cornell-box scene (16 x 1-spp passes, 4 bounces) beside the
oracle's CPU megakernel restatement. A third, "cache_vs_plain", is the
second half of the metric: equal-time relMSE of the learned radiance cache
against the plain tracer at 1920x1080 on the config-1 scene, against an
FP64 4096-spp reference: "cache_vs_plain" with the cache as configuration 4
states it (queried at bounce 2, trained on 1/16 of the paths extended 3
bounces), "cache_vs_plain_tuned" with the library's defaults (bounce 1,
1/64 of the paths). "restir_render" times ReSTIR DI frames of the
1M-triangle scene at 1920x1080. "full_stack_render" times the full stack at
scale: the 10M-triangle scene with its 16,384 emitters and a 4096 x 2048
envmap at 3840x2160, ReSTIR DI, 8 bounces and the radiance cache.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--bvh sah|lbvh] [--no-cpu] [--no-render] [--no-cache]
       [--cache-ref-spp 4096] [--cache-plain-spp 32] [--cache-warmup 2] [--no-restir] [--restir-layered]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hiprt-path-tracer_amd"))

METRIC = "Mrays/s closest-hit (1M tris); equal-time relMSE, cache vs plain, 1080p"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_ray(mean_nodes, mean_tris):
    # 32 B ray in + 16 B hit out + one 64-byte node per traversal step +
    # one 48-byte triangle record per intersection test.
    return 32.0 + 16.0 + 64.0 * mean_nodes + 48.0 * mean_tris


def step(what):
    """Names the leg about to run, on stderr, before its GPU work starts."""
    print(f"bench: {what}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-render", action="store_true", help="skip the plain-render timing leg")
    ap.add_argument("--no-cache", action="store_true", help="skip the radiance-cache vs plain leg")
    ap.add_argument("--cache-ref-spp", type=int, default=4096,
                    help="samples of the FP64 plain reference of the cache leg")
    ap.add_argument("--cache-plain-spp", type=int, default=32, help="plain samples whose time is the budget")
    ap.add_argument("--cache-warmup", type=int, default=2,
                    help="warm-up frames (training paths only) before the cache form accumulates (their time "
                         "counts)")
    ap.add_argument("--cache-query-bounce", type=int, default=None,
                    help="vertex at which the cache answers (default: the library's, 1)")
    ap.add_argument("--cache-train-every", type=int, default=None,
                    help="one training path in N (default: the library's, 64)")
    ap.add_argument("--cache-warmup-train-every", type=int, default=None,
                    help="one path in N runs in a warm-up frame (default: the library's, 16)")
    ap.add_argument("--frames", choices=["plain", "cache", "warmup"], default=None,
                    help="only render 3 + --steps samples of the config-1 scene at 1920x1080, plain, with the "
                         "radiance cache or as cache warm-up frames, and print ms per sample (a kernel-profiling "
                         "workload)")
    ap.add_argument("--shade-order", choices=["live", "material", "auto"], default=None,
                    help="with --frames: the path pass's shading order (default: the library's, auto)")
    ap.add_argument("--no-restir", action="store_true", help="skip the ReSTIR DI leg")
    ap.add_argument("--no-views", action="store_true", help="skip the batched-views leg (configuration 3)")
    ap.add_argument("--no-full-stack", action="store_true",
                    help="skip the full-stack leg (10M triangles, envmap, 3840x2160)")
    ap.add_argument("--views", type=int, default=96, help="views of the batched-views leg")
    ap.add_argument("--no-bake", action="store_true", help="skip the LUT baker leg")
    ap.add_argument("--restir-layered", action="store_true",
                    help="also time the ReSTIR DI leg with the config-2 scene's full layered materials")
    ap.add_argument("--bvh", choices=["lbvh", "sah"], default="sah",
                    help="BVH builder of the traversal leg: LBVH or top-down binned SAH")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--stack", type=int, default=0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (reference for the HBM byte counters)")
    args = ap.parse_args()
    cache_over = {k: v for k, v in (("query_bounce", args.cache_query_bounce),
                                    ("train_every", args.cache_train_every),
                                    ("warmup_train_every", args.cache_warmup_train_every)) if v is not None}
    if args.frames:
        return frames_only(args.frames, args.steps, cache_over, args.shade_order)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import mipt

    assert torch.cuda.is_available(), "bench.py needs a HIP GPU"
    if args.calibrate_copy:
        nbytes = 1 << 30
        src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "launches": 3}))
        return
    W, H = 1920, 1080
    n = W * H
    step("traversal leg: config 2 scene, 1920x1080 R_inc rays, closest hit")
    scene = mipt.generate_scene(2, W, H)
    ctx = mipt.Context(0)
    ctx.set_scene(scene)
    builder = mipt.MIPT_BVH_SAH_BINNED if args.bvh == "sah" else mipt.MIPT_BVH_LBVH
    bvh_name = "SAH_BINNED" if args.bvh == "sah" else "LBVH"
    ctx.build_bvh(builder)
    if args.block or args.stack:
        ctx.set_trace_config(args.block, args.stack)
    info = ctx.bvh_info()

    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    d_cam = mipt.empty_device(32 * n)
    d_rng = mipt.empty_device(4 * n)
    d_hits = mipt.empty_device(16 * n)
    d_rays = mipt.empty_device(32 * n)
    d_counts = mipt.empty_device(8 * n)
    ctx.camera_rays(scene.camera, W, H, 0, mipt.host_random_number(0), 0, d_cam, d_rng, sh)
    ctx.trace_closest(d_cam, d_hits, n, sh)
    ctx.bounce_rays(d_cam, d_hits, d_rng, d_rays, n, sh)
    ctx.trace_closest_counted(d_rays, d_hits, d_counts, n, sh)
    stream.synchronize()
    counts = mipt.from_device(d_counts, np.uint32)[: 2 * n].reshape(-1, 2).astype(np.float64)
    mean_nodes, mean_tris = counts[:, 0].mean(), counts[:, 1].mean()
    b_ray = algorithmic_bytes_per_ray(mean_nodes, mean_tris)

    for _ in range(args.warmup):
        ctx.trace_closest(d_rays, d_hits, n, sh)
    stream.synchronize()

    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in events:
        e0.record(stream)
        ctx.trace_closest(d_rays, d_hits, n, sh)
        e1.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [e0.elapsed_time(e1) for e0, e1 in events]
    mean_launch_s = float(np.mean(launch_ms)) / 1e3
    value = n * args.steps / wall / 1e6
    achieved_gbs = n * b_ray / mean_launch_s / 1e9
    hits = mipt.from_device(d_hits, mipt.HIT_DTYPE)[:n]

    cpu = None
    if not args.no_cpu:
        step("CPU baseline: oracle octree on every second ray")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        rays_host = mipt.from_device(d_rays, mipt.RAY_DTYPE)[:n]
        # The host cores this process may use (OMP_NUM_THREADS caps it on a
        # shared GPU host), not every core the machine reports.
        cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        tree = oracle.Octree(scene.vertices, scene.indices)
        sample = rays_host[::2]
        tc = time.perf_counter()
        cpu_hits = tree.trace(sample, nthreads=cores)
        cpu_s = time.perf_counter() - tc
        agree = float((cpu_hits["prim"] == hits[::2]["prim"]).mean())
        cpu = {"value": round(len(sample) / cpu_s / 1e6, 4), "unit": "Mrays/s", "cores": cores, "kind": "port",
               "sample": f"{len(sample)} R_inc rays (every 2nd of the timed set), octree of k-DOPs "
                         f"(max depth 32, <=8 tris/leaf), OpenMP dynamic over rays; {cpu_s:.2f} s; "
                         f"hit-ID agreement with GPU {agree:.6f}"}

    # HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    # (scaled by a plain copy), recorded in profiles/ for this kernel and
    # launch configuration; null when the configuration differs.
    traffic = None
    l2_hit_rate = None
    prof_source = None
    for name in ("trace_closest_config2_rinc.json", "trace_closest_config2_rinc_sah.json"):
        prof = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(prof) or args.block or args.stack:
            continue
        with open(prof) as f:
            pj = json.load(f)
        if pj.get("bvh", "LBVH") == bvh_name:
            traffic = pj.get("traffic_bytes_per_launch")
            l2_hit_rate = pj.get("l2_hit_rate")
            prof_source = pj.get("summary")
    # What bounds the kernel, from the counters: the algorithmic bytes are an
    # effective bandwidth (they grow with node visits, most of which L2 and
    # the Infinity Cache serve); the measured HBM bytes say whether HBM is
    # the limit.
    measured = None
    limiter = None
    if traffic:
        hbm_gbs = traffic / mean_launch_s / 1e9
        measured = {"hbm_gbs": round(hbm_gbs, 1), "hbm_frac": round(hbm_gbs / HBM_PEAK_GBS, 4),
                    "traffic_over_algorithmic": round(traffic / (n * b_ray), 4), "l2_hit_rate": l2_hit_rate}
        limiter = ("hbm bandwidth" if hbm_gbs / HBM_PEAK_GBS >= 0.5 else
                   f"dependent node-fetch latency and L2 (HBM traffic is {traffic / (n * b_ray) * 100:.1f} % of the "
                   f"algorithmic bytes and {hbm_gbs / HBM_PEAK_GBS * 100:.1f} % of HBM peak; L2 hit rate "
                   f"{(l2_hit_rate or 0) * 100:.1f} %)")

    if not args.no_render:
        step("plain render leg: config 1, 512x512, 16 x 1-spp passes")
    render = None if args.no_render else plain_render_leg(mipt, torch, np, args.no_cpu)
    with_restir = not (args.no_render or args.no_restir)
    if with_restir:
        step("ReSTIR DI leg: config 2 scene, 1920x1080, 5 bounces, diffuse + metal lobes")
    restir = restir_render_leg(mipt, torch, np, layered=False) if with_restir else None
    restir_fused = None
    if with_restir:
        step("ReSTIR DI leg: the same with light presampling and the fused spatiotemporal pass")
        restir_fused = restir_render_leg(mipt, torch, np, layered=False, fused=True)
    if with_restir and args.restir_layered:
        step("ReSTIR DI leg: config 2 scene, 1920x1080, 5 bounces, full layered materials")
        restir_layered = restir_render_leg(mipt, torch, np, layered=True)
    else:
        restir_layered = None
    full_stack = None
    if not (args.no_render or args.no_full_stack):
        step("full-stack leg: 10M-triangle scene, 4096x2048 envmap, 3840x2160, ReSTIR DI, 8 bounces, cache")
        full_stack = full_stack_leg(mipt, torch, np)
    views = None
    if not (args.no_render or args.no_views):
        step(f"batched views leg: {args.views} views of the config-1 scene, 4 spp, 5 bounces")
        views = views_render_leg(mipt, torch, np, args.views)
    bake = None
    if not args.no_bake:
        step("LUT baker leg: the five energy-compensation tables at the reference's settings")
        bake = lut_bake_leg(mipt, np, args.no_cpu)
    cache = cache_tuned = None
    if not args.no_cache:
        step(f"radiance cache vs plain legs: config 1, 1920x1080; FP64 reference of {args.cache_ref_spp} spp")
        ref, ref_s = plain_reference(mipt, torch, np, args.cache_ref_spp)
        ref_desc = (f"plain tracer, FP64 accumulation, {args.cache_ref_spp} spp ({ref_s:.1f} s on the GPU; "
                    f"bit-identical per sample to the oracle's CPU restatement)")
        # Configuration 4 as stated: queried at bounce 2, trained on 1/16 of
        # the paths, extended 3 bounces past the query vertex.
        stated = dict(query_bounce=2, train_every=16, train_extension=3)
        stated.update(cache_over)
        step("radiance cache leg as stated: query bounce 2, 1/16 training paths")
        cache = cache_vs_plain_leg(mipt, torch, np, ref, ref_desc, args.cache_plain_spp, args.cache_warmup, stated)
        step("radiance cache leg, tuned: the library defaults (query bounce 1, 1/64 training paths)")
        cache_tuned = cache_vs_plain_leg(mipt, torch, np, ref, ref_desc, args.cache_plain_spp, args.cache_warmup,
                                         {})

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded procedural scene, GPU-generated ray set)",
        "config": {"workload": "config 2 scene (1,040,384 tris) at 1920x1080, R_inc incoherent bounce rays, "
                               "closest hit", "rays_per_step": n, "bvh": bvh_name,
                   "block": args.block or 128,
                   "lds_stack": args.stack or 16},
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "achieved_kind": "algorithmic bytes (B_ray x rays) / mean launch time: an effective bandwidth",
                     "measured": measured, "limiter": limiter,
                     "traffic_source": prof_source if traffic else None,
                     "bytes_per_ray": round(b_ray, 1), "mean_nodes_per_ray": round(mean_nodes, 2),
                     "mean_tris_per_ray": round(mean_tris, 2), "mean_launch_ms": round(mean_launch_s * 1e3, 4)},
        "cpu_baseline": cpu,
        "bvh": {"nodes": info.node_count, "leaves": info.leaf_count, "max_depth": info.max_depth,
                "sah_cost": round(info.sah_cost, 2), "build_ms": round(info.build_ms, 2)},
        "hit_fraction": round(float((hits["prim"] >= 0).mean()), 4),
        "plain_render": render,
        "restir_render": restir,
        "restir_render_layered": restir_layered,
        "restir_render_fused": restir_fused,
        "views_render": views,
        "full_stack_render": full_stack,
        "lut_bake": bake,
        "cache_vs_plain": cache,
        "cache_vs_plain_tuned": cache_tuned,
    }
    print(json.dumps(out))
    ctx.close()


def frames_only(mode, samples, cache_over=None, shade_order=None):
    """Config-1 scene at 1920x1080, 4 bounces, uniform NEE + MIS: 3 warm-up and
    `samples` timed calls of mipt_render_sample, plain or with the radiance
    cache (default settings, cache_over replacing some), or of
    mipt_radiance_cache_warmup. Prints one JSON line."""
    import torch

    import mipt

    W, H = 1920, 1080
    scene = mipt.generate_scene(1, W, H)
    ctx = mipt.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    if shade_order:
        ctx.set_shading_order({"live": mipt.MIPT_SHADE_ORDER_LIVE, "material": mipt.MIPT_SHADE_ORDER_MATERIAL,
                               "auto": mipt.MIPT_SHADE_ORDER_AUTO}[shade_order])
    cache = mode != "plain"
    frame = ctx.cache_warmup if mode == "warmup" else ctx.render_sample
    if cache:
        ctx.set_radiance_cache(mipt.default_radiance_cache_settings(**(cache_over or {})))
    ctx.set_render_settings(mipt.default_render_settings(width=W, height=H, nb_bounces=4, accumulate_fp64=1,
                                                         radiance_cache=int(cache)))
    ctx.set_camera(scene.camera)
    for i in range(3):
        frame(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(samples):
        frame(3 + i)
    torch.cuda.synchronize()
    out = {"frames": mode, "samples": samples, "shade_order": shade_order or "auto",
           "ms_per_sample": (time.perf_counter() - t0) / samples * 1e3}
    if cache:
        q, t, dropped, loss = ctx.cache_stats()
        out.update({"queries": q, "training_records": t, "dropped_targets": dropped, "train_loss": loss})
    print(json.dumps(out))
    ctx.close()


def plain_reference(mipt, torch, np, ref_spp):
    """The reference image of the equal-time legs: the plain tracer on the
    config-1 scene at 1920x1080, 4 bounces, uniform NEE + MIS, accumulated
    in FP64 over ref_spp samples (the GPU tracer is bit-identical to the
    oracle's CPU restatement per sample, so this is the FP64 oracle image).
    Returns (mean image, seconds)."""
    W, H = 1920, 1080
    scene = mipt.generate_scene(1, W, H)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    ctx = mipt.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_render_settings(mipt.default_render_settings(width=W, height=H, nb_bounces=4, accumulate_fp64=1))
    ctx.set_camera(scene.camera)
    t0 = time.perf_counter()
    for s in range(ref_spp):
        ctx.render_sample(s, sh)
        if s % 512 == 511:
            stream.synchronize()
            step(f"  reference: {s + 1} of {ref_spp} samples")
    stream.synchronize()
    ref_s = time.perf_counter() - t0
    ref = ctx.read_framebuffer(fp64=True)[1] / ref_spp
    ctx.close()
    return ref, ref_s


def cache_vs_plain_leg(mipt, torch, np, ref, ref_desc, plain_spp, warmup, cache_over=None):
    """Equal-time relMSE, learned radiance cache vs plain, config-1 scene at
    1920x1080, 4 bounces, uniform NEE + MIS. relMSE = mean over pixels and
    channels of (x - r)^2 / (r^2 + 1e-2), reduced in float64, against the
    FP64 reference image `ref` (plain_reference). The plain form renders
    plain_spp samples; its wall time is the budget. The cache form (by default
    queried at bounce 1, trained every frame on 1/64 of the paths, batches of
    2^16, Adam lr 1e-2; cache_over replaces settings) first runs `warmup`
    warm-up frames (mipt_radiance_cache_warmup: 1/16 of the paths, all of them
    training paths, no image), whose time counts, then accumulates until it
    matches the plain relMSE or runs 3x the budget.
    Reported: both relMSE at equal time and the cache's time to reach the
    plain relMSE as a fraction of the plain time (interpolated between
    checkpoints in log-log). Wall times are host clock around synchronised
    sample loops, excluding the relMSE evaluations."""
    W, H = 1920, 1080
    scene = mipt.generate_scene(1, W, H)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream

    def make(cache):
        ctx = mipt.Context(0)
        ctx.set_scene(scene)
        ctx.build_bvh()
        if cache:
            ctx.set_radiance_cache(mipt.default_radiance_cache_settings(**(cache_over or {})))
        ctx.set_render_settings(mipt.default_render_settings(width=W, height=H, nb_bounces=4, accumulate_fp64=1,
                                                             radiance_cache=1 if cache else 0))
        ctx.set_camera(scene.camera)
        return ctx

    def relmse(ctx, spp):
        img = ctx.read_framebuffer(fp64=True)[1] / spp
        return float(np.mean((img - ref) ** 2 / (ref * ref + 1e-2)))

    checkpoints = sorted({1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, plain_spp})

    def run(ctx, first_sample, stop):
        """[(seconds, spp, relMSE)] at the checkpoints, until stop(curve)."""
        t = 0.0
        curve = []
        s = 0
        while True:
            stream.synchronize()
            t1 = time.perf_counter()
            ctx.render_sample(first_sample + s if s else 0, sh)
            stream.synchronize()
            t += time.perf_counter() - t1
            s += 1
            if s in checkpoints or s % 256 == 0:
                curve.append((t, s, relmse(ctx, s)))
                if stop(curve):
                    return curve

    ctx = make(False)
    plain = run(ctx, 0, lambda c: c[-1][1] >= plain_spp)
    ctx.close()
    budget, target = plain[-1][0], plain[-1][2]

    ctx = make(True)
    settings = mipt.default_radiance_cache_settings(**(cache_over or {}))
    tw = time.perf_counter()
    for i in range(warmup):
        ctx.cache_warmup(1000000 + i, sh)
    stream.synchronize()
    warm_s = time.perf_counter() - tw
    cached = run(ctx, 0, lambda c: (c[-1][0] + warm_s >= budget and c[-1][2] <= target)
                 or c[-1][0] + warm_s >= 3.0 * budget)
    queries, train, dropped, loss = ctx.cache_stats()
    ctx.close()
    cached = [(t + warm_s, s, e) for t, s, e in cached]

    def at_time(curve, t):
        # relMSE at time t (log-log interpolation; the last point beyond)
        prev = None
        for p in curve:
            if p[0] >= t:
                if prev is None:
                    return p[2]
                a = (np.log(t) - np.log(prev[0])) / (np.log(p[0]) - np.log(prev[0]))
                return float(np.exp(np.log(prev[2]) + a * (np.log(p[2]) - np.log(prev[2]))))
            prev = p
        return curve[-1][2]

    reach = None
    prev = (warm_s, 0, None)
    for p in cached:
        if p[2] <= target:
            if prev[2] is None:
                reach = p[0]
            else:
                a = (np.log(target) - np.log(prev[2])) / (np.log(p[2]) - np.log(prev[2]))
                reach = float(np.exp(np.log(prev[0]) + a * (np.log(p[0]) - np.log(prev[0]))))
            break
        prev = p
    return {"workload": f"config 1 scene (102,428 tris) 1920x1080, 4 bounces, uniform NEE + MIS; cache queried at "
                        f"bounce {settings.query_bounce}, trained on 1/{settings.train_every} of paths extended "
                        f"{settings.train_extension} bounces past it (batch "
                        f"2^16, Adam lr 1e-2, relative L2); {warmup} warm-up frames of 1/{settings.warmup_train_every}"
                        f" of paths",
            "reference": ref_desc,
            "plain_spp": plain_spp, "plain_seconds": round(budget, 4), "plain_relmse": target,
            "plain_ms_per_spp": round(budget / plain_spp * 1e3, 3),
            "cache_relmse_at_equal_time": at_time(cached, budget),
            "cache_seconds_to_plain_relmse": None if reach is None else round(reach, 4),
            "time_ratio_cache_over_plain": None if reach is None else round(reach / budget, 4),
            "cache_settings_overrides": dict(cache_over or {}),
            "cache_warmup_frames": warmup, "cache_warmup_seconds": round(warm_s, 4),
            "cache_ms_per_spp": round((cached[-1][0] - warm_s) / cached[-1][1] * 1e3, 3),
            "cache_queries_last_frame": queries, "cache_training_records_last_frame": train,
            "cache_dropped_targets": dropped, "cache_train_loss_last_frame": loss,
            "curve_plain": [[round(t, 4), s, e] for t, s, e in plain],
            "curve_cache": [[round(t, 4), s, e] for t, s, e in cached]}


def restir_render_leg(mipt, torch, np, layered, fused=False):
    """ReSTIR DI at scale: the 1,040,384-triangle scene at 1920x1080 with 4096
    power-sampled emitters; 32 light + 1 BSDF initial candidates, temporal
    reuse (M-cap 20), 2 spatial passes of 5 neighbours in 16 px, pairwise
    MIS, 5 bounces. With layered=False the materials keep their diffuse and
    metal lobes only (coat, sheen, specular, transmission and thin film set
    to 0); with layered=True they are the scene's full layered Principled
    materials. With fused=True the temporal pass and the first spatial pass
    are the one fused spatiotemporal pass, fed by the default light
    presampling (128 subsets of 1024 lights, 8x8 pixel tiles per subset).
    Times whole frames (camera + ReSTIR DI passes + path pass) and each
    part, with HIP events on the launch stream, after 2 warm-up frames."""
    W, H = 1920, 1080
    scene = mipt.generate_scene(2, W, H)
    if not layered:
        for m in scene.materials:
            m.coat = m.sheen = m.specular_transmission = m.diffuse_transmission = m.thin_film = m.specular = 0.0
    s = mipt.default_render_settings(width=W, height=H, nb_bounces=5,
                                     direct_light_sampling_strategy=mipt.MIPT_LSS_RESTIR_DI,
                                     direct_light_sampling_base_strategy=mipt.MIPT_LSS_BASE_POWER)
    rs = mipt.default_restir_di_settings(number_of_initial_light_candidates=32, m_cap=20, spatial_number_of_passes=2,
                                         do_fused_spatiotemporal=int(fused), do_light_presampling=int(fused))
    ctx = mipt.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_restir_di_settings(rs)
    ctx.set_render_settings(s)
    ctx.set_camera(scene.camera)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    draws = ctx.draws_per_sample()
    for f in range(2):
        ctx.render_sample(f, sh)
    stream.synchronize()
    frames = 5
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(frames)]
    t0 = time.perf_counter()
    for i in range(frames):
        f = 2 + i
        k = f * draws
        e = ev[i]
        e[0].record(stream)
        ctx.launch_pass(mipt.MIPT_PASS_CAMERA_RAYS, f, mipt.host_random_number(k), sh)
        e[1].record(stream)
        # the order and random numbers of mipt_render_sample
        if fused:
            ctx.launch_pass(mipt.MIPT_PASS_RESTIR_DI_LIGHTS_PRESAMPLING, f, mipt.host_random_number(k), sh)
        ctx.launch_pass(mipt.MIPT_PASS_RESTIR_DI_INITIAL_CANDIDATES, f, mipt.host_random_number(k + 1), sh)
        if fused:
            ctx.launch_pass(mipt.MIPT_PASS_RESTIR_DI_SPATIOTEMPORAL_REUSE, f, mipt.host_random_number(k + 4), sh)
            first_spatial = 1
        else:
            ctx.launch_pass(mipt.MIPT_PASS_RESTIR_DI_TEMPORAL_REUSE, f, mipt.host_random_number(k + 2), sh)
            first_spatial = 0
        for p in range(first_spatial, rs.spatial_number_of_passes):
            ctx.launch_pass(mipt.MIPT_PASS_RESTIR_DI_SPATIAL_REUSE(p), f,
                            mipt.host_random_number(k + 4 + p - first_spatial + (1 if fused else 0)), sh)
        e[2].record(stream)
        ctx.launch_pass(mipt.MIPT_PASS_PATH_TRACE, f, mipt.host_random_number(k + draws - 1), sh)
        e[3].record(stream)
    stream.synchronize()
    frame_ms = (time.perf_counter() - t0) / frames * 1e3
    cam_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    restir_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    path_ms = float(np.mean([e[2].elapsed_time(e[3]) for e in ev]))
    res = ctx.read_reservoirs(3)
    hit = ctx.read_gbuffer_prims() >= 0
    rej = ctx.rejected_samples()
    fb = ctx.read_framebuffer()
    ctx.close()
    return {"workload": "config 2 scene (1,040,384 tris, 4096 emitters) 1920x1080, ReSTIR DI 32+1 candidates, "
                        + ("light presampling 128 x 1024, fused spatiotemporal pass + 1 spatial pass"
                           if fused else "temporal M-cap 20, 2 spatial passes")
                        + " x 5 neighbours r=16, pairwise MIS, 5 bounces; "
                        + ("full layered Principled materials" if layered else "diffuse + metal lobes"),
            "frames": 2 + frames, "ms_per_frame": round(frame_ms, 3), "camera_pass_ms": round(cam_ms, 3),
            "restir_di_passes_ms": round(restir_ms, 3), "path_pass_ms": round(path_ms, 3),
            "rejected_samples": rej,
            "shaded_reservoirs": {"hit_pixels": int(hit.sum()),
                                  "M_min_on_hits": int(res["M"][hit].min()) if hit.any() else None,
                                  "M_max": int(res["M"].max()),
                                  "ucw_positive": int((res["ucw"] > 0).sum()),
                                  "ucw_nonfinite": int((~np.isfinite(res["ucw"])).sum())},
            "mean_radiance": [round(float(x), 5) for x in (fb.reshape(-1, 3) / (2 + frames)).mean(axis=0)]}


def full_stack_leg(mipt, torch, np):
    """The full stack at scale: the 10,479,104-triangle scene (8 fBm grids,
    3000 flattened instances, 16,384 emitters of power log-uniform over three
    decades, every Principled layer per object) under the 4096 x 2048
    procedural sun + sky envmap (alias-table sampled, ReSTIR DI envmap
    candidates with probability 0.5) at 3840x2160: ReSTIR DI with 32 light +
    1 BSDF candidates, temporal reuse (M-cap 20) and 2 spatial passes of 5
    neighbours in 16 px, 8 bounces, the radiance cache queried at bounce 2
    and trained on 1/16 of the paths. ReSTIR DI does not take dispersive
    glass yet: the dispersion of transmissive materials is set to 0 (the
    other layers stay). Times 3 frames after 2 warm-up frames and 8 cache
    warm-up frames; reports the SAH build of the 10M triangles."""
    W, H = 3840, 2160
    t0 = time.perf_counter()
    scene = mipt.generate_scene(5, W, H)
    gen_s = time.perf_counter() - t0
    dispersive = 0
    for m in scene.materials:
        if m.dispersion_scale > 0.0 and m.specular_transmission > 0.0:
            m.dispersion_scale = 0.0
            dispersive += 1
    env = mipt.generate_envmap(4096, 2048)
    ctx = mipt.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh(mipt.MIPT_BVH_SAH_BINNED)
    info = ctx.bvh_info()
    ctx.set_envmap(env)
    ctx.set_world_settings(mipt.default_world_settings(ambient_light_type=mipt.MIPT_AMBIENT_ENVMAP))
    ctx.set_radiance_cache(mipt.default_radiance_cache_settings(query_bounce=2, train_every=16, train_extension=3))
    rs = mipt.default_restir_di_settings(number_of_initial_light_candidates=32, m_cap=20, spatial_number_of_passes=2)
    ctx.set_restir_di_settings(rs)
    s = mipt.default_render_settings(width=W, height=H, nb_bounces=8,
                                     direct_light_sampling_strategy=mipt.MIPT_LSS_RESTIR_DI,
                                     direct_light_sampling_base_strategy=mipt.MIPT_LSS_BASE_POWER, radiance_cache=1)
    ctx.set_render_settings(s)
    ctx.set_camera(scene.camera)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    for i in range(8):
        ctx.cache_warmup(100000 + i, sh)
    for f in range(2):
        ctx.render_sample(f, sh)
    stream.synchronize()
    frames = 3
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for f in range(frames):
        ctx.render_sample(2 + f, sh)
    e1.record(stream)
    stream.synchronize()
    wall_ms = (time.perf_counter() - t0) / frames * 1e3
    gpu_ms = e0.elapsed_time(e1) / frames
    queries, train, dropped, loss = ctx.cache_stats()
    fb = ctx.read_framebuffer()
    rej = ctx.rejected_samples()
    ctx.close()
    return {"workload": "config 5: 10M-triangle scene (8 fBm grids + 3000 flattened instances + 16,384 emitters), "
                        "4096x2048 sun + sky envmap, 3840x2160, ReSTIR DI 32+1 candidates, temporal M-cap 20, "
                        "2 spatial passes, envmap candidates p = 0.5, 8 bounces, radiance cache (bounce 2, 1/16); "
                        f"dispersion off under ReSTIR DI ({dispersive} transmissive materials)",
            "triangles": int(scene.triangle_count), "emitters": int(scene.emissive.size),
            "scene_generation_s": round(gen_s, 2), "bvh_build_ms": round(info.build_ms, 2),
            "bvh_nodes": int(info.node_count), "bvh_sah_cost": round(info.sah_cost, 2),
            "frames_timed": frames, "ms_per_frame": round(gpu_ms, 2), "wall_ms_per_frame": round(wall_ms, 2),
            "cache_queries": int(queries), "cache_training_records": int(train), "cache_loss": float(loss),
            "rejected_samples": int(rej),
            "mean_radiance": [round(float(x), 5) for x in (fb.reshape(-1, 3) / (2 + frames)).mean(axis=0)]}


def lut_bake_leg(mipt, np, no_cpu):
    """The energy-compensation LUT baker: the conductor, glossy dielectric,
    glass (entering, exiting) and thin-glass tables at the reference's sizes
    and sample counts (GPUBakerConstants.h, *DirectionalAlbedoSettings.h),
    one lane per texel (the reference's summation order), each table baked
    into HBM and timed with hipEvents around its kernel launches; compared
    through the Radiance file's 8-bit truncation with the tables the
    reference ships. The CPU figure is the oracle's restatement on 256 texels
    of each table at the same sample counts, on every host core."""
    ctx = mipt.Context(0)
    tables = []
    total_ms = total_samples = 0.0
    for name, table in (("conductor", mipt.MIPT_LUT_GGX_CONDUCTOR),
                        ("glossy_dielectric", mipt.MIPT_LUT_GLOSSY_DIELECTRIC),
                        ("glass", mipt.MIPT_LUT_GGX_GLASS), ("glass_inverse", mipt.MIPT_LUT_GGX_GLASS_INVERSE),
                        ("thin_glass", mipt.MIPT_LUT_GGX_THIN_GLASS)):
        s = mipt.default_bake_settings(table)
        arr, info = ctx.bake_lut(s)
        shipped = mipt.shipped_lut(table)
        # the reference writes its tables through stb_image_write's RGBE
        # truncation (8-bit mantissa)
        m, e = np.frexp(arr)
        q = np.ldexp(np.floor((arr * (m.astype(np.float32) * np.float32(256.0) / arr)).astype(np.float32)),
                     e - 8).astype(np.float32)
        total_ms += info.ms
        total_samples += info.samples
        tables.append({"table": name, "size": [s.size_cos_theta, s.size_roughness, s.size_ior],
                       "samples_per_texel": info.launches * info.iterations_per_launch, "ms": round(info.ms, 2),
                       "gsamples_per_s": round(info.samples / info.ms * 1e-6, 2),
                       "equal_to_shipped": round(float((q == shipped).mean()), 5)})
    ctx.close()
    out = {"workload": "5 energy-compensation tables (1.82M texels) at the reference's sample counts",
           "tables": tables, "total_ms": round(total_ms, 1),
           "gsamples_per_s": round(total_samples / total_ms * 1e-6, 2)}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        cores = os.cpu_count() or 1
        cpu_samples = 0.0
        t0 = time.perf_counter()
        rng = np.random.default_rng(5)
        for table in range(5):
            s = mipt.default_bake_settings(table)
            n = s.size_cos_theta * s.size_roughness * s.size_ior
            tex = rng.integers(0, n, 256)
            ipk, launches, _ = oracle.bake_plan(table, s.size_cos_theta, s.size_roughness, s.size_ior,
                                                s.integration_sample_count)
            oracle.bake_texels(table, s.size_cos_theta, s.size_roughness, s.size_ior, s.integration_sample_count,
                               tex, nthreads=cores)
            cpu_samples += 256.0 * ipk * launches
        cpu_s = time.perf_counter() - t0
        out["cpu_oracle"] = {"gsamples_per_s": round(cpu_samples / cpu_s * 1e-9, 4), "cores": cores,
                             "sample": "256 texels of each table at the reference's sample counts",
                             "s": round(cpu_s, 2)}
    return out


def views_render_leg(mipt, torch, np, count):
    """Configuration 3: `count` views of the config-1 scene (heights
    log-uniform in [128, 1024] px, aspect in [1, 16/9]) rendered as one
    batch at 4 spp, 5 bounces, uniform NEE + MIS: every pass kernel and
    traversal launch covers the pixels of every view, the traversal launch
    configuration is chosen by the autotuner on incoherent bounce rays of the
    largest view, and the samples replay a captured hipGraph. Timed: 4-spp
    frames of the batch (graph replays, and eager launches), and the same
    views rendered one after the other (one view per launch sequence)."""
    scene = mipt.generate_scene(1, 64, 64)
    views = mipt.generate_views(count)
    sizes = [(v.width, v.height) for v in views]
    total = sum(w * h for w, h in sizes)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    ctx = mipt.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_render_settings(mipt.default_render_settings(nb_bounces=5, accumulate_fp64=1))
    # autotuner: incoherent bounce rays of the largest view
    big = max(range(count), key=lambda i: sizes[i][0] * sizes[i][1])
    w, h = sizes[big]
    n = w * h
    d_cam = mipt.empty_device(32 * n)
    d_rng = mipt.empty_device(4 * n)
    d_hits = mipt.empty_device(16 * n)
    d_rays = mipt.empty_device(32 * n)
    ctx.camera_rays(views[big].camera, w, h, 0, mipt.host_random_number(0), 0, d_cam, d_rng, sh)
    ctx.trace_closest(d_cam, d_hits, n, sh)
    ctx.bounce_rays(d_cam, d_hits, d_rng, d_rays, n, sh)
    stream.synchronize()
    tuning, table = ctx.autotune_trace(d_rays, n, 3, sh)
    tuned_on = f"{n:,} incoherent bounce rays of the largest view ({w}x{h})"
    del d_cam, d_rng, d_hits, d_rays
    ctx.set_views(views)

    def frame(first, spp):
        stream.synchronize()
        t0 = time.perf_counter()
        for s in range(first, first + spp):
            ctx.render_sample(s, sh)
        stream.synchronize()
        return time.perf_counter() - t0

    frame(0, 1)  # sample 0 runs eagerly (it resets the accumulation)
    ctx.set_graph_capture(True)
    frame(1, 2)
    graph_s = min(frame(3, 4), frame(7, 4))
    replays = ctx.graph_replays()
    ctx.set_graph_capture(False)
    eager_s = frame(11, 4)
    fb = ctx.read_framebuffer()
    mean = [round(float(x), 5) for x in (fb.reshape(-1, 3) / 15).mean(axis=0)]
    # the same views one at a time
    ctx.set_views([])
    seq_s = 0.0
    for v, (w, h) in zip(views, sizes):
        ctx.set_render_settings(mipt.default_render_settings(width=w, height=h, nb_bounces=5, accumulate_fp64=1))
        ctx.set_camera(v.camera)
        seq_s += frame(0, 4)
    ctx.close()
    return {"workload": f"{count} views of the config-1 scene (102,428 tris), {total:,} pixels, heights log-uniform "
                        f"[128, 1024], aspect [1, 16/9]; 4 spp, 5 bounces, uniform NEE + MIS; one wavefront over "
                        f"every view",
            "views": count, "pixels": total,
            "batch_graph_ms_per_4spp": round(graph_s * 1e3, 2),
            "batch_eager_ms_per_4spp": round(eager_s * 1e3, 2),
            "one_view_at_a_time_ms_per_4spp": round(seq_s * 1e3, 2),
            "batch_speedup_over_one_at_a_time": round(seq_s / graph_s, 3),
            "Mpaths_per_s": round(total * 4 / graph_s / 1e6, 1),
            "graph_replays": replays,
            "autotune": {"block": tuning.block_size, "lds_stack": tuning.lds_stack, "queue_tile": tuning.queue_tile,
                         "best_us": round(tuning.best_ms * 1e3, 1), "default_us": round(tuning.default_ms * 1e3, 1),
                         "configurations": tuning.configurations,
                         "rays": tuned_on},
            "mean_radiance": mean}


def plain_render_leg(mipt, torch, np, no_cpu):
    """Plain path tracing of the 512x512 Cornell-box scene (config 1): 16
    one-sample passes of CameraRays + path trace, 4 bounces, uniform light
    NEE + MIS. GPU wall time per 16-spp frame; the oracle's per-pixel
    megakernel (OpenMP over the host cores) on 2 of the 16 passes."""
    W = H = 512
    spp = 16
    scene = mipt.generate_scene(1, W, H)
    settings = mipt.default_render_settings(width=W, height=H, nb_bounces=4)
    ctx = mipt.Context(0)
    ctx.set_scene(scene)
    ctx.build_bvh()
    ctx.set_render_settings(settings)
    ctx.set_camera(scene.camera)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    for s in range(2):
        ctx.render_sample(s, sh)
    stream.synchronize()
    frames = 3
    t0 = time.perf_counter()
    for _ in range(frames):
        for s in range(spp):
            ctx.render_sample(s, sh)
    stream.synchronize()
    gpu_frame_ms = (time.perf_counter() - t0) / frames * 1e3
    fb = ctx.read_framebuffer()
    ctx.close()
    out = {"workload": "config 1 (102,428 tris) 512x512, 16 x 1-spp passes, 4 bounces, uniform NEE + MIS",
           "gpu_ms_per_frame": round(gpu_frame_ms, 3),
           "gpu_ms_per_pass": round(gpu_frame_ms / spp, 4),
           "gpu_Mpaths_per_s": round(W * H * spp / gpu_frame_ms / 1e3, 2),
           "mean_radiance": [round(float(x), 5) for x in (fb.reshape(-1, 3) / spp).mean(axis=0)]}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        R = oracle.Renderer(scene, settings)
        passes = 2
        tc = time.perf_counter()
        for s in range(passes):
            R.render_sample(s, mipt.host_random_number(2 * s), mipt.host_random_number(2 * s + 1), nthreads=cores)
        cpu_pass_s = (time.perf_counter() - tc) / passes
        out["cpu_baseline"] = {"ms_per_pass": round(cpu_pass_s * 1e3, 1),
                               "ms_per_frame_16spp": round(cpu_pass_s * spp * 1e3, 1), "cores": cores,
                               "kind": "port", "sample": f"{passes} of the 16 passes, full 512x512"}
        out["gpu_speedup_vs_cpu"] = round(cpu_pass_s * spp * 1e3 / gpu_frame_ms, 1)
    return out


if __name__ == "__main__":
    main()
