#!/usr/bin/env python3
"""Closest-hit ray-query throughput of the HIP backend on one MI355X.

Workload (one step = one launch over the whole batch): the 1,048,576-
triangle heightfield (1024x512 quads over [0,1]^2, z = 0.1 * fBm with 6
octaves, lacunarity 2, gain 0.5, seeded Perlin), binned-SAH BVH with 32
bins and leaves of at most 4 triangles, and 4,194,304 f32 rays with
origins uniform in the mesh box inflated by 10% and directions uniform on
the sphere, tmax = inf. Rays, BVH and hits stay resident in HBM; the
timed region is K back-to-back closest-hit launches.

Prints one JSON line (value = whole-job Mrays/s at N=1). The line also
carries the box-room path-tracing throughput, the 64-spp 1080p instanced
field, a 1080p fur ball of 300,000 hair segments, a 1080p textured
scene (the pattern zoo, with and without its bindings), the batched-jobs
comparison (32 rooms, half with clouds: one launch sequence, one
persistent launch, 32 sequential sessions) and the time to relMSE 0.01 at
1080p of the radiance cache vs plain wavefront path tracing.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "render.rs_amd"))

N_RAYS = 4194304
GRID = (1024, 512)
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "trace_closest_heightfield_1m_traffic.json")


# A one-GPU share of the host runs at most this many worker threads.
HOST_THREAD_CAP = 16


def host_threads():
    """Worker threads of the CPU baselines: every CPU this process may run
    on (its affinity mask), at most the one-GPU host share."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(HOST_THREAD_CAP, n))


def host_description():
    """CPU model and the CPUs this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"cpu_model": model, "affinity_cpus": affinity, "threads": host_threads(),
            "thread_cap": HOST_THREAD_CAP}


def cpu_baseline(pos, idx, rays):
    """The oracle (C++ f64 restatement of render.rs's CPU traversal) over
    the same rays on the host's cores; BVH build excluded."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    threads = host_threads()
    m = orc.Mesh(pos, idx, bins=32, max_leaf=4, sah_leaf_mult=1)
    t0 = time.perf_counter()
    m.trace_closest(rays, threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": round(rays.shape[0] / dt / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"all {rays.shape[0]} rays of the workload, closest hit, f64, "
                  f"{threads} threads; BVH build excluded; {dt:.2f} s",
        "host": host_description(),
    }


# Algorithmic bytes of the path tracer's trace kernels, per traced ray
# (the B model): the ray's path state and the record it writes, plus what
# its traversal fetches as the profiling kernels count it — 64-byte nodes,
# 48-byte triangle records, 80 bytes per instance entered (instance and
# mesh records) — and every loose object (64-byte triangles, 32-byte
# spheres, 128-byte quadrics), which each ray tests.
EXTEND_STATE_BYTES = 4 + 32 + 52   # queue entry; origin, direction; hit normal/t, material, two origins
SHADOW_STATE_BYTES = 4 + 48 + 32   # queue entry; origin, direction, contribution; radiance read + write
HBM_PEAK = 8000.0  # GB/s
MFMA_BF16_PEAK = 2500.0  # dense TFLOP/s
NRC_FWD_FLOP = 2 * (48 * 64 + 4 * 64 * 64 + 64 * 3)  # 39,296 per query
NRC_TRAIN_FLOP = 3 * NRC_FWD_FLOP                   # forward, input gradients, weight gradients
NRC_QUERY_BYTES = 64 + 16 * 8 * 8 + 16 + 32          # record; 16 levels x 8 corners x 2 f32; throughput; radiance


def trace_roofline(prof, scene):
    """B-model roofline of the extend and shadow kernels from a profile
    (rrhip_session_get_profile) of the scene description."""
    loose = 64 * scene.num_triangles + 32 * scene.num_spheres + 128 * scene.num_quadrics

    def model(kind, state):
        rays = prof[f"{kind}_rays"]
        b = (state + loose) * rays + 64 * prof[f"{kind}_nodes"] + 48 * prof[f"{kind}_triangles"] + \
            80 * prof[f"{kind}_instances"]
        ms = prof[f"{kind}_ms"]
        return {"rays": int(rays), "bytes_per_ray": round(b / max(rays, 1), 1),
                "node_fetches_per_ray": round(prof[f"{kind}_nodes"] / max(rays, 1), 2),
                "triangle_tests_per_ray": round(prof[f"{kind}_triangles"] / max(rays, 1), 2),
                "instances_per_ray": round(prof[f"{kind}_instances"] / max(rays, 1), 2),
                "ms": round(ms, 3), "gbs": round(b / (ms * 1e6), 1) if ms > 0 else None,
                "frac": round(b / (ms * 1e6) / HBM_PEAK, 4) if ms > 0 else None, "bytes": b}
    ext, sh = model("extend", EXTEND_STATE_BYTES), model("shadow", SHADOW_STATE_BYTES)
    total_b, total_ms = ext.pop("bytes") + sh.pop("bytes"), prof["extend_ms"] + prof["shadow_ms"]
    achieved = total_b / (total_ms * 1e6)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4), "traffic": None,
            "kernels": {"extend": ext, "shadow": sh, "shade_ms": round(prof["shade_ms"], 3)},
            "note": "extend + shadow kernels over one profiled frame: B-model bytes (path state, node / "
                    "triangle / instance fetches counted by the profiling kernels, loose objects) over "
                    "their event-timed device time"}


def profiled_frame(ps, params, start, count, nrc_config=None, ctx=None):
    """One profiled (untimed) render of samples [start, start + count)."""
    import rrhip
    se = ps.session(params)
    cache = None
    if nrc_config is not None:
        cache = rrhip.Nrc(ctx, nrc_config)
        se.set_nrc(cache)
        se.render(0, start)  # the cache at the state the timed runs reach there
    se.set_profiling(True)
    se.render(start, count)
    prof = se.profile()
    se.close()
    if cache is not None:
        cache.close()
    return prof


def box_room(ctx, args, with_cpu):
    """The box-room path-tracing workload (256x256, 16 spp, depth 8): one
    frame = rrhip_render_samples over 16 samples of every pixel."""
    import rrhip
    import scenes
    s = scenes.box_room(256, 256)
    d = s.desc()
    p = rrhip.RenderParams.make(max_depth=8)
    ps = rrhip.PtScene(ctx, d)
    se = ps.session(p)
    se.render(0, 16)  # warm-up frame
    st0 = se.stats()
    frames = max(3, args.steps // 4)
    t0 = time.perf_counter()
    for f in range(frames):
        se.render(16 * (f + 1), 16)
    dt = (time.perf_counter() - t0) / frames
    st1 = se.stats()
    rays = (st1["extension_rays"] + st1["shadow_rays"] - st0["extension_rays"] - st0["shadow_rays"]) / frames
    out = {
        "workload": "box room 256x256, 16 spp, max depth 8, NEE + MIS + RR (wavefront)",
        "ms_per_frame": round(dt * 1e3, 3),
        "device_ms_per_frame": round(st1["last_render_ms"], 3),
        "samples_per_s": round(256 * 256 * 16 / dt, 1),
        "mrays_per_s": round(rays / dt / 1e6, 1),
        "rays_per_frame": int(rays),
    }
    se.close()
    out["roofline"] = trace_roofline(profiled_frame(ps, p, 0, 16), d)
    ps.close()
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        threads = host_threads()
        t0 = time.perf_counter()
        _, crays = orc.render(d, p, 0, 16, threads=threads)
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"samples_per_s": round(256 * 256 * 16 / cdt, 1),
                               "mrays_per_s": round(crays / cdt / 1e6, 3), "cores": threads,
                               "kind": "port", "sample": f"one full frame, f64, {cdt:.2f} s"}
    return out


def instanced_pt(ctx, args, with_cpu):
    """The instanced path-tracing workload at its configured size: 10,000
    instances of a 20,480-triangle displaced icosphere (64 point lights
    through the light BVH + one rect light), 1920x1080, 64 spp, depth 12,
    f32 traversal, f64 accumulators; one frame = rrhip_render_samples over
    64 samples of every pixel in passes of 8 (the launch mode picked by the
    session autotuner, trials discarded). The roofline comes from a
    separate profiled 8-spp render of the same frame's first samples."""
    import rrhip
    import scenes
    W, H, SPP = 1920, 1080, 64
    s = scenes.instanced_field(W, H)
    d = s.desc()
    ps = rrhip.PtScene(ctx, d)
    p = rrhip.RenderParams.make(max_depth=12, samples_per_pass=8)
    se = ps.session(p)
    mode, tune_ms = se.autotune(8)
    se.render(0, 8)  # warm-up pass
    st0 = se.stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    se.render(8, SPP)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    st1 = se.stats()
    se.close()
    rays = st1["extension_rays"] + st1["shadow_rays"] - st0["extension_rays"] - st0["shadow_rays"]
    out = {
        "workload": "instanced field 1920x1080 (10,000 x 20,480-triangle instances, 64 point + 1 rect light), "
                    "64 spp, depth 12, NEE + MIS + RR (wavefront, passes of 8 spp)",
        "ms_per_frame": round(dt * 1e3, 1),
        "device_ms_per_frame": round(st1["last_render_ms"], 1),
        "mrays_per_s": round(rays / dt / 1e6, 1),
        "rays_per_frame": int(rays),
        "samples_per_s": round(W * H * SPP / dt, 1),
        "launch_mode": mode,
        "roofline": trace_roofline(profiled_frame(ps, p, 0, 8), d),
    }
    if with_cpu:
        # the oracle's plain path tracer on the same scene at 480x270 and
        # 2 spp, scaled by pixels and samples to the 64-spp 1080p frame;
        # setup (scene load, BVH builds) measured on an 8x8 render and
        # subtracted
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        threads = host_threads()
        small, tiny = scenes.instanced_field(480, 270), scenes.instanced_field(8, 8)
        t1 = time.perf_counter()
        orc.render(tiny.desc(), p, 0, 1, threads=threads)
        setup = time.perf_counter() - t1
        t1 = time.perf_counter()
        _, crays = orc.render(small.desc(), p, 0, 2, threads=threads)
        cdt = max(time.perf_counter() - t1 - setup, 1e-6)
        frame_s = cdt * (W * H) / (480 * 270) * SPP / 2
        out["cpu_baseline"] = {
            "value": round(crays / cdt / 1e6, 3), "unit": "Mrays/s", "frame_s": round(frame_s, 1),
            "cores": threads, "kind": "port",
            "sample": f"oracle plain PT, same scene at 480x270 x 2 spp ({cdt:.2f} s after {setup:.2f} s setup), "
                      f"scaled by pixels and samples to {frame_s:.0f} s per 64-spp 1080p frame",
        }
    ps.close()
    return out


def hair_curves(ctx, args, with_cpu):
    """Curves and hair fibers at 1920x1080: a 100,000-strand fur ball
    (300,000 rounded-cone segments in one curve BLAS, render.rs's fiber
    BSDF, a distant light over a ground), 16 spp, depth 8; one frame =
    rrhip_render_samples over 16 samples of every pixel in passes of 8.
    The B-model roofline counts segment records (48 B) where the triangle
    column says triangles."""
    import rrhip
    import scenes
    W, H, SPP = 1920, 1080, 16
    s = scenes.fur_ball(W, H)
    d = s.desc()
    ps = rrhip.PtScene(ctx, d)
    p = rrhip.RenderParams.make(max_depth=8, samples_per_pass=8)
    se = ps.session(p)
    se.render(0, 8)  # warm-up pass
    st0 = se.stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    se.render(8, SPP)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    st1 = se.stats()
    se.close()
    rays = st1["extension_rays"] + st1["shadow_rays"] - st0["extension_rays"] - st0["shadow_rays"]
    segs = int(d.curve_sets[0].num_segments)
    out = {
        "workload": f"fur ball 1920x1080 ({segs // 3:,} strands, {segs:,} rounded-cone segments, hair fiber "
                    "BSDF, distant light, ground), 16 spp, depth 8, NEE + MIS + RR (wavefront, passes of 8 spp)",
        "ms_per_frame": round(dt * 1e3, 1),
        "mrays_per_s": round(rays / dt / 1e6, 1),
        "rays_per_frame": int(rays),
        "samples_per_s": round(W * H * SPP / dt, 1),
        "roofline": trace_roofline(profiled_frame(ps, p, 0, 8), d),
    }
    if with_cpu:
        # the oracle on the same groom at 240x135 and 2 spp, scaled by
        # pixels and samples to the frame; setup (scene load, the curve
        # BLAS build) measured on an 8x8 render and subtracted
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        threads = host_threads()
        small, tiny = scenes.fur_ball(240, 135), scenes.fur_ball(8, 8)
        t1 = time.perf_counter()
        orc.render(tiny.desc(), p, 0, 1, threads=threads)
        setup = time.perf_counter() - t1
        t1 = time.perf_counter()
        _, crays = orc.render(small.desc(), p, 0, 2, threads=threads)
        cdt = max(time.perf_counter() - t1 - setup, 1e-6)
        frame_s = cdt * (W * H) / (240 * 135) * SPP / 2
        out["cpu_baseline"] = {
            "value": round(crays / cdt / 1e6, 3), "unit": "Mrays/s", "frame_s": round(frame_s, 1),
            "cores": threads, "kind": "port",
            "sample": f"oracle PT, same groom at 240x135 x 2 spp ({cdt:.2f} s after {setup:.2f} s setup), "
                      f"scaled by pixels and samples to {frame_s:.1f} s per 16-spp 1080p frame",
        }
    ps.close()
    return out


def textured(ctx, args, with_cpu):
    """Textures and the pattern graph at 1920x1080: the pattern zoo (every
    node kind; mesh, quadric, sphere and curve texture coordinates; EWA on
    a grazing ground; UDIM tiles), 16 spp, depth 8, and the same scene with
    its bindings removed (constant materials) to price the patterns: both
    frames run the same trace kernels, so the difference is st at the hits
    and the pattern evaluation in the shade kernel."""
    import rrhip
    import scenes
    W, H, SPP = 1920, 1080, 16
    p = rrhip.RenderParams.make(max_depth=8, samples_per_pass=8)

    def frame(s):
        d = s.desc()
        ps = rrhip.PtScene(ctx, d)
        se = ps.session(p)
        se.render(0, 8)  # warm-up pass
        st0 = se.stats()
        ctx.synchronize()
        t0 = time.perf_counter()
        se.render(8, SPP)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        st1 = se.stats()
        se.close()
        rays = st1["extension_rays"] + st1["shadow_rays"] - st0["extension_rays"] - st0["shadow_rays"]
        return ps, d, dt, rays

    ps, d, dt, rays = frame(scenes.pattern_zoo(W, H))
    out = {
        "workload": "pattern zoo 1920x1080 (textures with trilinear and EWA lookups, checker, fractal, mix, "
                    "colour correct, ramp, triplanar, UDIM; 7 bound materials), 16 spp, depth 8, NEE + MIS + RR "
                    "(wavefront, passes of 8 spp)",
        "ms_per_frame": round(dt * 1e3, 1),
        "mrays_per_s": round(rays / dt / 1e6, 1),
        "rays_per_frame": int(rays),
        "roofline": trace_roofline(profiled_frame(ps, p, 0, 8), d),
    }
    ps.close()
    plain = scenes.pattern_zoo(W, H)
    plain.bindings = []
    ps2, _, dt2, rays2 = frame(plain)
    ps2.close()
    out["unbound"] = {"ms_per_frame": round(dt2 * 1e3, 1), "mrays_per_s": round(rays2 / dt2 / 1e6, 1),
                      "note": "the same scene with every binding removed (constant materials)"}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        threads = host_threads()
        t1 = time.perf_counter()
        _, crays = orc.render(scenes.pattern_zoo(W, H).desc(), p, 8, SPP, threads=threads)
        cdt = max(time.perf_counter() - t1, 1e-6)
        out["cpu_baseline"] = {
            "value": round(crays / cdt / 1e6, 3), "unit": "Mrays/s", "frame_s": round(cdt, 2),
            "cores": threads, "kind": "port",
            "sample": f"oracle PT, the whole 16-spp 1080p frame ({cdt:.2f} s including the scene load)",
        }
    return out


def cache_kernels(prof, frames_profiled=1):
    """Per-frame radiance-cache kernel fractions from a profiled frame:
    inference against the dense bf16 MFMA peak (39,296 FLOP per query) and
    its hash-grid gathers against HBM (1,168 B per query: record, 16 levels
    x 8 corners x 2 f32, throughput, radiance update); the optimiser steps
    (117,888 FLOP per sample: forward, input and weight gradients)."""
    q, n = prof["cache_queries"], prof["cache_samples"]
    infer_s, steps_s = prof["cache_infer_ms"] / 1e3, prof["cache_steps_ms"] / 1e3
    out = {"queries": int(q), "training_samples": int(n)}
    if infer_s > 0:
        out["inference"] = {"ms": round(prof["cache_infer_ms"], 3),
                            "mfma_tflops": round(q * NRC_FWD_FLOP / infer_s / 1e12, 2),
                            "mfma_frac": round(q * NRC_FWD_FLOP / infer_s / 1e12 / MFMA_BF16_PEAK, 5),
                            "gather_gbs": round(q * NRC_QUERY_BYTES / infer_s / 1e9, 1),
                            "gather_frac": round(q * NRC_QUERY_BYTES / infer_s / 1e9 / HBM_PEAK, 4)}
    if steps_s > 0:
        out["training"] = {"ms": round(prof["cache_steps_ms"], 3),
                           "mfma_tflops": round(n * NRC_TRAIN_FLOP / steps_s / 1e12, 3),
                           "mfma_frac": round(n * NRC_TRAIN_FLOP / steps_s / 1e12 / MFMA_BF16_PEAK, 6),
                           "note": "the steps run on the cache's stream, overlapped with the next frame's tracing"}
    return out


def time_to_relmse(ctx, args, with_cpu):
    """Time to relMSE <= 0.01 at 1080p, radiance cache vs plain wavefront
    path tracing, on the instanced workload (10,000 instances of a 20,480-
    triangle displaced icosphere, 64 point lights + 1 rect light, depth 12).
    Both render 4 spp per frame from frame 0 into a running mean; after
    each frame the relMSE of the mean against a 4096-spp plain render
    (independent sample streams) is taken on the device, which finds the
    frame count F at which relMSE first reaches 0.01. The time is then
    that of a second, identical run of F frames back to back (fresh
    session and cache, same seeds, so the image is bit-identical: its
    relMSE is checked equal), wall clock from the first render call until
    the last frame's work, the cache's last optimiser steps included, is
    done. The curve's seconds are the first run's summed render calls."""
    import rrhip
    import scenes
    W, H = 1920, 1080
    s = scenes.instanced_field(W, H)
    d = s.desc()
    ps = rrhip.PtScene(ctx, d)
    params = rrhip.RenderParams.make(max_depth=12, samples_per_pass=4)
    t0 = time.perf_counter()
    ref = ps.session(rrhip.RenderParams.make(max_depth=12, samples_per_pass=4, sample_base=1 << 32))
    for k in range(0, args.ref_spp, 64):
        ref.render(k, min(64, args.ref_spp - k))
    ref_s = time.perf_counter() - t0

    def start(nrc_config):
        se = ps.session(params)
        cache = None
        if nrc_config is not None:
            cache = rrhip.Nrc(ctx, nrc_config)
            se.set_nrc(cache)
        return se, cache

    def run(nrc_config):
        se, cache = start(nrc_config)
        spent, frames, r = 0.0, 0, float("inf")
        curve = []
        while frames < args.max_frames:
            t1 = time.perf_counter()
            se.render(4 * frames, 4)
            spent += time.perf_counter() - t1
            frames += 1
            r = se.relmse(ref, 0.01)
            if frames in (1, 4, 16, 64, 128, 256):
                curve.append([frames, round(spent, 3), round(r, 5)])
            if r <= 0.01:
                break
        st = se.stats()
        out = {"seconds": None, "frames": frames, "spp": 4 * frames, "relmse": round(r, 5)}
        extra = {}
        if cache is not None:
            info = cache.info()
            extra["optimizer_steps"] = info["steps"]
            extra["cache_queries_per_frame"] = int(st["cache_queries"] / frames)
            extra["training_samples_per_frame"] = int(st["cache_samples"] / frames)
            cache.close()
        se.close()
        # the timed replay: the same frames back to back
        se, cache = start(nrc_config)
        ctx.synchronize()
        t1 = time.perf_counter()
        for f in range(frames):
            se.render(4 * f, 4)
        if cache is not None:
            cache.info()  # waits for the last pass's optimiser steps
        ctx.synchronize()
        wall = time.perf_counter() - t1
        r2 = se.relmse(ref, 0.01)
        if cache is not None:
            cache.close()
        se.close()
        if r <= 0.01 and r2 == r:
            out["seconds"] = round(wall, 3)
        out["replay_relmse_equal"] = r2 == r
        out["ms_per_frame"] = round(wall / frames * 1e3, 2)
        out["mrays_per_frame"] = round((st["extension_rays"] + st["shadow_rays"]) / frames / 1e6, 2)
        out["curve [frames, s, relMSE]"] = curve
        out.update(extra)
        return out

    plain = run(None)
    nrc = run(rrhip.NrcConfig.make())
    ref.close()
    # rooflines from profiled (untimed) frames: plain frame 0; the cached
    # frame after 16 frames of training
    plain["roofline"] = trace_roofline(profiled_frame(ps, params, 0, 4), d)
    pn = profiled_frame(ps, params, 64, 4, nrc_config=rrhip.NrcConfig.make(), ctx=ctx)
    nrc["roofline"] = trace_roofline(pn, d)
    nrc["cache_kernels"] = cache_kernels(pn)
    out = {
        "workload": "instanced field 1920x1080 (10,000 x 20,480-triangle instances, 64 point + 1 rect light), "
                    "4 spp/frame, depth 12; relMSE eps 0.01 vs a 4096-spp plain render",
        "reference_spp": args.ref_spp, "reference_render_s": round(ref_s, 2),
        "plain": plain, "nrc": nrc,
        "nrc_config": "16 levels x 2 features x 2^19, 48-64x5-3 MLP bf16 on MFMA, Adam 1e-2, 16 x 4096 samples "
                      "and 65536 training paths per frame (steps overlap the next frame), area-spread c = 1e-4",
        "speedup": (round(plain["seconds"] / nrc["seconds"], 3)
                    if plain["seconds"] and nrc["seconds"] else None),
    }
    if with_cpu and plain["seconds"]:
        # The oracle's plain path tracer on the same scene at 480x270 and
        # 4 spp, scaled by pixels to one 1080p frame; setup (scene load and
        # BVH builds) measured on an 8x8 render and subtracted.
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        threads = host_threads()
        small_scene = scenes.instanced_field(480, 270)  # keeps the arrays the desc points at alive
        tiny_scene = scenes.instanced_field(8, 8)
        t1 = time.perf_counter()
        orc.render(tiny_scene.desc(), params, 0, 1, threads=threads)
        setup = time.perf_counter() - t1
        t1 = time.perf_counter()
        _, crays = orc.render(small_scene.desc(), params, 0, 4, threads=threads)
        cdt = max(time.perf_counter() - t1 - setup, 1e-6)
        frame_s = cdt * (W * H) / (480 * 270)
        out["cpu_baseline"] = {
            "value": round(frame_s * plain["frames"], 1), "unit": "s (plain PT to relMSE 0.01, extrapolated)",
            "cores": threads, "kind": "port",
            "sample": f"oracle plain PT, same scene at 480x270 x 4 spp ({cdt:.2f} s after {setup:.2f} s setup, "
                      f"{crays / cdt / 1e6:.2f} Mrays/s), scaled by pixel count to {frame_s:.2f} s per 1080p "
                      f"frame x the {plain['frames']} frames the GPU plain run needed",
        }
    ps.close()
    return out


def batched_jobs(ctx, args, with_cpu):
    """32 heterogeneous room jobs (1-40 spheres, 64^2..512^2, 4-64 spp,
    every other one filled with a 128^3 fBm cloud) rendered (a) one after
    another, each by its own session, (b) as one batch whose passes
    advance every unfinished job through a single launch sequence, and (c)
    in one persistent launch. Completion times are per job from the start
    of the whole workload."""
    import numpy as np
    import rrhip
    import scenes
    jobs = scenes.batch_rooms(32, seed=1)
    pss = [rrhip.PtScene(ctx, s.desc()) for s, _, _ in jobs]
    prm = [rrhip.RenderParams.make(max_depth=d) for _, _, d in jobs]
    # warm-up (first launches of each kernel)
    w = pss[0].session(prm[0])
    w.render(0, 1)
    w.close()
    seq_done, elapsed, seq_rays = [], 0.0, 0
    for ps, p, (_, spp, _) in zip(pss, prm, jobs):
        se = ps.session(p)
        t0 = time.perf_counter()
        se.render(0, spp)
        elapsed += time.perf_counter() - t0
        seq_done.append(elapsed * 1e3)
        st = se.stats()
        seq_rays += st["extension_rays"] + st["shadow_rays"]
        se.close()
    b = rrhip.Batch(ctx, [(ps, spp, p) for ps, p, (_, spp, _) in zip(pss, prm, jobs)])
    b.render()  # warm-up
    t0 = time.perf_counter()
    job_ms, total_ms = b.render()
    wall_ms = (time.perf_counter() - t0) * 1e3
    bst = b.stats()
    brays = bst["extension_rays"] + bst["shadow_rays"]
    # (c) every job in one persistent launch
    b.render_persistent()  # warm-up
    t0 = time.perf_counter()
    pjob_ms, ptotal_ms = b.render_persistent()
    pwall_ms = (time.perf_counter() - t0) * 1e3
    pst = b.stats()
    prays = pst["extension_rays"] + pst["shadow_rays"]

    def pct(v, q):
        return round(float(np.percentile(v, q)), 2)

    out = {
        "workload": "32 box rooms: 1-40 spheres r~U(0.05,0.3), 64^2-512^2, 4-64 spp, half with a 128^3 fBm cloud "
                    "(delta tracking, HG g 0.3, depth 64), others depth 8",
        "samples": int(sum(s.npix * spp for s, spp, _ in jobs)),
        "sequential": {"total_ms": round(elapsed * 1e3, 2), "mrays_per_s": round(seq_rays / elapsed / 1e6, 1),
                       "p50_ms": pct(seq_done, 50), "p99_ms": pct(seq_done, 99)},
        "batched": {"total_ms": round(wall_ms, 2), "device_ms": round(total_ms, 2),
                    "mrays_per_s": round(brays / (wall_ms / 1e3) / 1e6, 1), "p50_ms": pct(job_ms, 50),
                    "p99_ms": pct(job_ms, 99), "passes": bst["passes"], "samples_per_pass": bst["samples_per_pass"]},
        "persistent": {"total_ms": round(pwall_ms, 2), "device_ms": round(ptotal_ms, 2),
                       "mrays_per_s": round(prays / (pwall_ms / 1e3) / 1e6, 1), "p50_ms": pct(pjob_ms, 50),
                       "p99_ms": pct(pjob_ms, 99), "launches": 1,
                       "note": "one persistent launch: lanes pull (job, pixel) items shortest job first and "
                               "regenerate paths in place; job completion from the device clock"},
        "rays": int(brays),
    }
    b.close()
    if with_cpu:
        # The oracle renders every job at 1 spp; the sequential CPU time is
        # that sum scaled by each job's spp (cost is linear in samples).
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        threads = host_threads()
        est, sample_s = 0.0, 0.0
        for (s, spp, d) in jobs:
            t1 = time.perf_counter()
            orc.render(s.desc(), rrhip.RenderParams.make(max_depth=d), 0, 1, threads=threads)
            dt = time.perf_counter() - t1
            sample_s += dt
            est += dt * spp
        out["cpu_baseline"] = {"value": round(est, 1), "unit": "s (32 sequential oracle renders, extrapolated)",
                               "cores": threads, "kind": "port",
                               "sample": f"every job at 1 spp ({sample_s:.1f} s), scaled by its spp"}
    for ps in pss:
        ps.close()
    return out


def copy_calibration(ctx, args):
    """K launches of the plain copy kernel over 512 MiB (reads and writes
    exactly that much per launch)."""
    nbytes = 512 << 20
    src = ctx.alloc(nbytes)
    dst = ctx.alloc(nbytes)
    for _ in range(args.warmup):
        ctx.device_copy(dst, src, nbytes)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.device_copy(dst, src, nbytes)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"mode": "copy", "bytes_per_launch": nbytes, "ms_per_launch": round(dt * 1e3, 4),
                      "GB/s (read+write)": round(2 * nbytes / dt / 1e9, 1)}), flush=True)
    ctx.free(src)
    ctx.free(dst)
    ctx.close()


def nrc_steps(ctx, args):
    """K optimiser steps of the radiance cache (default configuration,
    bounds [0, 1]^3) on one fixed batch of --nrc-batch random samples
    staged on the device once: the step chain the time-to-relMSE frames
    run (training, grid-gradient scatter, MLP reduction + Adam, grid Adam,
    bf16 pack), timed with HIP events on the cache's stream."""
    import numpy as np
    import rrhip
    n = args.nrc_batch
    rng = np.random.default_rng(args.seed)
    q = np.zeros((n, 16), np.float32)
    q[:, 0:3] = rng.random((n, 3))
    q[:, 3] = rng.random(n)
    for a in (4, 8):
        d = rng.normal(size=(n, 3))
        q[:, a:a + 3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    q[:, 12:15] = rng.random((n, 3))
    t = rng.random((n, 3)).astype(np.float32)
    c = rrhip.Nrc(ctx, rrhip.NrcConfig.make(lo=(0, 0, 0), hi=(1, 1, 1)))
    for _ in range(args.warmup):
        c.train_step(q, t)
    ms = c.time_steps(q, t, args.steps)
    flop = n * 3 * 2 * (48 * 64 + 4 * 64 * 64 + 64 * 3)
    print(json.dumps({"mode": "nrc-step", "samples_per_step": n, "steps": args.steps,
                      "us_per_step": round(ms / args.steps * 1e3, 1),
                      "mlp_tflops": round(flop / (ms / args.steps / 1e3) / 1e12, 2)}), flush=True)
    c.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tuning", default=None,
                    help="block,stack,rays_per_wave[,width] instead of the autotuner (profiling runs)")
    ap.add_argument("--no-cost", action="store_true",
                    help="skip the counting pass (profiling runs)")
    ap.add_argument("--no-pt", action="store_true", help="skip the box-room path-tracing line")
    ap.add_argument("--no-relmse", action="store_true", help="skip the time-to-relMSE comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-jobs comparison")
    ap.add_argument("--no-instanced", action="store_true", help="skip the 64-spp instanced-field frame")
    ap.add_argument("--no-hair", action="store_true", help="skip the fur-ball (curves and hair) frame")
    ap.add_argument("--no-textured", action="store_true", help="skip the textured (pattern zoo) frame")
    ap.add_argument("--ref-spp", type=int, default=4096, help="spp of the relMSE reference render")
    ap.add_argument("--max-frames", type=int, default=400, help="frame cap of the time-to-relMSE runs")
    ap.add_argument("--mode", choices=("trace", "copy", "nrc-step"), default="trace",
                    help="copy: time the plain copy kernel that calibrates HBM counters; nrc-step: time the "
                         "radiance cache's optimiser steps alone (profiling runs)")
    ap.add_argument("--nrc-batch", type=int, default=4096, help="samples per step of --mode nrc-step")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the ray-query workload is replicated-BVH / "
              "split-rays data parallel with no exchange, and this build runs on a single "
              "MI355X. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import rrhip
    ctx = rrhip.Context(0)
    if args.mode == "copy":
        copy_calibration(ctx, args)
        return
    if args.mode == "nrc-step":
        nrc_steps(ctx, args)
        return
    pos, idx = rrhip.gen_heightfield(GRID[0], GRID[1], args.seed)
    rays = rrhip.gen_query_rays(pos.min(0), pos.max(0), 0.1, args.seed, N_RAYS)
    scene = ctx.upload_mesh(pos, idx, bins=32, max_leaf=4)
    info = scene.info()
    d_rays = ctx.to_device(rays)
    d_hits = ctx.alloc(N_RAYS * 16)
    d_occ = ctx.alloc(N_RAYS)
    # Launch configuration from the autotuner (setup, not timed); the
    # any-hit kernel is tuned on its own before it is timed.
    if args.tuning:
        vals = [int(x) for x in args.tuning.split(",")]
        scene.set_tuning(*vals)  # block, stack, rays_per_wave[, width]
        tuning = scene.get_tuning()
    else:
        tuning, _ = scene.autotune(d_rays, N_RAYS, d_hits)

    for _ in range(args.warmup):
        scene.trace_closest_device(d_rays, N_RAYS, d_hits)
    ctx.synchronize()
    t0 = time.perf_counter()
    ev_ms = scene.time_trace_device(d_rays, N_RAYS, d_hits, args.steps)
    ctx.synchronize()
    wall_s = time.perf_counter() - t0
    ms_per_step = wall_s * 1e3 / args.steps
    kernel_ms = ev_ms / args.steps
    value = N_RAYS / (ms_per_step / 1e3) / 1e6

    tuning_any = scene.get_tuning()
    if not args.tuning:
        tuning_any, _ = scene.autotune(d_rays, N_RAYS, d_occ, any_hit=True)
    for _ in range(args.warmup):
        scene.trace_any_device(d_rays, N_RAYS, d_occ)
    any_ms = scene.time_trace_device(d_rays, N_RAYS, d_occ, args.steps, any_hit=True) / args.steps

    # PCIe-inclusive rate of the host-buffer entry point (rays in, hits out
    # through pageable host memory); reported beside, never as, the value.
    scene.trace_closest(rays)
    t0 = time.perf_counter()
    for _ in range(3):
        scene.trace_closest(rays)
    host_mrays = N_RAYS * 3 / (time.perf_counter() - t0) / 1e6

    if args.no_cost:
        nodes, tris = 0, 0
    else:
        nodes, tris = scene.trace_cost(rays)
    node_bytes = 128 if tuning.get("width", 2) == 4 else 64  # Node4 or Node2 per visit
    alg_bytes = N_RAYS * (32 + 16) + node_bytes * nodes + 48 * tris
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9
    # Counter figures of the committed profile, used only when it was taken
    # on this workload at the launch configuration timed here (the file
    # holds one entry per profiled configuration).
    traffic, limiter = None, None
    if os.path.exists(TRAFFIC_FILE):
        want = {k: tuning.get(k) for k in ("block", "stack", "rays_per_wave", "width")}
        for t in json.load(open(TRAFFIC_FILE)).get("closest_hit", []):
            if t.get("rays") == N_RAYS and t.get("triangles") == idx.shape[0] and t.get("tuning") == want:
                traffic = t.get("hbm_bytes_per_launch")
                limiter = {"resource": "VALU issue", "frac": t.get("valu_issue_frac"),
                           "how": "SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x kernel cycles)",
                           "fabric_gbs": round(traffic / (t["avg_duration_us"] * 1e3), 1),
                           "l2_hit_rate": t.get("l2_hit_rate"), "source": t.get("source")}

    out = {
        "metric": "closest-hit Mrays/s",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 re-test of candidate triangles)",
        "data": "synthetic",
        "config": {
            "workload": "ray query: 1,048,576-triangle fBm heightfield, 4,194,304 rays, closest hit",
            "rays": N_RAYS,
            "triangles": int(idx.shape[0]),
            "bvh": f"binned SAH 32 bins, leaf<=4, {info['num_nodes']} x 64B nodes, depth {info['depth']} "
                   f"(collapsed four-wide: {info['depth4']} levels of 128B nodes)",
            "parallelism": "single GPU",
            "tuning": tuning,
        },
        "kernel_ms": round(kernel_ms, 4),
        "any_hit_mrays_s": round(N_RAYS / (any_ms / 1e3) / 1e6, 2),
        "any_hit_tuning": tuning_any,
        "host_buffers_mrays_s": round(host_mrays, 1),
        "traversal": {"node_bytes": node_bytes, "node_fetches_per_ray": round(nodes / N_RAYS, 3),
                      "triangle_fetches_per_ray": round(tris / N_RAYS, 3),
                      "algorithmic_bytes_per_ray": round(alg_bytes / N_RAYS, 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "note": "achieved = algorithmic node/triangle/ray bytes per launch / kernel time; "
                             "the binding resource measured by counters is in 'limiter'"},
        "limiter": limiter,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pos, idx, rays)
    if not args.no_pt:
        out["path_tracing"] = box_room(ctx, args, not args.no_cpu_baseline)
    if not args.no_instanced:
        out["instanced_path_tracing"] = instanced_pt(ctx, args, not args.no_cpu_baseline)
    if not args.no_hair:
        out["hair_curves"] = hair_curves(ctx, args, not args.no_cpu_baseline)
    if not args.no_textured:
        out["textured"] = textured(ctx, args, not args.no_cpu_baseline)
    if not args.no_batch:
        out["batched_jobs"] = batched_jobs(ctx, args, not args.no_cpu_baseline)
    if not args.no_relmse:
        out["time_to_relmse"] = time_to_relmse(ctx, args, not args.no_cpu_baseline)
    for p in (d_rays, d_hits, d_occ):
        ctx.free(p)
    scene.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
