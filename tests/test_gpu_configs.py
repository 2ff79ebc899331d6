"""The path-tracing workloads at their configured sizes on the GPU:

* instanced field (10,000 instances of the 20,480-triangle displaced
  icosphere, 64 point lights + 1 rect light, 1920x1080, depth 12): two
  32 x 32 tiles of the full image at 128 spp against the oracle rendering
  the same pixels at their whole-image seeds (independent streams), every
  pixel-channel within 3 sigma (at most 1 % outside) and tile means within
  3 sigma;
* batched jobs (32 rooms, 64^2-512^2, 4-64 spp, every other one with a
  128^3 fBm cloud): each job's accumulator equals its own session's bit
  for bit, and four jobs (two clouds, two surface rooms) match the
  oracle; the same in one persistent launch (those four bit for bit
  against sessions rendering them in one pass, and against the oracle);
* radiance cache on the instanced field at 1920x1080 with the default
  cache (65,536 training paths, 16 optimiser steps of 4,096 samples per
  frame): finite non-negative images, exactly 65,536 training samples per
  frame, bit-reproducible frames, and relMSE against a plain 256-spp
  reference falling as frames accumulate.
"""
import sys

import numpy as np
import pytest

import rrhip
import scenes
from pt_stats import image_mean_z, mean_var, oracle_render, oracle_render_pixels, sigma_test

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def check_parity(acc_gpu, acc_cpu, label, max_frac=0.01):
    z = sigma_test(acc_gpu, acc_cpu)
    frac = float((z > 3).mean())
    zi = image_mean_z(acc_gpu, acc_cpu)
    mg, _, _ = mean_var(acc_gpu)
    mc, _, _ = mean_var(acc_cpu)
    say(f"{label}: outside 3 sigma {frac:.4%}, max z {z.max():.3f}, mean z {np.round(zi, 2)}, "
        f"means gpu {mg.mean(0).round(5)} cpu {mc.mean(0).round(5)}")
    assert frac <= max_frac, f"{label}: {frac:.4%} of pixel-channels outside 3 sigma"
    assert (zi < 3).all(), f"{label}: means differ by {zi} sigma"


def tile(x0, y0, w, h, width):
    ys, xs = np.meshgrid(np.arange(y0, y0 + h), np.arange(x0, x0 + w), indexing="ij")
    return (ys * width + xs).ravel().astype(np.uint32)


def test_instanced_field_full_size_tiles_match_oracle(ctx):
    W, H, SPP = 1920, 1080, 128
    say(f"instanced field full size {W}x{H} (10000 x 20480 triangles), {SPP} spp; two 32x32 tiles vs oracle")
    s = scenes.instanced_field(W, H)
    d = s.desc()
    p = rrhip.RenderParams.make(max_depth=12, samples_per_pass=8)
    ps = rrhip.PtScene(ctx, d)
    se = ps.session(p)
    se.render(0, SPP)
    acc = se.accum()
    st = se.stats()
    se.close()
    ps.close()
    say(f"gpu stats {st}")
    assert np.isfinite(acc).all() and (acc[:, :6] >= 0).all() and (acc[:, 6] == SPP).all()
    for label, (x0, y0) in (("centre tile", (944, 524)), ("lower-left tile", (400, 900))):
        px = tile(x0, y0, 32, 32, W)
        acc_c, rays = oracle_render_pixels(d, rrhip.RenderParams.make(max_depth=12), px, 0, SPP)
        say(f"{label}: oracle rays per sample {rays / px.size / SPP:.2f}")
        check_parity(acc[px], acc_c, label)


def test_batched_rooms_full_workload(ctx):
    say("batched jobs: 32 rooms (64^2-512^2, 4-64 spp, 16 with 128^3 clouds) vs sessions and oracle")
    jobs = scenes.batch_rooms(32, seed=1)
    pss = [rrhip.PtScene(ctx, s.desc()) for s, _, _ in jobs]
    spp_pass = 8
    b = rrhip.Batch(ctx, [(ps, spp, rrhip.RenderParams.make(max_depth=d)) for ps, (_, spp, d) in zip(pss, jobs)],
                    samples_per_pass=spp_pass)
    job_ms, total = b.render()
    st = b.stats()
    say(f"batch: {total:.1f} ms, {st['passes']} passes, job ms p50 {np.percentile(job_ms, 50):.1f}")
    rays = 0
    for j, (ps, (s, spp, d)) in enumerate(zip(pss, jobs)):
        se = ps.session(rrhip.RenderParams.make(max_depth=d, samples_per_pass=spp_pass))
        se.render(0, spp)
        want = se.accum()
        sst = se.stats()
        rays += sst["extension_rays"] + sst["shadow_rays"]
        se.close()
        got = b.accum(j)
        assert got.tobytes() == want.tobytes(), f"job {j} differs from its session"
        assert (got[:, 6] == spp).all()
    assert rays == st["extension_rays"] + st["shadow_rays"]
    clouds = [j for j, (s, _, _) in enumerate(jobs) if s.media]
    rooms = [j for j, (s, _, _) in enumerate(jobs) if not s.media]
    cost = {j: jobs[j][0].npix * jobs[j][1] for j in range(len(jobs))}
    pick = sorted(clouds, key=cost.get)[:2] + sorted(rooms, key=cost.get)[:2]
    oracle = {}
    for j in pick:
        s, spp, d = jobs[j]
        oracle[j], _ = oracle_render(s.desc(), rrhip.RenderParams.make(max_depth=d), 0, spp)
        check_parity(b.accum(j), oracle[j], f"job {j} ({s.camera.width}^2, {spp} spp, "
                                            f"{'cloud' if s.media else 'surfaces'})")
    # the same workload in one persistent launch: every job's sample count,
    # the ray totals, the picked jobs bit for bit against sessions rendering
    # them in one pass, and against the oracle
    say("the 32 rooms in one persistent launch")
    job_ms, total = b.render_persistent()
    pst = b.stats()
    say(f"persistent: {total:.1f} ms, job ms p50 {np.percentile(job_ms, 50):.1f} p99 {np.percentile(job_ms, 99):.1f}")
    assert pst["extension_rays"] + pst["shadow_rays"] > 0
    for j, (s, spp, d) in enumerate(jobs):
        assert (b.accum(j)[:, 6] == spp).all()
    for j in pick:
        s, spp, d = jobs[j]
        se = pss[j].session(rrhip.RenderParams.make(max_depth=d, samples_per_pass=spp))
        se.render(0, spp)
        assert b.accum(j).tobytes() == se.accum().tobytes(), f"job {j}: persistent launch differs from its session"
        se.close()
        check_parity(b.accum(j), oracle[j], f"persistent job {j}")
    b.close()
    for ps in pss:
        ps.close()


def test_radiance_cache_at_1080p(ctx):
    W, H = 1920, 1080
    say("radiance cache on the instanced field at 1920x1080, default cache, 4 spp frames")
    s = scenes.instanced_field(W, H)
    ps = rrhip.PtScene(ctx, s.desc())
    ref = ps.session(rrhip.RenderParams.make(max_depth=12, samples_per_pass=8, sample_base=1 << 32))
    ref.render(0, 256)
    cfg = rrhip.NrcConfig.make()
    per_frame = cfg.batch * cfg.steps_per_pass
    assert per_frame == 65536 and cfg.train_paths == 65536

    def run(frames):
        c = rrhip.Nrc(ctx, cfg)
        se = ps.session(rrhip.RenderParams.make(max_depth=12, samples_per_pass=4))
        se.set_nrc(c)
        curve = []
        for f in range(frames):
            se.render(4 * f, 4)
            curve.append(se.relmse(ref))
        out = (se.accum(), se.stats(), c.info(), curve)
        se.close()
        c.close()
        return out

    acc, st, info, curve = run(8)
    say(f"relMSE by frame {np.round(curve, 4)}; stats {st}; steps {info['steps']}")
    assert np.isfinite(acc).all() and (acc[:, :6] >= 0).all() and (acc[:, 6] == 32).all()
    assert st["cache_samples"] == 8 * per_frame
    assert st["cache_queries"] > 8 * W * H
    assert info["steps"] == 8 * cfg.steps_per_pass
    assert curve[-1] < 0.6 * curve[0] and curve[-1] < curve[3]
    acc2, _, _, curve2 = run(8)
    assert acc2.tobytes() == acc.tobytes() and curve2 == curve
    ref.close()
    ps.close()
