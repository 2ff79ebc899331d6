"""Batched jobs: every job of a batch equals, bit for bit, the same job
rendered by its own session with the batch's samples_per_pass (same
per-path code, streams and accumulation order), whatever the other jobs;
in one persistent launch, the session with samples_per_pass = the job's
spp."""
import sys

import numpy as np
import pytest

import rrhip
import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def mixed_jobs():
    """Six rooms: three with clouds (one with a point light in the cloud:
    two-ray shadow records), one with a hull of ink and a cutout sphere
    (medium stacks and blocker walks); (scene, spp, depth)."""
    rng = np.random.default_rng(3)
    jobs = []
    for j, (res, spp) in enumerate([(32, 5), (48, 12), (16, 3), (40, 7), (24, 9), (32, 1)]):
        spheres = [(tuple(rng.uniform(-0.7, 0.7, 3)), float(rng.uniform(0.05, 0.3))) for _ in range(j * 3 + 1)]
        if j % 2:
            s = scenes.cloud_room(res, res, n=16, seed=j, spheres=spheres)
            if j == 3:  # a point light in the cloud: two-ray shadow records (equiangular)
                s.add_light(scenes.light(rrhip.LIGHT_POINT, intensity=2.0, p=(0.2, 0.1, 0.0)))
            depth = 64
        else:
            s = scenes.box_room(res, res)
            s.spheres = []
            for c, r in spheres:
                s.add_sphere(c, r, 0)
            if j == 4:  # a hull of ink and a cutout sphere: medium stacks and blocker walks
                ink = s.add_medium((1.5, 0.6, 0.2), (0.2, 0.2, 0.2), 0.0)
                s.add_sphere((0.4, -0.5, 0.0), 0.35, s.add_material(scenes.hull(ink)))
                s.add_sphere((-0.3, 0.2, 0.3), 0.25, s.add_material(scenes.matte((0.8, 0.5, 0.3), presence=0.5)))
            depth = 8
        jobs.append((s, spp, depth))
    return jobs


def test_batch_jobs_equal_their_own_sessions(ctx):
    say("batch of 6 mixed rooms (3 with clouds, one with a point light in it) vs 6 sessions")
    jobs = mixed_jobs()
    pss = [rrhip.PtScene(ctx, s.desc()) for s, _, _ in jobs]
    spp_pass = 4
    b = rrhip.Batch(ctx, [(ps, spp, rrhip.RenderParams.make(max_depth=d)) for ps, (_, spp, d) in zip(pss, jobs)],
                    samples_per_pass=spp_pass)
    job_ms, total = b.render()
    st = b.stats()
    say(f"batch: {total:.2f} ms, per job {np.round(job_ms, 2)}, stats {st}")
    assert st["passes"] == 3 and (job_ms <= total + 1e-6).all()
    rays = 0
    for j, (ps, (s, spp, d)) in enumerate(zip(pss, jobs)):
        se = ps.session(rrhip.RenderParams.make(max_depth=d, samples_per_pass=spp_pass))
        se.render(0, spp)
        want = se.accum()
        got = b.accum(j)
        assert got.tobytes() == want.tobytes(), f"job {j} differs"
        sst = se.stats()
        rays += sst["extension_rays"] + sst["shadow_rays"]
        se.close()
    assert rays == st["extension_rays"] + st["shadow_rays"]
    # a second render of the same batch reproduces it
    first = b.accum(1)
    b.render()
    assert b.accum(1).tobytes() == first.tobytes()
    b.close()
    for ps in pss:
        ps.close()


def test_persistent_launch_equals_sessions(ctx):
    say("one persistent launch of the 6 mixed rooms vs sessions rendering each job in one pass")
    jobs = mixed_jobs()
    pss = [rrhip.PtScene(ctx, s.desc()) for s, _, _ in jobs]
    b = rrhip.Batch(ctx, [(ps, spp, rrhip.RenderParams.make(max_depth=d)) for ps, (_, spp, d) in zip(pss, jobs)])
    job_ms, total = b.render_persistent()
    st = b.stats()
    say(f"persistent: {total:.2f} ms, per job {np.round(job_ms, 2)}, rays {st['extension_rays']} + {st['shadow_rays']}")
    assert (job_ms > 0).all() and (job_ms <= total * 1.05 + 0.05).all()
    ext = shd = 0
    for j, (ps, (s, spp, d)) in enumerate(zip(pss, jobs)):
        se = ps.session(rrhip.RenderParams.make(max_depth=d, samples_per_pass=spp))
        se.render(0, spp)
        want = se.accum()
        sst = se.stats()
        ext += sst["extension_rays"]
        shd += sst["shadow_rays"]
        se.close()
        assert b.accum(j).tobytes() == want.tobytes(), f"job {j} differs from its session"
    assert (st["extension_rays"], st["shadow_rays"]) == (ext, shd)
    first = b.accum(3)
    b.render_persistent()
    assert b.accum(3).tobytes() == first.tobytes()
    b.close()
    fixed = rrhip.Batch(ctx, [(pss[0], 2, rrhip.RenderParams.make(max_depth=8, ray_offset=1e-4))])
    with pytest.raises(rrhip.RrhipError) as e:
        fixed.render_persistent()
    assert e.value.status == 5
    fixed.close()
    for ps in pss:
        ps.close()
