"""Session profiling (rrhip_session_set_profiling): counting trace kernels
and per-kernel device times must leave the render bit for bit unchanged,
count what the traversal fetches (every ray visits at least the root of a
non-empty hierarchy; instanced rays enter instances, loose-only scenes
none), and time every bounce's kernels."""
import sys

import numpy as np
import pytest

import rrhip
import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def render(ctx, scene, spp, depth, profile, nrc=None, spp_pass=0):
    ps = rrhip.PtScene(ctx, scene.desc())
    se = ps.session(rrhip.RenderParams.make(max_depth=depth, samples_per_pass=spp_pass))
    cache = rrhip.Nrc(ctx, rrhip.NrcConfig.make(train_paths=4096, batch=1024, steps_per_pass=2)) if nrc else None
    if cache:
        se.set_nrc(cache)
    if profile:
        se.set_profiling(True)
    se.render(0, spp)
    acc = se.accum()
    prof = se.profile() if profile else None
    st = se.stats()
    se.close()
    if cache:
        cache.close()
    ps.close()
    return acc, prof, st


def test_profiling_leaves_the_render_unchanged_and_counts_the_instanced_field(ctx):
    s = scenes.instanced_field(192, 108, n_instances=400)
    say("profiling: instanced field 192x108 (400 instances), 4 spp, depth 12, with and without profiling")
    a0, _, st0 = render(ctx, s, 4, 12, False)
    a1, p, st1 = render(ctx, s, 4, 12, True)
    say(f"profile {p}")
    assert a0.tobytes() == a1.tobytes()
    assert p["extend_rays"] == st1["extension_rays"] and p["shadow_rays"] == st1["shadow_rays"]
    assert p["extend_nodes"] >= p["extend_rays"] and p["shadow_nodes"] > 0
    assert p["extend_instances"] > 0 and p["extend_triangles"] > 0
    assert p["passes"] >= 1 and p["bounce_rounds"] == 12 * p["passes"]
    assert p["extend_ms"] > 0 and p["shade_ms"] > 0 and p["shadow_ms"] > 0
    assert p["cache_infer_ms"] == 0 and p["cache_steps_ms"] == 0


def test_profiling_with_loose_objects_and_a_cache(ctx):
    s = scenes.box_room(64, 64)
    say("profiling: box room 64x64 (loose objects only), 8 spp; then with a radiance cache")
    a0, _, _ = render(ctx, s, 8, 8, False, spp_pass=4)
    a1, p, _ = render(ctx, s, 8, 8, True, spp_pass=4)
    assert a0.tobytes() == a1.tobytes()
    assert p["extend_nodes"] == 0 and p["extend_instances"] == 0 and p["extend_rays"] > 0
    c0, _, _ = render(ctx, s, 8, 8, False, nrc=True, spp_pass=4)
    c1, pc, _ = render(ctx, s, 8, 8, True, nrc=True, spp_pass=4)
    say(f"cache profile {pc}")
    assert c0.tobytes() == c1.tobytes()
    assert pc["cache_queries"] > 0 and pc["cache_infer_ms"] > 0 and pc["cache_steps_ms"] > 0
    ps = rrhip.PtScene(ctx, s.desc())
    se = ps.session(rrhip.RenderParams.make(max_depth=8))
    with pytest.raises(rrhip.RrhipError):
        se.profile()  # profiling is off
    se.close()
    ps.close()
