"""Film (AOV) mode of the GPU path tracer against the oracle's film
(pt::render_film, src/raytracer/pt/mod.rs:163-276).

Bars: the first-hit AOVs (albedo, normal, depth, id and its coverage) are
functions of the camera sample, which both sides draw from the same PCG32
stream, so they agree to f32 precision except on the few pixels where an
f32 and an f64 camera ray fall on different sides of a silhouette (at most
2 % of the pixels, by at most 2 samples' worth of coverage); the diffuse /
specular split is a Monte Carlo estimate: whole-image channel means within
3 sigma of the oracle's, sigma taken from two independent GPU renders;
diffuse + specular == beauty; a material with no specular-like lobe has
diffuse == beauty bit for bit.
"""
import sys

import numpy as np
import pytest

import orc
import rrhip
import scenes

pytestmark = pytest.mark.gpu

W = H = 32
R = 0.8


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def gpu_film(ctx, scene, params, start, count, mode=0):
    ps = rrhip.PtScene(ctx, scene.desc())
    se = ps.session(params)
    if mode:
        se.set_launch_mode(mode)
    se.set_film(True)
    se.render(start, count)
    film, acc = se.film(), se.accum()
    se.close()
    ps.close()
    return film, acc


def first_hit_agree(g, o, spp, label):
    """albedo / normal / depth / id / coverage equal to f32 precision except
    on silhouette pixels."""
    same_cov = (g[:, 16] == o[:, 16]) & (np.abs(g[:, 17] - o[:, 17]) < 0.5 / spp)
    ok = same_cov & (np.abs(g[:, 9:15] - o[:, 9:15]).max(1) < 2e-4) & \
        (np.abs(g[:, 15] - o[:, 15]) <= 2e-5 * np.maximum(o[:, 15], 1.0))
    bad = ~ok
    say(f"{label}: first-hit AOVs differ on {bad.sum()} of {len(ok)} pixels")
    assert bad.mean() <= 0.02, f"{label}: {bad.sum()} pixels differ"
    # where they differ, only by a silhouette's worth of samples
    assert np.all(np.abs(g[bad, 17] - o[bad, 17]) <= 2.0 / spp + 1e-12) or np.all(g[bad, 16] != o[bad, 16])


def split_agree(ctx, scene, params, spp, o, label):
    """diffuse and specular image means within 3 sigma of the oracle's."""
    g1, _ = gpu_film(ctx, scene, params, 0, spp)
    g2, _ = gpu_film(ctx, scene, rrhip.RenderParams.make(max_depth=params.max_depth, sample_base=1 << 40), 0, spp)
    npx = g1.shape[0]
    for name, sl in (("diffuse", slice(3, 6)), ("specular", slice(6, 9))):
        sig = np.std(g1[:, sl] - g2[:, sl], axis=0) / np.sqrt(npx)  # >= sd of the gpu - oracle mean
        d = np.abs(g1[:, sl].mean(0) - o[:, sl].mean(0))
        z = np.where(sig > 0, d / np.maximum(sig, 1e-300), np.where(d > 1e-9, np.inf, 0.0))
        say(f"{label} {name}: gpu {g1[:, sl].mean(0).round(5)} oracle {o[:, sl].mean(0).round(5)} z {z.round(2)}")
        assert (z < 3).all(), f"{label}: {name} means differ by {z} sigma"
    np.testing.assert_allclose(g1[:, 3:6] + g1[:, 6:9], g1[:, 0:3], rtol=1e-5, atol=1e-7)
    return g1


def test_film_first_hit_aovs_match_oracle(ctx):
    say("film: albedo / normal / depth / id on a sphere in a dome, 32x32, 16 spp")
    m = scenes.material(diffuse_gain=0.8, diffuse_color=(0.5, 0.25, 1.0), specular_ior=1.5,
                        refraction_color=(1.0, 1.0, 1.0), glass_gain=0.25, object_id=7)
    s = scenes.furnace(W, H, mat=m, radius=R)
    p = rrhip.RenderParams.make(max_depth=4)
    g, acc = gpu_film(ctx, s, p, 0, 16)
    o = orc.render_film(s.desc(), p, 0, 16)
    first_hit_agree(g, o, 16, "sphere")
    # the film's beauty is the accumulator's mean image
    np.testing.assert_array_equal(g[:, 0:3], acc[:, 0:3] / acc[:, 6:7])
    assert set(np.unique(g[:, 16])) <= {0.0, 7.0}


def test_film_diffuse_only_material_is_all_diffuse(ctx):
    say("film: matte sphere, diffuse == beauty")
    s = scenes.furnace(W, H, mat=scenes.matte((0.7, 0.6, 0.5)), radius=R)
    g, _ = gpu_film(ctx, s, rrhip.RenderParams.make(max_depth=6), 0, 16)
    np.testing.assert_array_equal(g[:, 3:6], g[:, 0:3])
    assert np.all(g[:, 6:9] == 0.0)


def test_film_metal_split_matches_oracle(ctx):
    say("film: metal sphere in a dome, diffuse / specular split, 64 spp")
    s = scenes.furnace(W, H, mat=scenes.metal((0.9, 0.8, 0.7), 0.2), radius=R)
    p = rrhip.RenderParams.make(max_depth=6)
    o = orc.render_film(s.desc(), p, 0, 64)
    g = split_agree(ctx, s, p, 64, o, "metal")
    assert g[:, 6:9].sum() > 0.2 * g[:, 0:3].sum()  # the metal's reflections: a large share


def test_film_cornell_with_ids(ctx):
    say("film: cornell fixture with object ids, 48x48, 32 spp")
    s = scenes.cornell_fixture(48, 48)
    for i, m in enumerate(s.materials):
        m.id = 10 + i
    p = rrhip.RenderParams.make(max_depth=6)
    o = orc.render_film(s.desc(), p, 0, 32)
    g = split_agree(ctx, s, p, 32, o, "cornell")
    first_hit_agree(g, o, 32, "cornell")
    assert len(np.unique(g[:, 16])) >= 5


@pytest.mark.parametrize("through", ["hull", "cutout"])
def test_film_hull_and_cutout(ctx, through):
    say(f"film: first real hit behind a {through}")
    s = scenes.Scene(W, H, 30.0, eye=(0, 0, 0), look_at=(0, 0, 1))
    s.add_light(scenes.light(rrhip.LIGHT_DOME, intensity=1.0))
    back = s.add_material(scenes.matte((0.5, 0.5, 0.5), object_id=2))
    s.add_polygon([(-9, -9, 8), (9, -9, 8), (9, 9, 8), (-9, 9, 8)], back)
    if through == "hull":
        med = s.add_medium((0.0, 0.0, 0.0), (1e-9, 1e-9, 1e-9), 0.0)
        front = s.add_material(scenes.hull(med))
    else:
        front = s.add_material(scenes.matte((0.9, 0.1, 0.1), presence=0.0, object_id=3))
    s.add_polygon([(-9, -9, 3), (9, -9, 3), (9, 9, 3), (-9, 9, 3)], front)
    p = rrhip.RenderParams.make(max_depth=3)
    g, _ = gpu_film(ctx, s, p, 0, 8)
    o = orc.render_film(s.desc(), p, 0, 8)
    first_hit_agree(g, o, 8, through)
    assert np.all(g[:, 16] == (2 if through == "hull" else 3))


def test_film_medium_first_scatter(ctx):
    """A scatter in a medium is the first real hit (pt/mod.rs:559-570):
    albedo sigma_s / sigma_t, normal towards the viewer, id 0."""
    say("film: cloud room, first-hit AOVs in expectation")
    s = scenes.cloud_room(24, 24, n=32)
    p = rrhip.RenderParams.make(max_depth=8)
    o = orc.render_film(s.desc(), p, 0, 32)
    g1, _ = gpu_film(ctx, s, p, 0, 32)
    g2, _ = gpu_film(ctx, s, rrhip.RenderParams.make(max_depth=8, sample_base=1 << 40), 0, 32)
    npx = g1.shape[0]
    for name, sl in (("albedo", slice(9, 12)), ("normal", slice(12, 15)), ("depth", slice(15, 16)),
                     ("diffuse", slice(3, 6))):
        sig = np.std(g1[:, sl] - g2[:, sl], axis=0) / np.sqrt(npx)
        d = np.abs(g1[:, sl].mean(0) - o[:, sl].mean(0))
        z = d / np.maximum(sig, 1e-12)
        say(f"cloud {name}: gpu {g1[:, sl].mean(0).round(4)} oracle {o[:, sl].mean(0).round(4)} z {z.round(2)}")
        assert (z < 3).all(), f"cloud {name} differs by {z} sigma"


def test_film_graph_mode_and_clearing(ctx):
    say("film: graph launches equal plain launches; load_accum clears the film")
    s = scenes.cornell_fixture(24, 24)
    p = rrhip.RenderParams.make(max_depth=5, samples_per_pass=4)
    a, _ = gpu_film(ctx, s, p, 0, 12, mode=0)
    b, _ = gpu_film(ctx, s, p, 0, 12, mode=1)
    np.testing.assert_array_equal(a, b)
    ps = rrhip.PtScene(ctx, s.desc())
    se = ps.session(p)
    with pytest.raises(rrhip.RrhipError):
        se.film()  # film mode is off
    se.set_film(True)
    se.render(0, 4)
    assert se.film()[:, 0:3].max() > 0
    se.load_accum(None)
    f = se.film()
    assert np.all(f == 0.0)
    se.render(0, 12)
    np.testing.assert_array_equal(se.film(), a)
    se.set_film(False)
    se.close()
    ps.close()


def test_film_rejects_radiance_cache(ctx):
    say("film: not with a radiance cache")
    s = scenes.box_room(32, 32)
    ps = rrhip.PtScene(ctx, s.desc())
    se = ps.session(rrhip.RenderParams.make(max_depth=4))
    cache = rrhip.Nrc(ctx, rrhip.NrcConfig.make(train_paths=1024))
    se.set_film(True)
    with pytest.raises(rrhip.RrhipError):
        se.set_nrc(cache)
    se.set_film(False)
    se.set_nrc(cache)
    with pytest.raises(rrhip.RrhipError):
        se.set_film(True)
    se.set_nrc(None)
    se.close()
    cache.close()
    ps.close()
