"""Curves and hair fibers on the GPU against the oracle (render.rs
src/geometry/curves.rs, src/raytracer/pt/hair.rs, pt/mod.rs:738-800):

* the device fiber BSDF (rrhip_hair_bsdf, the shade kernel's code) against
  the oracle's f64 hair.rs on random fiber-frame queries: f and pdf to f32
  precision, and samples from the same uniforms landing on the same
  direction with the same value and pdf (a few queries may pick another
  lobe where a uniform falls within f32 rounding of a lobe boundary, or
  sit at the azimuth's wrap);
* first hits on every kind of curve geometry (capsule, taper, diced Bezier
  strand, points, a rotated and scaled instance) through film mode: normal,
  depth, albedo, id and coverage per pixel equal the oracle's except on
  silhouette pixels;
* render.rs's hair parity scene (tests/metal_pt_tests.rs:246-313: 400
  strands, 1,200 segments, hair BSDF under a distant light over a ground)
  and the curve zoo (hair fibers and matte curves under a rect light and a
  dome) within 3 sigma per pixel of the oracle (at most 1 % of
  pixel-channels outside, image means within 3 sigma), with error-bounded
  origins and with render.rs's fixed 1e-4 push; render.rs's own bar for
  this scene is a 4 % mean and 0.12 RMSE at 160 spp;
* hair scenes are refused by the radiance cache and by batches.
"""
import sys

import numpy as np
import pytest

import orc
import rrhip
import scenes
from pt_stats import image_mean_z, mean_var, oracle_render, sigma_test

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def parity(g, c, label, max_frac=0.01):
    z = sigma_test(g, c)
    frac = float((z > 3).mean())
    zi = image_mean_z(g, c)
    mg, _, _ = mean_var(g)
    mc, _, _ = mean_var(c)
    say(f"{label}: outside 3 sigma {frac:.4%}, max z {z.max():.2f}, image-mean z {np.round(zi, 2)}, "
        f"means gpu {mg.mean(0).round(5)} cpu {mc.mean(0).round(5)}")
    assert frac <= max_frac, f"{label}: {frac:.4%} of pixel-channels outside 3 sigma"
    assert (zi < 3).all(), f"{label}: image means differ by {zi} sigma"


def gpu_render(ctx, scene, params, start, count):
    ps = rrhip.PtScene(ctx, scene.desc())
    se = ps.session(params)
    se.render(start, count)
    acc = se.accum()
    st = se.stats()
    se.close()
    ps.close()
    return acc, st


HAIR_CASES = [
    dict(beta_m=0.3, beta_n=0.3, color=(0.35, 0.2, 0.06)),   # hair.rs's defaults
    dict(beta_m=0.6, beta_n=0.5, absorption=(0.0, 0.0, 0.0)),  # white, every lobe in the Bessel branch
    dict(beta_m=0.12, beta_n=0.2, color=(0.5, 0.3, 0.12), eta=1.6),  # narrow lobes
]


def random_dirs(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_device_fiber_bsdf_matches_oracle(ctx):
    rng = np.random.default_rng(21)
    n = 20000
    for case in HAIR_CASES:
        m = scenes.hair(**case)
        hp = dict(sigma_a=tuple(m.hair_sigma_a), beta_m=m.hair_beta_m, beta_n=m.hair_beta_n, eta=m.hair_eta)
        say(f"fiber BSDF {case}: {n} queries")
        wo = random_dirs(rng, n).astype(np.float32)
        wi = random_dirs(rng, n).astype(np.float32)
        h = rng.uniform(-1, 1, n).astype(np.float32)
        u = rng.random((n, 4)).astype(np.float32)
        f, pdf, ws, fs, pdfs = ctx.hair_bsdf(m, wo, wi, h, u)
        fo, po = orc.hair_eval(wo.astype(float), wi.astype(float), h.astype(float), **hp)
        scale = np.maximum(np.abs(fo).max(1), 1e-3)
        ef = np.abs(f - fo).max(1) / scale
        ep = np.abs(pdf - po) / np.maximum(po, 1e-3)
        say(f"  f: median rel err {np.median(ef):.2e}, 99.9 % {np.quantile(ef, 0.999):.2e}; "
            f"pdf: median {np.median(ep):.2e}, 99.9 % {np.quantile(ep, 0.999):.2e}")
        assert np.quantile(ef, 0.999) < 1e-3 and np.quantile(ep, 0.999) < 1e-3
        assert (ef < 2e-2).mean() > 0.9999 and (ep < 2e-2).mean() > 0.9999
        wso, fso, pso, oko = orc.hair_sample(wo.astype(float), h.astype(float), u.astype(float), **hp)
        okg = pdfs > 0
        same = okg & oko & (np.sum(ws * wso, 1) > 1 - 1e-5)
        say(f"  samples: {same.mean():.5f} same direction; ok gpu {okg.mean():.5f} oracle {oko.mean():.5f}")
        assert same.mean() > 0.998
        es = np.abs(pdfs[same] - pso[same]) / np.maximum(pso[same], 1e-3)
        assert np.quantile(es, 0.999) < 5e-3


def film_gpu(ctx, scene, params, spp):
    ps = rrhip.PtScene(ctx, scene.desc())
    se = ps.session(params)
    se.set_film(True)
    se.render(0, spp)
    film = se.film()
    se.close()
    ps.close()
    return film


def test_curve_first_hits_match_oracle(ctx):
    W = H = 48
    spp = 8
    say(f"curve zoo {W}x{H} first hits (film), {spp} spp")
    s = scenes.curve_zoo(W, H)
    p = rrhip.RenderParams.make(max_depth=2)
    g = film_gpu(ctx, s, p, spp)
    o = orc.render_film(s.desc(), rrhip.RenderParams.make(max_depth=2, ray_offset=1e-9), 0, spp)
    same_cov = (g[:, 16] == o[:, 16]) & (np.abs(g[:, 17] - o[:, 17]) < 0.5 / spp)
    ok = same_cov & (np.abs(g[:, 9:15] - o[:, 9:15]).max(1) < 2e-3) & \
        (np.abs(g[:, 15] - o[:, 15]) <= 2e-5 * np.maximum(o[:, 15], 1.0))
    bad = ~ok
    hit = o[:, 17] > 0
    say(f"first-hit AOVs differ on {bad.sum()} of {len(ok)} pixels ({hit.sum()} see geometry)")
    assert bad.mean() <= 0.03
    # every curve kind is seen: the ids of the fiber, red and blue materials
    # and the ground lead pixels
    ids = set(np.unique(o[:, 16]).astype(int))
    assert {1, 2, 3, 4} <= ids, ids


@pytest.mark.parametrize("offset", [0.0, 1e-4])
def test_hair_ball_matches_oracle(ctx, offset):
    spp = 128 if offset == 0.0 else 64
    say(f"hair ball 64x64 (400 strands, 1200 segments), {spp} spp, ray_offset {offset}")
    s = scenes.hair_ball(64, 64)
    p = rrhip.RenderParams.make(max_depth=8, ray_offset=offset)
    g, st = gpu_render(ctx, s, p, 0, spp)
    say(f"gpu stats {st}")
    assert np.isfinite(g).all() and (g[:, 6] == spp).all()
    c, _ = oracle_render(s.desc(), p, 0, spp)
    parity(g, c, f"hair ball offset {offset}")


@pytest.mark.parametrize("hair_fiber", [True, False])
def test_curve_zoo_matches_oracle(ctx, hair_fiber):
    spp = 64
    say(f"curve zoo 64x64, {'hair fibers' if hair_fiber else 'matte curves'}, {spp} spp")
    s = scenes.curve_zoo(64, 64, hair_fiber=hair_fiber)
    p = rrhip.RenderParams.make(max_depth=6)
    g, _ = gpu_render(ctx, s, p, 0, spp)
    c, _ = oracle_render(s.desc(), p, 0, spp)
    parity(g, c, f"curve zoo ({'hair' if hair_fiber else 'matte'})")


def test_hair_refused_by_cache_and_batches(ctx):
    s = scenes.hair_ball(16, 16)
    ps = rrhip.PtScene(ctx, s.desc())
    se = ps.session(rrhip.RenderParams.make(max_depth=4))
    c = rrhip.Nrc(ctx, rrhip.NrcConfig.make())
    with pytest.raises(rrhip.RrhipError) as e:
        se.set_nrc(c)
    assert e.value.status == 5  # RRHIP_ERR_UNSUPPORTED
    with pytest.raises(rrhip.RrhipError) as e:
        rrhip.Batch(ctx, [(ps, 4, rrhip.RenderParams.make(max_depth=4))], samples_per_pass=4)
    assert e.value.status == 5  # RRHIP_ERR_UNSUPPORTED
    c.close()
    se.close()
    ps.close()
