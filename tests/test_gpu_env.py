"""The dome light with a lat-long environment map (EnvMap,
src/scene/envmap.rs; pt/mod.rs:375-387, 645-654, 1138-1172) on the GPU
against the oracle: renders lit by the reference's own sky.hdr fixture
agree per pixel within 3 sigma (at most 1 % of pixel-channels outside,
image means within 3 sigma); a constant map around an albedo-1 diffuse
sphere is a white furnace (1 within 1e-3), whatever the map's
non-uniform sampling density."""
import os
import sys

import numpy as np
import pytest

import rrhip
import scenes
from pt_stats import image_mean_z, mean_var, oracle_render, sigma_test
import rgbe

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def gpu_render(ctx, scene, params, start, count):
    ps = rrhip.PtScene(ctx, scene.desc())
    se = ps.session(params)
    se.render(start, count)
    acc = se.accum()
    se.close()
    ps.close()
    return acc


def sky_scene(w, h):
    """Three spheres (matte, rough metal, glass) on a grey ground under the
    sky fixture (dome intensity 0.35, as bust.rib / pan.rib use it)."""
    s = scenes.Scene(w, h, 40.0, eye=(0, 1.2, -5), look_at=(0, 0.4, 0))
    s.add_light(scenes.light(rrhip.LIGHT_DOME, intensity=0.35))
    s.set_env(rgbe.read_hdr(os.path.join(HERE, "golden", "sky.hdr")))
    ground = s.add_material(scenes.matte((0.6, 0.6, 0.6)))
    s.add_polygon([(-20, 0, -20), (20, 0, -20), (20, 0, 20), (-20, 0, 20)], ground)
    s.add_sphere((-1.4, 0.6, 0), 0.6, s.add_material(scenes.matte((0.8, 0.3, 0.2))))
    s.add_sphere((0.0, 0.6, 0.4), 0.6, s.add_material(scenes.metal((0.9, 0.85, 0.8), 0.3)))
    s.add_sphere((1.4, 0.6, 0), 0.6, s.add_material(scenes.glass(1.5, 0.1)))
    return s


def test_sky_dome_matches_oracle(ctx):
    say("sky.hdr dome: 64x48, 64 spp, GPU vs oracle")
    s = sky_scene(64, 48)
    p = rrhip.RenderParams.make(max_depth=8)
    g = gpu_render(ctx, s, p, 0, 64)
    c, _ = oracle_render(s.desc(), p, 0, 64)
    z = sigma_test(g, c)
    frac = float((z > 3).mean())
    zi = image_mean_z(g, c)
    mg, _, _ = mean_var(g)
    mc, _, _ = mean_var(c)
    say(f"outside 3 sigma {frac:.4%}, image-mean z {np.round(zi, 2)}, means gpu {mg.mean(0).round(4)} "
        f"cpu {mc.mean(0).round(4)}")
    assert frac <= 0.01 and (zi < 3).all()
    assert mg.mean() > 0.05


def test_constant_map_is_a_white_furnace(ctx):
    say("constant environment map furnace: 48x48, 512 spp")
    s = scenes.furnace(48, 48)
    s.set_env(np.ones((16, 32, 3), np.float32))
    acc = gpu_render(ctx, s, rrhip.RenderParams.make(max_depth=12), 0, 512)
    m, _, _ = mean_var(acc)
    centre = m.reshape(48, 48, 3)[6:42, 6:42].reshape(-1, 3).mean(0)
    say(f"furnace centre {centre}")
    assert np.abs(centre - 1.0).max() < 1e-3


def test_env_map_validation(ctx):
    s = scenes.box_room(16, 16)  # no dome light
    s.set_env(np.ones((4, 8, 3), np.float32))
    with pytest.raises(rrhip.RrhipError):
        rrhip.PtScene(ctx, s.desc())
    s2 = scenes.furnace(16, 16)
    bad = np.ones((4, 8, 3), np.float32)
    bad[1, 2, 0] = -1.0
    s2.set_env(bad)
    with pytest.raises(rrhip.RrhipError):
        rrhip.PtScene(ctx, s2.desc())
