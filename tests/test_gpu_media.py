"""Participating media on the GPU (atmosphere with a voxel fBm cloud:
delta tracking, ratio-tracked shadow transmittance, Henyey-Greenstein
scattering) against the f64 oracle and closed forms."""
import sys

import numpy as np
import pytest

import rrhip
import scenes
from pt_stats import image_mean_z, mean_var, oracle_render, sigma_test

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def render(ctx, s, params, start, count):
    ps = rrhip.PtScene(ctx, s.desc())
    se = ps.session(params)
    se.render(start, count)
    acc = se.accum()
    se.close()
    ps.close()
    return acc


def test_cloud_room_matches_oracle(ctx):
    say("cloud room 64x64 (32^3 fBm grid, g 0.3), depth 64, 256 spp vs oracle")
    s = scenes.cloud_room(64, 64, n=32)
    p = rrhip.RenderParams.make(max_depth=64)
    acc_g = render(ctx, s, p, 0, 256)
    acc_c, _ = oracle_render(s.desc(), p, 0, 256)
    z = sigma_test(acc_g, acc_c)
    frac = float((z > 3).mean())
    zi = image_mean_z(acc_g, acc_c)
    mg, _, _ = mean_var(acc_g)
    mc, _, _ = mean_var(acc_c)
    say(f"outside 3 sigma {frac:.4%}, image-mean z {np.round(zi, 2)}, means gpu {mg.mean(0).round(5)} "
        f"cpu {mc.mean(0).round(5)}")
    assert frac <= 0.01 and (zi < 3).all()


def test_absorbing_slab_matches_beer_lambert(ctx):
    """A purely absorbing medium of constant density over [-1,1]^3 seen
    against a unit dome: every pixel is exp(-sigma_a L), L the chord of its
    camera ray through the box (delta tracking turns each real collision
    into a zero-weight scatter)."""
    say("absorbing box 48x48 vs exp(-sigma L), 512 spp")
    W = H = 48
    s = scenes.Scene(W, H, 40.0, eye=(0.3, 0.2, -4.0), look_at=(0, 0, 0), background=(0, 0, 0))
    s.add_light(scenes.light(rrhip.LIGHT_DOME, intensity=1.0, color=(1, 1, 1)))
    ones = np.ones((4, 4, 4), np.float32)
    s.atmosphere = s.add_medium((0.8, 0.8, 0.8), (0.0, 0.0, 0.0), 0.0, density=ones)
    acc = render(ctx, s, rrhip.RenderParams.make(max_depth=8, firefly_clamp=0.0), 0, 512)
    m, var, n = mean_var(acc)
    # chord lengths of the (jittered) pixel rays: average exp(-sigma L) over
    # each pixel's footprint by supersampling the same camera on the host
    d = s.desc()
    cam = d.camera
    fwd = np.array(cam.look_at) - np.array(cam.eye)
    fwd /= np.linalg.norm(fwd)
    right = np.cross(np.array(cam.up), fwd)
    right /= np.linalg.norm(right)
    up = np.cross(fwd, right)
    hh = np.tan(np.radians(cam.fov_deg) / 2)
    hw = hh * W / H
    k = 8
    sub = (np.arange(k) + 0.5) / k
    want = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            px = x + sub[None, :]
            py = y + sub[:, None]
            u = (2 * px / W - 1) * hw
            v = (1 - 2 * py / H) * hh
            dirs = fwd[None, None, :] + u[..., None] * right + v[..., None] * up
            dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
            o = np.array(cam.eye)
            inv = 1.0 / dirs
            t0 = (-1 - o) * inv
            t1 = (1 - o) * inv
            tn = np.minimum(t0, t1).max(-1)
            tf = np.maximum(t0, t1).min(-1)
            L = np.clip(tf - np.maximum(tn, 0), 0, None)
            want[y, x] = np.exp(-0.8 * L).mean()
    got = m[:, 0].reshape(H, W)
    err = np.abs(got - want)
    se = np.sqrt(var[:, 0] / n[:, 0]).reshape(H, W)
    say(f"max |mean - exp(-sigma L)| {err.max():.4f}, median {np.median(err):.4f}")
    # the per-pixel statistical bound plus the footprint-quadrature error
    assert (err <= 4 * se + 0.01).mean() > 0.99
    assert abs(got.mean() - want.mean()) < 3e-3


def test_scattering_cloud_in_a_white_furnace(ctx):
    """A conservative (sigma_a = 0) cloud lit by a unit dome with nothing
    else in the scene scatters energy without loss: every pixel's
    expectation is 1."""
    say("white furnace with a pure-scattering fBm cloud, 64x64, 512 spp")
    s = scenes.Scene(64, 64, 30.0, eye=(0, 0, -5), look_at=(0, 0, 0))
    s.add_light(scenes.light(rrhip.LIGHT_DOME, intensity=1.0, color=(1, 1, 1)))
    dens = rrhip.gen_density_grid(32, 2, frequency=2.0, octaves=4, coverage=0.6, sharpness=3.0)
    s.atmosphere = s.add_medium((0, 0, 0), (3.0, 3.0, 3.0), 0.3, density=dens)
    acc = render(ctx, s, rrhip.RenderParams.make(max_depth=64, firefly_clamp=0.0), 0, 512)
    m, var, n = mean_var(acc)
    mean = m.mean()
    se = np.sqrt(var.mean() / n.mean() / m.size)  # standard error of the image mean
    say(f"furnace mean {mean:.5f} (standard error {se:.5f})")
    assert abs(mean - 1.0) < max(4 * se, 2e-3)


def test_single_scatter_towards_a_point_light_matches_quadrature(ctx):
    """Equiangular sampling towards a point light MIS-paired with
    collision NEE (pt/mod.rs:501-598): with one bounce, a 0.01-degree
    camera down a homogeneous atmosphere holds the single-scattered
    radiance, the 1-D integral along the axis."""
    from test_oracle_media import single_scatter_quadrature, single_scatter_scene
    say("single scatter towards a point light: 64x64 x 64 spp vs quadrature and oracle")
    s = single_scatter_scene(64, 64)
    p = rrhip.RenderParams.make(max_depth=1, firefly_clamp=0.0)
    acc = render(ctx, s, p, 0, 64)
    m, var, n = mean_var(acc)
    mean = m.mean(0)
    se = np.sqrt(var.mean(0) / n.mean() / m.shape[0])
    want = single_scatter_quadrature()
    acc_c, _ = oracle_render(single_scatter_scene(8, 8).desc(), p, 0, 4096)
    mc = (acc_c[:, :3] / acc_c[:, 6:7]).mean(0)
    say(f"gpu {mean.round(6)} (se {se.round(6)}), quadrature {want.round(6)}, oracle {mc.round(6)}")
    assert np.all(np.abs(mean - want) < 4 * se + 3e-4 * want)
    assert np.all(np.abs(mc - want) < 4 * se + 3e-4 * want)


def test_point_light_in_a_cloud_matches_oracle(ctx):
    say("cloud room 48x48 with a point light inside the atmosphere, depth 16, 256 spp vs oracle")
    s = scenes.cloud_room(48, 48, n=16)
    s.add_light(scenes.light(rrhip.LIGHT_POINT, intensity=2.0, color=(1.0, 0.9, 0.7), p=(0.3, 0.2, 0.1)))
    p = rrhip.RenderParams.make(max_depth=16)
    acc_g = render(ctx, s, p, 0, 256)
    acc_c, _ = oracle_render(s.desc(), p, 0, 256)
    z = sigma_test(acc_g, acc_c)
    frac = float((z > 3).mean())
    zi = image_mean_z(acc_g, acc_c)
    mg, _, _ = mean_var(acc_g)
    mc, _, _ = mean_var(acc_c)
    say(f"outside 3 sigma {frac:.4%}, image-mean z {np.round(zi, 2)}, means gpu {mg.mean(0).round(5)} "
        f"cpu {mc.mean(0).round(5)}")
    assert frac <= 0.01 and (zi < 3).all()


def test_media_rejections(ctx):
    s2 = scenes.cloud_room(16, 16, n=8)
    ps = rrhip.PtScene(ctx, s2.desc())
    se = ps.session(rrhip.RenderParams.make(max_depth=16))
    with pytest.raises(rrhip.RrhipError):
        se.set_nrc(rrhip.Nrc(ctx, rrhip.NrcConfig.make(log2_table=12, train_paths=64)))
    se.close()
    ps.close()
    # cutouts without media: no cache either
    s3 = scenes.box_room(16, 16)
    s3.add_sphere((0, 0, 0), 0.2, s3.add_material(scenes.matte((0.5, 0.5, 0.5), presence=0.5)))
    ps3 = rrhip.PtScene(ctx, s3.desc())
    se3 = ps3.session(rrhip.RenderParams.make(max_depth=8))
    cache = rrhip.Nrc(ctx, rrhip.NrcConfig.make(log2_table=12, train_paths=64))
    with pytest.raises(rrhip.RrhipError) as e:
        se3.set_nrc(cache)
    assert e.value.status == 5
    cache.close()
    se3.close()
    ps3.close()
    # an interior medium index past the media, a cutout outside [0, 1]
    s4 = scenes.box_room(16, 16)
    s4.add_sphere((0, 0, 0), 0.2, s4.add_material(scenes.hull(3)))
    with pytest.raises(rrhip.RrhipError):
        rrhip.PtScene(ctx, s4.desc())
    s5 = scenes.box_room(16, 16)
    s5.add_sphere((0, 0, 0), 0.2, s5.add_material(scenes.matte((0.5, 0.5, 0.5), presence=1.5)))
    with pytest.raises(rrhip.RrhipError):
        rrhip.PtScene(ctx, s5.desc())


def test_hull_sphere_absorbs_by_its_chord(ctx):
    """Hull crossings push and pop the medium (pt/mod.rs:662-683): an
    absorbing interior transmits exp(-sigma_a chord) of the dome."""
    from test_oracle_media import center_chords, hull_sphere_scene
    say("hull sphere with an absorbing interior, 24x24, 1024 spp vs exp(-sigma chord)")
    W = H = 24
    s = hull_sphere_scene(W, H, (0.8, 0.5, 0.2), (0, 0, 0))
    acc = render(ctx, s, rrhip.RenderParams.make(max_depth=8, firefly_clamp=0.0), 0, 1024)
    m, var, n = mean_var(acc)
    se = np.sqrt(var / n)
    chord, clean = center_chords(W, H)
    want = np.exp(-np.outer(chord, [0.8, 0.5, 0.2]))
    inside = clean & (chord > 0)
    z = np.abs(m - want)[inside] / np.maximum(se[inside], 1e-12)
    out = m[clean & (chord == 0)]
    say(f"inside: max z {z.max():.2f}, mean z {z.mean():.2f}; outside max |m - 1| {np.abs(out - 1).max():.2e}")
    assert np.abs(out - 1).max() < 1e-6
    assert z.max() < 5.5 and z.mean() < 1.2


def test_hull_white_furnace(ctx):
    from test_oracle_media import hull_sphere_scene
    say("conservative cloud inside a hull in a white furnace, 32x32, 512 spp")
    s = hull_sphere_scene(32, 32, (0, 0, 0), (2.0, 2.0, 2.0), 0.3)
    acc = render(ctx, s, rrhip.RenderParams.make(max_depth=64, firefly_clamp=0.0), 0, 512)
    m, var, n = mean_var(acc)
    mean = m.mean()
    se = np.sqrt(var.mean() / n.mean() / m.size)
    say(f"furnace mean {mean:.5f} (standard error {se:.5f})")
    assert abs(mean - 1.0) < max(4 * se, 2e-3)


def test_cutouts_attenuate_shadows_and_pass_camera_rays(ctx):
    from test_oracle_media import cutout_scene
    say("cutout quad: shadow attenuation 1 - presence (exact), camera pass-through 1 - presence")
    W = H = 16
    acc = render(ctx, cutout_scene(W, H, 0.6), rrhip.RenderParams.make(max_depth=4, firefly_clamp=0.0), 0, 16)
    m = (acc[:, :3] / acc[:, 6:7])[:, 0].reshape(H, W)
    lit, shaded = m[:, W // 2 + 2:], m[:, :W // 2 - 2]
    say(f"lit {lit.mean():.6f} (want {0.5 / np.pi:.6f}), shaded {shaded.mean():.6f} (want {0.2 / np.pi:.6f})")
    assert np.allclose(lit, 0.5 / np.pi, rtol=2e-5)
    assert np.allclose(shaded, 0.5 / np.pi * 0.4, rtol=2e-5)
    s2 = scenes.Scene(64, 64, 20.0, eye=(0, 0, -2), look_at=(0, 0, 0))
    s2.add_light(scenes.light(rrhip.LIGHT_DOME, intensity=1.0))
    q = s2.add_material(scenes.material(diffuse_gain=0.0, specular_ior=1.0, presence=0.3))
    s2.add_polygon([(-3, -3, 0), (3, -3, 0), (3, 3, 0), (-3, 3, 0)], q)
    acc2 = render(ctx, s2, rrhip.RenderParams.make(max_depth=4, firefly_clamp=0.0), 0, 256)
    mean = acc2[:, 0].sum() / acc2[:, 6].sum()
    assert abs(mean - 0.7) < 4 * np.sqrt(0.21 / acc2[:, 6].sum()), mean


def transparent_room(w, h):
    """The box room with a glass sphere holding a scattering cloud (pushed
    and popped on refraction), a hull sphere around an absorbing medium,
    and a cutout sphere of presence 0.5."""
    s = scenes.box_room(w, h)
    s.spheres = []
    cloud = s.add_medium((0.05, 0.05, 0.05), (1.5, 1.2, 0.9), 0.3)
    ink = s.add_medium((1.5, 0.6, 0.2), (0, 0, 0), 0.0)
    glass = s.add_material(scenes.material(diffuse_gain=0.0, specular_ior=1.0, glass_gain=1.0, glass_ior=1.3,
                                           glass_roughness=0.1, interior=cloud))
    s.add_sphere((-0.4, -0.6, 0.2), 0.35, glass)
    s.add_sphere((0.45, -0.55, -0.1), 0.4, s.add_material(scenes.hull(ink)))
    s.add_sphere((0.0, 0.3, 0.4), 0.25, s.add_material(scenes.matte((0.8, 0.6, 0.2), presence=0.5)))
    return s


def test_transparent_room_matches_oracle(ctx):
    say("room with a cloud-filled glass sphere, a hull sphere of ink and a cutout sphere, 48x48, 256 spp vs oracle")
    s = transparent_room(48, 48)
    p = rrhip.RenderParams.make(max_depth=16)
    acc_g = render(ctx, s, p, 0, 256)
    acc_c, _ = oracle_render(s.desc(), p, 0, 256)
    z = sigma_test(acc_g, acc_c)
    frac = float((z > 3).mean())
    zi = image_mean_z(acc_g, acc_c)
    mg, _, _ = mean_var(acc_g)
    mc, _, _ = mean_var(acc_c)
    say(f"outside 3 sigma {frac:.4%}, image-mean z {np.round(zi, 2)}, means gpu {mg.mean(0).round(5)} "
        f"cpu {mc.mean(0).round(5)}")
    assert frac <= 0.01 and (zi < 3).all()


def test_procedural_density_rejections(ctx):
    # a medium with both a grid and the procedural field, too many
    # octaves, an unknown procedural kind, non-finite parameters
    def scene(**kw):
        s = scenes.box_room(16, 16)
        mi = s.add_medium((0.1, 0.1, 0.1), (0.5, 0.5, 0.5), fbm={})
        for k, v in kw.items():
            setattr(s.media[mi], k, v)
        s.add_sphere((0, 0, 0), 0.2, s.add_material(scenes.hull(mi)))
        return s
    rrhip.PtScene(ctx, scene().desc()).close()  # the valid one builds
    ones = np.ones((2, 2, 2), np.float32)
    for kw in (dict(nx=2, ny=2, nz=2, density=ones.ctypes.data), dict(octaves=65), dict(procedural=2),
               dict(frequency=float("nan")), dict(sharpness=float("inf"))):
        with pytest.raises(rrhip.RrhipError):
            rrhip.PtScene(ctx, scene(**kw).desc())
