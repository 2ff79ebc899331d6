"""The oracle's film (pt::render_film, src/raytracer/pt/mod.rs:163-276)
pinned by what the geometry and the material parameters fix:

* albedo AOV = the closed form of pt/mod.rs:690-696 for the material hit
  (diffuse + primary-specular F0 + refraction, capped at 1), 0 on a miss;
* normal AOV = the sphere's outward normal at the camera-ray hit, depth =
  the eye-to-sphere distance along the ray (closed-form ray / sphere
  intersection per sample, averaged);
* id AOV = the material's id with full coverage inside the silhouette, 0
  outside;
* the diffuse split: a material with no specular-like lobe puts every
  contribution in the diffuse AOV (diffuse == beauty bit for bit,
  specular == 0); a smooth metal's dome light arrives through its first
  (specular) continuation, so diffuse holds only what is emitted before
  it and diffuse + specular == beauty.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "render.rs_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import orc  # noqa: E402
import rrhip  # noqa: E402
import scenes  # noqa: E402

W = H = 32


R = 0.8


def sphere_in_dome(mat):
    """The furnace layout (pbr_tests.rs:31-54) with a smaller sphere: radius
    R, 4 units down +z, inside a constant dome of radiance 1, 30 degree fov."""
    return scenes.furnace(W, H, mat=mat, radius=R)


def camera_dirs(desc, pixels, samples):
    xy = np.array([(p % W, p // W) for p in pixels for _ in samples], np.uint32)
    ss = np.array([s for _ in pixels for s in samples], np.uint64)
    o, d = orc.camera_rays(desc, xy, ss)
    return o.reshape(len(pixels), len(samples), 3), d.reshape(len(pixels), len(samples), 3)


def sphere_hits(o, d, c=np.array([0.0, 0.0, 4.0]), r=R):
    oc = o - c
    b = np.sum(oc * d, -1)
    cc = np.sum(oc * oc, -1) - r * r
    disc = b * b - cc
    t = -b - np.sqrt(np.maximum(disc, 0.0))
    return disc > 0, t, (o + d * t[..., None] - c) / r


def test_albedo_normal_depth_id_closed_form():
    m = scenes.material(diffuse_gain=0.8, diffuse_color=(0.5, 0.25, 1.0), specular_ior=1.5,
                        refraction_color=(1.0, 1.0, 1.0), glass_gain=0.25, object_id=7)
    s = sphere_in_dome(m)
    d = s.desc()
    spp = 16
    film = orc.render_film(d, rrhip.RenderParams.make(max_depth=4), 0, spp)
    f0 = ((1.5 - 1.0) / 2.5) ** 2
    want_albedo = np.minimum(0.8 * np.array([0.5, 0.25, 1.0]) + f0 + 0.25 * 1.0, 1.0)
    pixels = list(range(W * H))
    o, dirs = camera_dirs(d, pixels, list(range(spp)))
    hit, t, n = sphere_hits(o, dirs)
    full = hit.all(1)
    none = ~hit.any(1)
    assert full.sum() > 200 and none.sum() > 100
    # albedo: the closed form where every sample hits, 0 where none does
    np.testing.assert_allclose(film[full, 9:12], np.broadcast_to(want_albedo, (full.sum(), 3)), rtol=0, atol=1e-12)
    assert np.all(film[none, 9:15] == 0.0) and np.all(film[none, 15] == 0.0)
    # normal (outward = towards the camera side) and depth: the per-sample
    # closed form averaged over the pixel's samples
    want_n = n[full].mean(1)
    want_t = t[full].mean(1)
    np.testing.assert_allclose(film[full, 12:15], want_n, rtol=0, atol=1e-9)
    np.testing.assert_allclose(film[full, 15], want_t, rtol=1e-9)
    # id: 7 with full coverage inside, 0 (nothing hit) outside
    assert np.all(film[full, 16] == 7) and np.all(film[full, 17] == 1.0)
    assert np.all(film[none, 16] == 0) and np.all(film[none, 17] == 1.0)
    # partially covered pixels: coverage is the fraction of samples that hit
    part = ~full & ~none
    frac = hit[part].mean(1)
    cov = np.where(film[part, 16] == 7, film[part, 17], 1.0 - film[part, 17])
    np.testing.assert_allclose(cov, frac, rtol=0, atol=1e-12)


def test_diffuse_only_material_puts_everything_in_diffuse():
    s = sphere_in_dome(scenes.matte((0.7, 0.6, 0.5)))
    film = orc.render_film(s.desc(), rrhip.RenderParams.make(max_depth=6), 0, 8)
    assert np.array_equal(film[:, 3:6], film[:, 0:3])
    assert np.all(film[:, 6:9] == 0.0)
    assert film[:, 0:3].max() > 0.3


def test_metal_dome_light_goes_to_specular():
    m = scenes.metal((0.9, 0.9, 0.9), 0.05)
    s = sphere_in_dome(m)
    d = s.desc()
    film = orc.render_film(d, rrhip.RenderParams.make(max_depth=6), 0, 32)
    pixels = list(range(W * H))
    o, dirs = camera_dirs(d, pixels, list(range(32)))
    hit = sphere_hits(o, dirs)[0]
    on, off = hit.all(1), ~hit.any(1)
    beauty, diffuse, spec = film[:, 0:3], film[:, 3:6], film[:, 6:9]
    np.testing.assert_allclose(diffuse + spec, beauty, rtol=1e-12, atol=1e-15)
    # off the sphere the dome is seen directly (no continuation): diffuse
    assert np.all(spec[off] == 0.0) and np.all(beauty[off] == 1.0)
    # on the sphere most radiance arrives through the specular first bounce
    assert spec[on].sum() > 5.0 * diffuse[on].sum()


def test_film_beauty_matches_render():
    """The film's beauty is render()'s mean image for the same samples."""
    s = scenes.cornell_fixture(24, 24)
    d = s.desc()
    p = rrhip.RenderParams.make(max_depth=5)
    film = orc.render_film(d, p, 0, 6)
    acc, _ = orc.render(d, p, 0, 6)
    np.testing.assert_allclose(film[:, 0:3], acc[:, 0:3] / acc[:, 6:7], rtol=1e-12, atol=0)


@pytest.mark.parametrize("through", ["hull", "cutout"])
def test_hull_is_looked_through_but_cutout_is_recorded(through):
    """pt/mod.rs:662-700: a volume hull crossing adds its distance and is
    not the first real hit; a cutout surface is recorded even when the ray
    passes it."""
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
    d = s.desc()
    film = orc.render_film(d, rrhip.RenderParams.make(max_depth=3), 0, 4)
    o, dirs = camera_dirs(d, list(range(W * H)), list(range(4)))
    if through == "hull":
        want_t = (8.0 / dirs[..., 2] - 1e-4).mean(1)  # the crossing continues 1e-4 past the hull
        np.testing.assert_allclose(film[:, 15], want_t, rtol=1e-9)
        assert np.all(film[:, 16] == 2)
        np.testing.assert_allclose(film[:, 9:12], 0.5, atol=1e-12)
    else:
        want_t = (3.0 / dirs[..., 2]).mean(1)
        np.testing.assert_allclose(film[:, 15], want_t, rtol=1e-9)
        assert np.all(film[:, 16] == 3)
        np.testing.assert_allclose(film[:, 9:12], np.broadcast_to([0.9, 0.1, 0.1], (W * H, 3)), atol=1e-12)
