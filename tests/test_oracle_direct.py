"""Pins the oracle's next-event estimation and its MIS pairing with emitter
hits (render.rs pt/mod.rs:1032-1209, 1000-1023, 709-723) with closed-form
direct illumination of a Lambertian plane (tests/direct_light.py): a rect,
a disk and a sphere light (both strategies, then the light-sampling arm
alone against its power-heuristic share), a point and a distant light,
and three lights of different kinds under the light BVH's pick pmf.

The closed forms are first checked against brute-force quadrature, and
each render must also reject a plausible mistake's expectation (a 2 %
error in the total, the balance heuristic in place of the power
heuristic), so the bars have the power to see those mistakes."""
import numpy as np
import pytest

import direct_light as dl
import orc
import rrhip

W, H = 24, 18
SPP = 512


def render(lights, depth, spp=SPP, base=0):
    s = dl.scene(lights, W, H)
    p = rrhip.RenderParams.make(max_depth=depth, firefly_clamp=0.0, sample_base=base)
    acc, _ = orc.render(s.desc(), p, 0, spp)
    return acc


def test_pixel_points_are_the_camera_rays():
    s = dl.scene(dl.point_light(), W, H)
    xs, ys = np.meshgrid(np.arange(W), np.arange(H))
    xy = np.stack([xs.ravel(), ys.ravel()], 1)
    o, d = orc.camera_rays(s.desc(), xy, np.zeros(len(xy), np.uint64))
    assert np.allclose(d, [0, -1, 0], atol=1e-15)
    t = o[:, 1] / -d[:, 1]
    assert np.allclose(o + d * t[:, None], dl.pixel_points(W, H), atol=1e-12)


@pytest.mark.parametrize("case", ["rect", "disk"])
def test_area_form_factor_equals_quadrature(case):
    light = dl.CASES[case]()[0]
    pts = dl.pixel_points(W, H)[::7]
    quad = dl.nee_arm(light, pts, n=96, heuristic="none")
    assert np.allclose(quad, dl.expected_total([light], pts), rtol=1e-9, atol=0)


def test_disk_on_axis_closed_form():
    light = dl.Disk(0.0, 0.0, 0.7, 0.5, (1.0, 1.0, 1.0))
    e = light.irradiance(np.zeros((1, 3)))[0, 0]
    assert abs(e - np.pi * 0.25 / (0.25 + 0.49)) < 1e-14
    near = light.irradiance(np.array([[1e-6, 0, 0]]))[0, 0]
    assert abs(near - e) < 1e-9


def test_sphere_irradiance_equals_quadrature():
    light = dl.sphere_light()[0]
    pts = dl.pixel_points(W, H)[::5]
    quad = dl.sphere_nee_arm(light, pts, n=96, heuristic="none")
    assert np.allclose(quad, dl.expected_total([light], pts), rtol=1e-9, atol=0)
    # the two arms of an unclipped sphere sum to the irradiance; the
    # emitter-hit arm from the offset origin differs from it by O(1e-4)
    nee = dl.sphere_nee_arm(light, pts)
    bsdf = dl.sphere_bsdf_arm(light, pts)
    assert np.allclose(nee + bsdf, dl.expected_total([light], pts), rtol=5e-4, atol=0)


def expected(case, lights, pts):
    """Both strategies: rho/pi E(x); for the sphere light as render.rs
    computes it (its shadow rays clip the light's own sphere)."""
    if case == "sphere":
        return dl.sphere_total(lights[0], pts)
    return dl.expected_total(lights, pts)


@pytest.mark.parametrize("case", list(dl.CASES))
def test_both_strategies_equal_the_closed_form(case):
    lights = dl.CASES[case]()
    acc = render(lights, 2)
    want = expected(case, lights, dl.pixel_points(W, H))
    dl.check(acc, want, f"oracle {case} depth 2")
    assert dl.rejects(acc, want * 0.98) and dl.rejects(acc, want * 1.02)


@pytest.mark.parametrize("case", ["rect", "disk"])
def test_light_arm_equals_its_power_heuristic_share(case):
    lights = dl.CASES[case]()
    acc = render(lights, 1)
    pts = dl.pixel_points(W, H)
    want = dl.nee_arm(lights[0], pts)
    dl.check(acc, want, f"oracle {case} NEE arm")
    assert dl.rejects(acc, dl.nee_arm(lights[0], pts, heuristic="balance"))
    # the arm alone is a strict share of the total
    total = dl.expected_total(lights, pts)
    assert np.all(want < total) and want.sum() > 0.3 * total.sum()


def test_sphere_light_arm_equals_its_power_heuristic_share():
    # without the emitter sphere no shadow ray can clip the light itself
    lights = dl.sphere_light(geometry=False)
    acc = render(lights, 1)
    pts = dl.pixel_points(W, H)
    dl.check(acc, dl.sphere_nee_arm(lights[0], pts), "oracle sphere NEE arm")
    assert dl.rejects(acc, dl.sphere_nee_arm(lights[0], pts, heuristic="balance"))


def test_sphere_light_shadow_rays_clip_their_own_sphere():
    """render.rs's shadow ray to a sphere-light sample leaves from p + n
    1e-4 but is aimed from p, so near the silhouette it can enter the
    light's own sphere before its end (pt/mod.rs:1118-1124, 1212-1224).
    With the emitter sphere present the light arm equals the cone integral
    less exactly that blocked share, and not the unclipped integral."""
    pts = dl.pixel_points(W, H)
    light = dl.sphere_light()[0]
    acc = render([light], 1, spp=2048)
    clipped = dl.sphere_nee_arm(light, pts, clip=True)
    dl.check(acc, clipped, "oracle sphere NEE arm, emitter sphere present")
    full = dl.sphere_nee_arm(light, pts)
    assert dl.rejects(acc, full)
    loss = 1 - clipped.sum() / full.sum()
    assert 5e-4 < loss < 5e-3, loss


@pytest.mark.parametrize("case", ["point", "distant"])
def test_delta_lights_have_no_emitter_arm(case):
    lights = dl.CASES[case]()
    a1 = render(lights, 1)
    a2 = render(lights, 2)
    # a delta light's NEE value is exact: every sample of a pixel is equal
    assert np.array_equal(a1[:, :3], a2[:, :3])
    want = dl.expected_total(lights, dl.pixel_points(W, H))
    n = a1[:, 6:7]
    # f64 hit points carry ~1e-16 of rounding; the light's value ~1e-11
    assert np.allclose(a1[:, :3] / n, want, rtol=1e-10, atol=1e-15)
