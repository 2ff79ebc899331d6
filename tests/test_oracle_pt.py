"""Pins the oracle's path tracer, BSDF lobes and light BVH with render.rs's
own tests: bxdf.rs:389-487 (pdf normalisation, lobe furnace, estimator
consistency), light_sampler.rs:322-371 (pmf consistency, near-light
importance), camera.rs:213-236 (filter support), pbr_tests.rs:31-89 (white
furnace, metal/glass energy) and pt_tests.rs:17-95 (Cornell GI sanity,
determinism, convergence), plus closed forms (Lambert f = rho/pi)."""
import ctypes

import numpy as np
import pytest

import orc
import rrhip
import scenes

LUM = np.array([0.2126, 0.7152, 0.0722])


def params_diffuse():
    return scenes.material(diffuse_gain=1.0, diffuse_color=(1, 1, 1), specular_ior=1.0)


def plastic(color, roughness):
    """pbr.rs:112-120 (PbrParams::from_plastic)."""
    return scenes.material(diffuse_gain=0.9, diffuse_color=color, specular_ior=1.5,
                           specular_roughness=max(roughness, 0.02))


def norm(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_lambert_closed_form():
    m = scenes.matte((0.6, 0.3, 0.15))
    wo = norm([0.2, -0.4, 0.8])
    wi = norm([-0.3, 0.1, 0.6])
    f, pdf = orc.bsdf_eval(m, wo, wi, 1.5)
    assert np.allclose(f, np.array([0.6, 0.3, 0.15]) / np.pi, rtol=1e-14)
    assert abs(pdf - wi[2] / np.pi) < 1e-15


def test_pdf_integrates_to_one():
    # bxdf.rs:389-422: 64 x 128 midpoint quadrature over the hemisphere.
    wo = norm([0.3, -0.15, 0.94])
    nt, npd = 64, 128
    th = (np.arange(nt) + 0.5) / nt * (np.pi / 2)
    ph = (np.arange(npd) + 0.5) / npd * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    wi = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dA = np.sin(T).ravel() * (np.pi / 2 / nt) * (2 * np.pi / npd)
    for mat in (params_diffuse(), scenes.metal((0.9, 0.7, 0.4), 0.3), plastic((0.5, 0.2, 0.2), 0.15)):
        _, pdf = orc.bsdf_eval_n(mat, wo, wi, 1.5)
        assert abs((pdf * dA).sum() - 1.0) < 0.05


def furnace_estimate(mat, seed, stream, n=200000):
    wo = norm([0.2, 0.1, 0.97])
    wi, f, pdf, ok = orc.bsdf_sample_n(mat, wo, 1.5, seed, stream, n)
    est = np.zeros(n)
    est[ok] = (f[ok] @ LUM) * np.abs(wi[ok, 2]) / pdf[ok]
    return est.mean()


def test_lobe_furnace_estimates():
    # bxdf.rs:424-450
    assert abs(furnace_estimate(params_diffuse(), 7, 11) - 1.0) < 0.02
    metal = furnace_estimate(scenes.metal((1, 1, 1), 0.4), 7, 12)
    assert 0.75 < metal <= 1.02
    assert furnace_estimate(plastic((1, 1, 1), 0.2), 7, 13) <= 1.03


def test_estimator_consistency():
    # bxdf.rs:452-487: BSDF-sampled vs uniform-hemisphere estimates agree.
    p = scenes.metal((0.8, 0.6, 0.3), 0.35)
    wo = norm([-0.3, 0.2, 0.93])
    n = 400000
    wi, f, pdf, ok = orc.bsdf_sample_n(p, wo, 1.5, 3, 9, n)
    by_bsdf = np.where(ok, (f @ LUM) * np.abs(wi[:, 2]) / np.where(ok, pdf, 1), 0).mean()
    u = orc.pcg32_seq(3, 10, 2 * n).astype(np.float64).reshape(n, 2) / 2.0**32
    z = u[:, 0]
    r = np.sqrt(np.maximum(1 - z * z, 0))
    phi = 2 * np.pi * u[:, 1]
    wu = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    fu, _ = orc.bsdf_eval_n(p, wo, wu, 1.5)
    by_uniform = ((fu @ LUM) * wu[:, 2] * 2 * np.pi).mean()
    assert abs(by_bsdf - by_uniform) / by_uniform < 0.03


def point_light(x, y, z, intensity):
    return scenes.light(rrhip.LIGHT_POINT, intensity=intensity, p=(x, y, z))


def test_light_bvh_pmf_consistency_and_normalisation():
    # light_sampler.rs:322-355
    lights = []
    for i in range(40):
        a = i * 0.7
        lights.append(point_light(np.cos(a) * (5.0 + (i % 7)), float(i % 5),
                                  np.sin(a) * (5.0 + (i % 3) * 3.0), 0.5 + (i % 4)))
    lights.append(scenes.light(rrhip.LIGHT_DOME, intensity=0.5))
    ls = orc.LightSampler(lights)
    p = (1.0, 0.5, -2.0)
    total = sum(ls.pmf(p, i) for i in range(len(lights)))
    assert abs(total - 1.0) < 1e-9
    for k in range(200):
        idx, pmf = ls.sample(p, (k + 0.5) / 200)
        re = ls.pmf(p, idx)
        assert abs(pmf - re) < 1e-9 * max(pmf, re, 1e-12)


def test_light_bvh_prefers_near_lights():
    # light_sampler.rs:357-371
    lights = [point_light(0.0, 1.0, 0.0, 10.0)]
    lights += [point_light(500.0 + i, 1.0, 500.0, 10.0) for i in range(63)]
    assert orc.LightSampler(lights).pmf((0.0, 0.0, 0.5), 0) > 0.25


def test_sphere_and_triangle_intersections():
    s = scenes.Scene(4, 4, 45, (0, 0, 0), (0, 0, 1))
    m = s.add_material(scenes.matte((1, 1, 1)))
    s.add_sphere((0, 0, 5), 1.0, m)
    s.add_triangle((-1, -1, 2), (1, -1, 2), (0, 1, 2), m)
    d = s.desc()
    t, point, normal, front, mat = orc.scene_intersect(d, (0, 0, 0), (0, 0, 1))
    # Triangle first; its winding faces +z, so the ray meets the back side and
    # the reported normal is flipped toward the ray (extra.rs:700-705).
    assert abs(t - 2.0) < 1e-12 and not front and normal[2] < -0.99
    t, _, normal, front, _ = orc.scene_intersect(d, (3, 0, 5), (-1, 0, 0))
    assert abs(t - 2.0) < 1e-12 and normal[0] > 0.99 and front
    t, _, normal, front, _ = orc.scene_intersect(d, (0, 0.5, 5), (0, 1, 0))  # from inside
    assert abs(t - 0.5) < 1e-12 and not front
    assert orc.scene_intersect(d, (0, 0, 0), (0, 0, -1)) is None


def mean_center(acc, w, h, margin):
    img = (acc[:, :3] / acc[:, 6:7]).reshape(h, w, 3)
    return img[margin:h - margin, margin:w - margin].reshape(-1, 3).mean(0)


def test_white_furnace_diffuse():
    # pbr_tests.rs:31-54 (an albedo-1 diffuse sphere in a unit dome renders 1)
    f = scenes.furnace(64, 64)
    acc, _ = orc.render(f.desc(), rrhip.RenderParams.make(max_depth=10), 0, 128)
    assert np.abs(mean_center(acc, 64, 64, 8) - 1.0).max() < 0.02
    # render.rs's PxrSurface defaults add an F0 = 0.04 specular coat (pbr.rs:133-140)
    f = scenes.furnace(64, 64, mat=scenes.material(diffuse_gain=1, diffuse_color=(1, 1, 1)))
    acc, _ = orc.render(f.desc(), rrhip.RenderParams.make(max_depth=10), 0, 128)
    assert np.abs(mean_center(acc, 64, 64, 8) - 1.0).max() < 0.02


def test_furnace_energy_metal_and_glass():
    # pbr_tests.rs:57-89
    for mat in (scenes.material(specular_face_color=(1, 1, 1), specular_roughness=0.4, diffuse_gain=0),
                scenes.material(diffuse_gain=0, glass_gain=1, glass_ior=1.5, glass_roughness=0.1)):
        f = scenes.furnace(48, 48, mat=mat)
        acc, _ = orc.render(f.desc(), rrhip.RenderParams.make(max_depth=10), 0, 96)
        avg = mean_center(acc, 48, 48, 6).mean()
        assert 0.55 <= avg <= 1.05


def test_cornell_gi_sanity():
    # pt_tests.rs:17-54 on the cornell.rib fixture
    s = scenes.cornell_fixture(96, 96)
    acc, _ = orc.render(s.desc(), rrhip.RenderParams.make(max_depth=10), 0, 48)
    img = (acc[:, :3] / acc[:, 6:7]).reshape(96, 96, 3)
    lum = img.reshape(-1, 3).mean(0) @ LUM
    assert 0.05 < lum < 2.0
    left = img[24:72, :32].reshape(-1, 3).mean(0)
    right = img[24:72, 64:].reshape(-1, 3).mean(0)
    assert left[0] > left[1] and right[1] > right[0]
    assert img[4, 4].sum() > 0.03


def test_deterministic():
    # pt_tests.rs:57-69
    s = scenes.box_room(32, 32)
    d = s.desc()
    p = rrhip.RenderParams.make(max_depth=8)
    a, _ = orc.render(d, p, 0, 8)
    b, _ = orc.render(d, p, 0, 8, threads=3)
    assert a.tobytes() == b.tobytes()


def test_sample_ranges_compose():
    # accum.rs:1-7: per-(pixel, sample) seeding makes range partitions exact.
    s = scenes.box_room(24, 24)
    d = s.desc()
    p = rrhip.RenderParams.make(max_depth=8)
    whole, _ = orc.render(d, p, 0, 8)
    part, _ = orc.render(d, p, 0, 3)
    part, _ = orc.render(d, p, 3, 5, accum=part)
    assert np.allclose(whole, part, rtol=1e-12, atol=0)


def test_pixel_list_render_equals_the_whole_image():
    """orc.render_pixels gives the listed pixels exactly the sums a whole-
    image render gives them (same full-image seeds)."""
    s = scenes.box_room(24, 16)
    d = s.desc()
    p = rrhip.RenderParams.make(max_depth=8)
    whole, _ = orc.render(d, p, 0, 6)
    px = np.array([0, 5, 23, 24, 200, 383, 17], np.uint32)
    part, _ = orc.render_pixels(d, p, px, 0, 6, threads=3)
    assert part.tobytes() == whole[px].tobytes()


def test_converges_with_spp():
    # pt_tests.rs:72-95
    s = scenes.cornell_fixture(32, 32)
    d = s.desc()
    p = rrhip.RenderParams.make(max_depth=10)

    def img(start, n):
        acc, _ = orc.render(d, p, start, n)
        return acc[:, :3] / acc[:, 6:7]
    ref = img(1000, 256)
    coarse = np.sqrt(((img(0, 4) - ref) ** 2).mean())
    fine = np.sqrt(((img(0, 64) - ref) ** 2).mean())
    assert fine < coarse * 0.6


def _plant(acc, factor):
    """Scales each pixel's samples by `factor` (per pixel, broadcast over
    channels): sums by f, sums of squares by f^2."""
    out = acc.copy()
    out[:, :3] *= factor
    out[:, 3:6] *= factor * factor
    return out


def _checker(w, block=8):
    yy, xx = np.meshgrid(np.arange(w), np.arange(w), indexing="ij")
    return ((((yy // block) + (xx // block)) % 2) * 2 - 1).ravel()[:, None].astype(np.float64)


@pytest.mark.parametrize("w,spp,uniform_bias,local_bias", [(64, 16, 0.05, 0.20), (96, 256, 0.02, 0.10)])
def test_sigma_harness_detects_a_planted_bias(w, spp, uniform_bias, local_bias):
    """The 3-sigma comparison of the GPU parity tests (at the sample counts
    they use: 16 spp for the box room, 256 for most others), fed two
    independently seeded oracle renders of the box room as the GPU tests
    feed it (pt_stats.oracle_render): they pass both bars. Planted on the
    CPU into one of them: a uniform bias fails the image-mean bar (5 % at
    64 x 64 x 16 spp, 2 % at 96 x 96 x 256 spp), and a bias of +-local_bias
    in alternating 8 x 8 blocks, which cancels in the image mean, fails
    the per-pixel bar (more than 1 % of pixel-channels beyond 3 sigma):
    +-20 % at 16 spp, +-10 % at 256 spp."""
    from pt_stats import image_mean_z, oracle_render, sigma_test
    s = scenes.box_room(w, w)
    d = s.desc()
    p = rrhip.RenderParams.make(max_depth=8)
    a, _ = orc.render(d, p, 0, spp)
    b, _ = oracle_render(d, p, 0, spp)
    assert not np.array_equal(a[:, :3], b[:, :3])
    assert (sigma_test(a, b) > 3).mean() < 0.01
    assert (image_mean_z(a, b) < 3).all()
    assert (image_mean_z(a, _plant(b, 1 + uniform_bias)) > 3).any()
    local = _plant(b, 1 + local_bias * _checker(w))
    assert (image_mean_z(a, local) < 3).all()
    assert (sigma_test(a, local) > 3).mean() > 0.01


def test_thin_lens_focuses_on_the_focal_plane():
    """camera.rs:150-179: every ray of a pixel (fixed sub-pixel position)
    passes through the same point of the focal plane, eye + dir *
    focal_distance; origins fill the aperture disk uniformly (E r^2 =
    R^2 / 2) in the right/up plane."""
    s = scenes.box_room(32, 24)
    s.camera.lens_radius = 0.1
    s.camera.focal_distance = 1.5
    s.camera.filter_width = 1e-9  # pixel centres
    d = s.desc()
    n = 4000
    xy = np.tile([[20, 7]], (n, 1))
    o, dirs = orc.camera_rays(d, xy, np.arange(n))
    eye = np.array([0, 0, -0.9])
    fwd = np.array([0, 0, 1.0])
    t = (1.5 - (o - eye) @ fwd) / (dirs @ fwd)
    hit = o + dirs * t[:, None]
    assert np.abs(hit - hit[0]).max() < 1e-9
    off = o - eye
    assert np.abs(off @ fwd).max() < 1e-15
    r2 = (off ** 2).sum(1)
    assert r2.max() <= 0.1 ** 2 * (1 + 1e-12)
    assert abs(r2.mean() / 0.01 - 0.5) < 0.02
    # pinhole: the same pixel centre without a lens passes through that point too
    s.camera.lens_radius = 0.0
    o0, d0 = orc.camera_rays(s.desc(), xy[:1], np.arange(1))
    t0 = 1.5 / (d0[0] @ fwd)
    assert np.allclose(o0[0] + d0[0] * t0, hit[0], atol=1e-12)


def test_orthographic_rays_are_parallel_on_the_screen_window():
    """camera.rs:154-159: origin eye + right u hw + up v hh, direction
    forward."""
    s = scenes.box_room(16, 8)
    s.camera.projection = 1
    s.camera.ortho_half = (ctypes.c_double * 2)(0.8, 0.4)
    s.camera.filter_width = 1e-9
    xs, ys = np.meshgrid(np.arange(16), np.arange(8))
    xy = np.stack([xs.ravel(), ys.ravel()], 1)
    o, dirs = orc.camera_rays(s.desc(), xy, np.zeros(len(xy), np.uint64))
    assert np.allclose(dirs, [0, 0, 1], atol=1e-15)
    u = (xy[:, 0] + 0.5) / 16 * 2 - 1
    v = 1 - (xy[:, 1] + 0.5) / 8 * 2
    want = np.stack([u * 0.8, v * 0.4, np.full(len(u), -0.9)], 1)
    assert np.allclose(o, want, atol=1e-9)


def moving_quad_scene(w, h, camera_motion=False, deform=False):
    """An orthographic camera (window 2 x 1) looking down +z at a black
    unit quad instance at z = 2 whose centre slides from x = -1 (shutter
    open) to x = +1 (close) — or, with camera_motion, a static quad at the
    origin seen by rays that camera motion shifts by the same amount; or,
    with deform, a static instance of a mesh whose vertices move by the
    same amount (mesh.rs:96-111) — against a unit dome: pixel x = X is
    covered during a known fraction of the shutter."""
    s = scenes.Scene(w, h, 40.0, eye=(0, 0, 0), look_at=(0, 0, 1))
    s.camera.projection = 1
    s.camera.ortho_half = (ctypes.c_double * 2)(2.0, 1.0)
    s.camera.filter_width = 1e-9
    s.add_light(scenes.light(rrhip.LIGHT_DOME, intensity=1.0))
    black = s.add_material(scenes.material(diffuse_gain=0.0, specular_ior=1.0))
    corners = np.array([(-0.5, -0.5, 0), (0.5, -0.5, 0), (0.5, 0.5, 0), (-0.5, 0.5, 0)])
    if deform:
        quad = s.add_mesh(corners + (-1, 0, 0), [(0, 1, 2), (0, 2, 3)], positions1=corners + (1, 0, 0))
        s.add_instance(quad, black, [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 2]])
        return s
    quad = s.add_mesh(corners, [(0, 1, 2), (0, 2, 3)])
    if camera_motion:
        s.add_instance(quad, black, [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 2]])
        # camera at x = -1: the quad sits at image x = +1 at t = 0; rays at
        # time t move by lerp(I, M, t), M a shift of +2 in x, so the quad
        # slides to image x = -1 at t = 1 (covered_fraction mirrored)
        m = np.array([[1, 0, 0, 2.0], [0, 1, 0, 0], [0, 0, 1, 0]])
        s.camera.motion_inv = (ctypes.c_double * 12)(*m.ravel())
        s.camera.has_motion_inv = 1
        s.camera.eye = (ctypes.c_double * 3)(-1.0, 0, 0)
        s.camera.look_at = (ctypes.c_double * 3)(-1.0, 0, 1)
    else:
        s.add_instance(quad, black, [[1, 0, 0, -1], [0, 1, 0, 0], [0, 0, 1, 2]],
                       [[1, 0, 0, 1], [0, 1, 0, 0], [0, 0, 1, 2]])
    return s


def covered_fraction(x):
    """The quad's centre at time t is -1 + 2t: it covers x for t in
    [(x + 0.5) / 2, (x + 1.5) / 2] within [0, 1]."""
    lo = np.clip((x + 0.5) / 2, 0, 1)
    hi = np.clip((x + 1.5) / 2, 0, 1)
    return hi - lo


@pytest.mark.parametrize("mode", ["instance", "camera", "deform"])
def test_motion_blur_covers_pixels_for_their_shutter_fraction(mode):
    """mesh.rs:294-305 (instance transform lerped at the ray's shutter
    time), camera.rs:134-146 (camera motion) and mesh.rs:96-111 (vertices
    lerped towards positions1): each pixel of a row sees the dome for 1 -
    (fraction of the shutter the quad covers it)."""
    W = 32
    camera_motion = mode == "camera"
    s = moving_quad_scene(W, 4, camera_motion, deform=mode == "deform")
    acc, _ = orc.render(s.desc(), rrhip.RenderParams.make(max_depth=2, firefly_clamp=0.0), 0, 4096)
    m = (acc[:, 0] / acc[:, 6]).reshape(4, W)
    x = ((np.arange(W) + 0.5) / W * 2 - 1) * 2.0
    want = 1 - covered_fraction(-x if camera_motion else x)
    se = np.sqrt(np.maximum(want * (1 - want), 1e-12) / (2 * 4096))
    got = m[1:3].mean(0)  # rows at y = -0.25, 0.25 (the quad spans |y| <= 0.5)
    assert np.array_equal(m[0], np.ones(W)) and np.array_equal(m[3], np.ones(W))
    assert np.all(np.abs(got - want) < 4.5 * se + 1e-9), np.c_[x, got, want]
