"""GPU wavefront path tracer against the f64 oracle.

Bars (f32 GPU arithmetic vs f64 CPU, same per-(pixel, sample) PCG32
streams): every pixel-channel mean within 3 sigma of the oracle's except a
fraction bounded by the heavy tails of 3 sigma at the sample counts used
(at most 1 %); whole-image channel means within 3 sigma; the white furnace
(albedo-1 diffuse sphere in a unit dome) equal to 1 within 1e-3; renders
bit-reproducible.
"""
import ctypes
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


def gpu_render(ctx, scene, params, start, count):
    d = scene.desc()
    ps = rrhip.PtScene(ctx, d)
    se = ps.session(params)
    se.render(start, count)
    acc = se.accum()
    st = se.stats()
    se.close()
    ps.close()
    return acc, st


def check_parity(acc_gpu, acc_cpu, max_frac=0.01, label=""):
    z = sigma_test(acc_gpu, acc_cpu)
    frac = float((z > 3).mean())
    zi = image_mean_z(acc_gpu, acc_cpu)
    mg, _, _ = mean_var(acc_gpu)
    mc, _, _ = mean_var(acc_cpu)
    say(f"{label}: outside 3 sigma {frac:.4%}, max z {z.max():.3f}, image-mean z {np.round(zi, 2)}, "
        f"means gpu {mg.mean(0).round(5)} cpu {mc.mean(0).round(5)}")
    assert frac <= max_frac, f"{label}: {frac:.4%} of pixel-channels outside 3 sigma"
    assert (zi < 3).all(), f"{label}: image means differ by {zi} sigma"


def test_white_furnace_to_1e3(ctx):
    say("white furnace 64x64, 512 spp")
    s = scenes.furnace(64, 64)
    acc, _ = gpu_render(ctx, s, rrhip.RenderParams.make(max_depth=10), 0, 512)
    m, _, _ = mean_var(acc)
    center = m.reshape(64, 64, 3)[8:56, 8:56].reshape(-1, 3).mean(0)
    say(f"furnace centre mean {center}")
    assert np.abs(center - 1.0).max() < 1e-3


def test_box_room_matches_oracle(ctx):
    say("box room 256x256, 16 spp, depth 8 (the C0 workload) vs oracle")
    s = scenes.box_room(256, 256)
    p = rrhip.RenderParams.make(max_depth=8)
    acc_g, st = gpu_render(ctx, s, p, 0, 16)
    acc_c, rays = oracle_render(s.desc(), p, 0, 16)
    say(f"gpu stats {st}; oracle rays {rays}")
    check_parity(acc_g, acc_c, label="box room")
    total = st["extension_rays"] + st["shadow_rays"]
    assert abs(total - rays) / rays < 0.01


def test_box_room_high_spp_matches_oracle(ctx):
    say("box room 96x96, 256 spp vs oracle")
    s = scenes.box_room(96, 96)
    p = rrhip.RenderParams.make(max_depth=8)
    acc_g, _ = gpu_render(ctx, s, p, 0, 256)
    acc_c, _ = oracle_render(s.desc(), p, 0, 256)
    check_parity(acc_g, acc_c, label="box room 256 spp")


def test_cornell_fixture_gi_and_parity(ctx):
    say("cornell.rib fixture 96x96, 48 spp")
    s = scenes.cornell_fixture(96, 96)
    p = rrhip.RenderParams.make(max_depth=10)
    acc_g, _ = gpu_render(ctx, s, p, 0, 48)
    img = (acc_g[:, :3] / acc_g[:, 6:7]).reshape(96, 96, 3)
    lum = img.reshape(-1, 3).mean(0) @ np.array([0.2126, 0.7152, 0.0722])
    assert 0.05 < lum < 2.0
    left = img[24:72, :32].reshape(-1, 3).mean(0)
    right = img[24:72, 64:].reshape(-1, 3).mean(0)
    assert left[0] > left[1] and right[1] > right[0]
    assert img[4, 4].sum() > 0.03
    acc_c, _ = oracle_render(s.desc(), p, 0, 48)
    check_parity(acc_g, acc_c, label="cornell")


@pytest.mark.parametrize("which", ["conductor", "glass"])
def test_furnace_metal_and_glass(ctx, which):
    say(f"{which} furnace 48x48, 256 spp vs oracle")
    mat = (scenes.material(specular_face_color=(1, 1, 1), specular_roughness=0.4, diffuse_gain=0)
           if which == "conductor" else
           scenes.material(diffuse_gain=0, glass_gain=1, glass_ior=1.5, glass_roughness=0.1))
    s = scenes.furnace(48, 48, mat=mat)
    p = rrhip.RenderParams.make(max_depth=10)
    acc_g, _ = gpu_render(ctx, s, p, 0, 256)
    m, _, _ = mean_var(acc_g)
    avg = m.reshape(48, 48, 3)[6:42, 6:42].mean()
    assert 0.55 <= avg <= 1.05
    acc_c, _ = oracle_render(s.desc(), p, 0, 256)
    check_parity(acc_g, acc_c, label=which)


def lights_scene():
    s = scenes.Scene(64, 64, 50.0, eye=(0, 1.5, -4), look_at=(0, 0, 0))
    floor = s.add_material(scenes.matte((0.7, 0.7, 0.7), diffuse_roughness=0.4))
    shiny = s.add_material(scenes.material(diffuse_gain=0.5, diffuse_color=(0.6, 0.3, 0.2),
                                           specular_ior=1.5, specular_roughness=0.3,
                                           clearcoat_gain=0.5, fuzz_gain=0.3))
    s.add_polygon([(-3, 0, -3), (3, 0, -3), (3, 0, 3), (-3, 0, 3)], floor)
    s.add_sphere((0, 0.6, 0), 0.6, shiny)
    s.add_sphere((1.2, 0.3, -0.5), 0.3, s.add_material(scenes.glass(1.5, 0.2)))
    s.add_light(scenes.light(rrhip.LIGHT_POINT, intensity=5.0, color=(1, 0.9, 0.8), p=(1, 2, -1)))
    s.add_light(scenes.light(rrhip.LIGHT_DISTANT, intensity=0.5, p=(-0.3, -1, 0.2), radius=0.05))
    s.add_light(scenes.light(rrhip.LIGHT_SPHERE, intensity=3.0, color=(0.4, 0.6, 1.0), p=(-1.5, 1.5, 0), radius=0.3))
    s.add_light(scenes.light(rrhip.LIGHT_DISK, intensity=4.0, p=(0, 2.5, 0), e1=(0.4, 0, 0), e2=(0, 0, 0.4)))
    s.add_light(scenes.light(rrhip.LIGHT_DOME, intensity=0.1))
    return s


def test_every_light_kind_and_lobe_matches_oracle(ctx):
    say("point / distant / sphere / disk / dome lights, Oren-Nayar, clearcoat, fuzz, glass")
    s = lights_scene()
    p = rrhip.RenderParams.make(max_depth=10)
    acc_g, _ = gpu_render(ctx, s, p, 0, 256)
    acc_c, _ = oracle_render(s.desc(), p, 0, 256)
    check_parity(acc_g, acc_c, label="lights")


def test_deterministic_and_sample_ranges_compose(ctx):
    say("determinism, sample ranges, accumulator round trip")
    s = scenes.box_room(48, 48)
    d = s.desc()
    ps = rrhip.PtScene(ctx, d)
    p = rrhip.RenderParams.make(max_depth=8, samples_per_pass=4)
    a = ps.session(p)
    a.render(0, 16)
    acc_a = a.accum()
    b = ps.session(p)
    b.render(0, 16)
    assert b.accum().tobytes() == acc_a.tobytes()
    c = ps.session(rrhip.RenderParams.make(max_depth=8, samples_per_pass=16))
    c.render(0, 5)
    ckpt = c.accum()
    c.load_accum(None)
    assert not c.accum().any()
    c.load_accum(ckpt)
    c.render(5, 11)
    assert np.allclose(c.accum(), acc_a, rtol=1e-12, atol=1e-12)
    for x in (a, b, c):
        x.close()
    ps.close()


def test_instanced_field_matches_oracle(ctx):
    say("instanced field: 300 instances of a 1280-triangle displaced icosphere, 96x54, 128 spp")
    s = scenes.instanced_field(96, 54, n_instances=300, levels=3, n_point_lights=16)
    p = rrhip.RenderParams.make(max_depth=12)
    acc_g, st = gpu_render(ctx, s, p, 0, 128)
    acc_c, rays = oracle_render(s.desc(), p, 0, 128)
    say(f"gpu rays {st['extension_rays'] + st['shadow_rays']} oracle rays {rays}")
    check_parity(acc_g, acc_c, label="instanced field")


def test_instanced_field_full_size_properties(ctx):
    """The full instanced workload (10,000 instances of the 20,480-triangle
    mesh, 64 point lights + 1 area light, 1920x1080, depth 12) at 2 spp:
    finite non-negative radiance, reproducible bits, and the primary-ray
    count equal to the pixel count per sample."""
    say("instanced field full size: 10000 x 20480 triangles, 1920x1080, 2 spp")
    s = scenes.instanced_field()
    d = s.desc()
    ps = rrhip.PtScene(ctx, d)
    p = rrhip.RenderParams.make(max_depth=12, samples_per_pass=2)
    a = ps.session(p)
    a.render(0, 2)
    acc = a.accum()
    st = a.stats()
    say(f"stats {st}")
    assert np.isfinite(acc).all() and (acc[:, :6] >= 0).all()
    assert (acc[:, 6] == 2).all()
    assert st["extension_rays"] >= 2 * 1920 * 1080
    b = ps.session(p)
    b.render(0, 2)
    assert b.accum().tobytes() == acc.tobytes()
    a.close()
    b.close()
    ps.close()


def test_relmse_on_device_matches_numpy(ctx):
    """relMSE = mean over pixels and channels of (x - r)^2 / (r^2 + eps),
    x and r the per-pixel means of two sessions, reduced in f64 on the
    device; checked against the same sum over the read-back accumulators,
    against zero for a session and itself, and against a session that
    differs by a known constant factor (closed form)."""
    say("relMSE on the device: box room 64x64, 16 spp vs 256 spp")
    s = scenes.box_room(64, 64)
    ps = rrhip.PtScene(ctx, s.desc())
    a = ps.session(rrhip.RenderParams.make(max_depth=8))
    a.render(0, 16)
    r = ps.session(rrhip.RenderParams.make(max_depth=8, sample_base=1 << 32))
    r.render(0, 256)
    acc_a, acc_r = a.accum(), r.accum()
    x = acc_a[:, :3] / acc_a[:, 6:7]
    ref = acc_r[:, :3] / acc_r[:, 6:7]
    for eps in (0.01, 1.0):
        want = float(np.mean((x - ref) ** 2 / (ref * ref + eps)))
        got = a.relmse(r, eps)
        say(f"eps {eps}: device {got!r} numpy {want!r}")
        assert abs(got - want) <= 1e-12 * want
    assert a.relmse(a) == 0.0
    # Scaling every sum of the reference by 2 makes x = r / 2 wherever the
    # two accumulators agree: load a's own sums, doubled, into r.
    doubled = acc_a.copy()
    doubled[:, :6] *= 2.0
    r.load_accum(doubled)
    want = float(np.mean((x - 2 * x) ** 2 / ((2 * x) ** 2 + 0.01)))
    assert abs(a.relmse(r) - want) <= 1e-12 * want
    a.close()
    r.close()
    ps.close()


def test_launch_modes_equal_stream_launches(ctx):
    """Launch modes (bit 0: one captured hipGraph per pass size, replayed;
    bit 1: queue sorted by shading class) give the accumulators of plain
    launches bit for bit, replay after replay, with other sessions' eager
    work between the replays; also with a radiance cache attached (its
    optimiser kernels are in the graph). The autotuner leaves the
    accumulator as it was."""
    say("launch modes vs stream launches (box room 48x48, instanced field with cache)")
    s = scenes.box_room(48, 48)
    ps = rrhip.PtScene(ctx, s.desc())
    p = rrhip.RenderParams.make(max_depth=8, samples_per_pass=4)
    ref = ps.session(p)
    ref.render(0, 14)  # passes of 4, 4, 4, 2
    other = ps.session(rrhip.RenderParams.make(max_depth=8, samples_per_pass=3))
    for mode in (1, 2, 3):
        g = ps.session(p)
        g.set_launch_mode(mode)
        for start, count in ((0, 6), (6, 5), (11, 3)):
            g.render(start, count)
            other.render(start, count)  # eager work between replays
        assert np.allclose(g.accum(), ref.accum(), rtol=1e-12, atol=1e-12), mode
        g2 = ps.session(p)
        g2.set_launch_mode(mode)
        g2.render(0, 14)
        assert g2.accum().tobytes() == ref.accum().tobytes(), mode
        g.close()
        g2.close()
    before = ref.accum()
    mode, ms = ref.autotune(4)
    say(f"autotune picked mode {mode} ({ms:.3f} ms for 4 spp)")
    assert 0 <= mode <= 3 and ms > 0
    assert ref.accum().tobytes() == before.tobytes()
    for x in (ref, other):
        x.close()
    ps.close()
    f = scenes.instanced_field(96, 54, n_instances=300, levels=3, n_point_lights=16)
    pf = rrhip.PtScene(ctx, f.desc())
    out = []
    for mode in (0, 3):
        c = rrhip.Nrc(ctx, rrhip.NrcConfig.make(log2_table=14, train_paths=2048, batch=2048))
        se = pf.session(rrhip.RenderParams.make(max_depth=12, samples_per_pass=2))
        se.set_nrc(c)
        se.set_launch_mode(mode)
        for k in range(4):
            se.render(2 * k, 2)
        out.append((se.accum(), c.get_params()))
        se.close()
        c.close()
    assert out[0][0].tobytes() == out[1][0].tobytes()
    assert out[0][1][0].tobytes() == out[1][1][0].tobytes()
    pf.close()


def test_thin_lens_and_orthographic_cameras_match_oracle(ctx):
    """Camera::generate_ray_lens (camera.rs:150-179): a thin lens focused
    on the back spheres' depth and an orthographic screen window, each
    rendered on the device and by the oracle."""
    import ctypes
    say("box room 96x96, thin lens (r 0.05, focus 1.0) and orthographic, 128 spp vs oracle")
    p = rrhip.RenderParams.make(max_depth=8)
    s = scenes.box_room(96, 96)
    s.camera.lens_radius = 0.05
    s.camera.focal_distance = 1.0
    acc_g, _ = gpu_render(ctx, s, p, 0, 128)
    acc_c, _ = oracle_render(s.desc(), p, 0, 128)
    check_parity(acc_g, acc_c, label="thin lens")
    s2 = scenes.box_room(96, 96)
    s2.camera.projection = 1
    s2.camera.ortho_half = (ctypes.c_double * 2)(0.9, 0.9)
    acc_g, _ = gpu_render(ctx, s2, p, 0, 128)
    acc_c, _ = oracle_render(s2.desc(), p, 0, 128)
    check_parity(acc_g, acc_c, label="orthographic")
    # invalid cameras are refused
    s3 = scenes.box_room(8, 8)
    s3.camera.lens_radius = 0.1  # focal_distance 0
    with pytest.raises(rrhip.RrhipError):
        rrhip.PtScene(ctx, s3.desc())
    s4 = scenes.box_room(8, 8)
    s4.camera.projection = 1  # ortho_half 0
    with pytest.raises(rrhip.RrhipError):
        rrhip.PtScene(ctx, s4.desc())


@pytest.mark.parametrize("mode", ["instance", "camera", "deform"])
def test_motion_blur_matches_shutter_fractions(ctx, mode):
    """Moving instance (mesh.rs:294-305), moving camera (camera.rs:
    134-146) or deforming mesh (mesh.rs:96-111): each pixel sees the dome
    for 1 - the fraction of the shutter the sliding black quad covers it."""
    from test_oracle_pt import covered_fraction, moving_quad_scene
    say(f"motion blur ({mode}): 32x4, 4096 spp vs shutter fractions")
    W = 32
    camera_motion = mode == "camera"
    s = moving_quad_scene(W, 4, camera_motion, deform=mode == "deform")
    acc, _ = gpu_render(ctx, s, rrhip.RenderParams.make(max_depth=2, firefly_clamp=0.0), 0, 4096)
    m = (acc[:, 0] / acc[:, 6]).reshape(4, W)
    x = ((np.arange(W) + 0.5) / W * 2 - 1) * 2.0
    want = 1 - covered_fraction(-x if camera_motion else x)
    se = np.sqrt(np.maximum(want * (1 - want), 1e-12) / (2 * 4096))
    got = m[1:3].mean(0)
    say(f"max |got - want| {np.abs(got - want).max():.4f}, max z {(np.abs(got - want) / np.maximum(se, 1e-9)).max():.2f}")
    assert np.array_equal(m[0], np.ones(W)) and np.array_equal(m[3], np.ones(W))
    assert np.all(np.abs(got - want) < 4.5 * se + 1e-6)


def deforming_scene(width=64, height=64, offset=(0.0, 0.0, 0.0)):
    """A displaced icosphere whose vertices swell by 1.35 and shift by 0.4
    in x over the shutter (Mesh::positions1), carrying rest-pose vertex
    normals that render.rs ignores while it deforms, rotated and scaled by
    its instance, over a ground under a rect and a distant light."""
    s = scenes.Scene(width, height, 40.0, eye=(0.0 + offset[0], 1.0 + offset[1], -5.0 + offset[2]),
                     look_at=offset)
    s.add_light(scenes.light(rrhip.LIGHT_RECT, intensity=7.0, p=(-1 + offset[0], 3 + offset[1], -1 + offset[2]),
                             e1=(2, 0, 0), e2=(0, 0, 2)))
    s.add_light(scenes.light(rrhip.LIGHT_DISTANT, intensity=1.0, p=(0.3, -1, 0.2)))
    g = s.add_material(scenes.matte((0.6, 0.6, 0.6)))
    o = np.asarray(offset)
    s.add_polygon([tuple(o + (-6, -1.2, -6)), tuple(o + (6, -1.2, -6)), tuple(o + (6, -1.2, 6)),
                   tuple(o + (-6, -1.2, 6))], g)
    pos, idx, nrm = rrhip.gen_icosphere(2, 3)
    pos1 = pos * 1.35 + np.array([0.4, 0.0, 0.0], np.float32)
    mesh = s.add_mesh(pos, idx, nrm, positions1=pos1)
    m = s.add_material(scenes.material(diffuse_color=(0.7, 0.3, 0.2), specular_roughness=0.25))
    xf = scenes._place((0, 0, 0), 25, 0.9)
    xf[:, 3] += o
    s.add_instance(mesh, m, xf)
    return s


@pytest.mark.parametrize("offset", [0.0, 1e-4])
def test_deforming_mesh_matches_oracle(ctx, offset):
    """mesh.rs:96-111, 199-220: vertices lerped at each ray's shutter time,
    geometric normals at that time, BLAS boxes over both shutter ends; the
    GPU's error-bounded origins carry the lerp's rounding."""
    say(f"deforming icosphere 64x64, 128 spp, ray_offset {offset}")
    s = deforming_scene()
    p = rrhip.RenderParams.make(max_depth=6, ray_offset=offset)
    g, _ = gpu_render(ctx, s, p, 0, 128)
    c, _ = oracle_render(s.desc(), p, 0, 128)
    check_parity(g, c, label=f"deforming icosphere offset {offset}")


def test_moving_instances_match_oracle(ctx):
    """A small instanced field where every third instance moves (a
    translation plus a different rotation at shutter close), against the
    oracle."""
    say("instanced field 96x54 with moving instances, 128 spp vs oracle")
    s = scenes.instanced_field(96, 54, n_instances=300, levels=3, n_point_lights=16)
    rng = np.random.default_rng(5)
    for k, inst in enumerate(s.instances):
        if k % 3:
            continue
        m0 = np.array(list(inst.object_to_world), float).reshape(3, 4)
        m1 = m0.copy()
        m1[:, 3] += rng.uniform(-3, 3, 3)
        c, sn = np.cos(0.3), np.sin(0.3)
        m1[:, :3] = m1[:, :3] @ np.array([[c, 0, sn], [0, 1, 0], [-sn, 0, c]])
        inst.object_to_world1 = (ctypes.c_double * 12)(*m1.ravel())
        inst.moving = 1
    p = rrhip.RenderParams.make(max_depth=12)
    acc_g, _ = gpu_render(ctx, s, p, 0, 128)
    acc_c, _ = oracle_render(s.desc(), p, 0, 128)
    check_parity(acc_g, acc_c, label="moving instances")


def test_near_mirror_specular_stays_finite(ctx):
    """PxrSurface roughness at render.rs's floor of 0.005 (alpha 2.5e-5):
    half vectors on the normal must not make the GGX density infinite in
    f32 (a regression test: a fractal-driven roughness reaching the floor
    once produced a NaN pixel); parity with the oracle."""
    s = scenes.Scene(48, 48, 40.0, eye=(0, 0.5, -4), look_at=(0, 0, 0))
    s.add_light(scenes.light(rrhip.LIGHT_DOME, intensity=0.5))
    s.add_light(scenes.light(rrhip.LIGHT_SPHERE, intensity=20.0, p=(1.5, 2.5, -2.0), radius=0.3))
    mirror = s.add_material(scenes.material(diffuse_gain=0.0, specular_face_color=(0.95, 0.9, 0.8),
                                            specular_roughness=0.005))
    coat = s.add_material(scenes.material(diffuse_color=(0.2, 0.3, 0.6), specular_roughness=0.005,
                                          clearcoat_gain=1.0, clearcoat_roughness=0.005))
    s.add_sphere((-0.6, 0, 0), 0.55, mirror)
    s.add_sphere((0.6, 0, 0), 0.55, coat)
    ground = s.add_material(scenes.matte((0.5, 0.5, 0.5)))
    s.add_polygon([(-5, -0.55, -5), (5, -0.55, -5), (5, -0.55, 5), (-5, -0.55, 5)], ground)
    p = rrhip.RenderParams.make(max_depth=6)
    g, _ = gpu_render(ctx, s, p, 0, 1024)
    say(f"near-mirror: finite {np.isfinite(g).all()}")
    assert np.isfinite(g).all()
    c, _ = oracle_render(s.desc(), p, 0, 256)
    check_parity(g, c, label="near-mirror spheres")


def test_fresh_session_starts_from_zero_after_a_freed_one(ctx):
    # a session's accumulators and ray counters are cleared on the stream
    # its work runs on: a session created where a busy one was freed reads
    # zero before rendering, and its profile counts exactly its own rays
    s = scenes.box_room(64, 64)
    ps = rrhip.PtScene(ctx, s.desc())
    p = rrhip.RenderParams.make(max_depth=8)
    busy = ps.session(p)
    busy.render(0, 64)
    assert busy.stats()["extension_rays"] > 0
    busy.close()
    se = ps.session(p)
    assert not se.accum().any()
    st0 = se.stats()
    assert st0["extension_rays"] == st0["shadow_rays"] == st0["samples_rendered"] == 0
    se.set_profiling(True)
    se.render(0, 4)
    prof, st = se.profile(), se.stats()
    say(f"fresh session: profile {prof['extend_rays']} / {prof['shadow_rays']} rays, "
        f"stats {st['extension_rays']} / {st['shadow_rays']}")
    assert 0 < prof["extend_rays"] == st["extension_rays"]
    assert 0 < prof["shadow_rays"] == st["shadow_rays"]
    se.close()
    ps.close()
