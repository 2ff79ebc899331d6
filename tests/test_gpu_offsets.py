"""Error-bounded ray origins (the default) against render.rs's fixed push of
1e-4 along the shading normal (src/raytracer/pt/mod.rs:30, 976-977,
1212-1224), on the box room moved far from the world origin.

Between 4096 and 8192 one f32 ulp is 4.9e-4, so a fixed 1e-4 push leaves
the origin inside the hit point's rounding error: continuation and shadow
rays start on the wrong side of the surface about half of the time and hit
the surface they leave. Error-bounded origins push by the bound itself,
here 1-4 mm; the room's light hangs 5 cm below the ceiling instead of 1 cm
so that the pushes stay far below its smallest feature (f32 at 6e3 cannot
resolve a 1 cm gap with any push that is safe). Bars: the moved room
(loose triangles and spheres, and again with the walls as an instanced
mesh) equals the oracle's room at the origin within 3 sigma per pixel (at
most 1 % outside) and in the image means; at the origin the default and
the fixed push each match the oracle with the same origin rule; far from
the origin the fixed push renders more than 5 % darker (over 5 sigma).
"""
import sys

import numpy as np
import pytest

import rrhip
import scenes
from pt_stats import image_mean_z, mean_var, oracle_render, sigma_test

pytestmark = pytest.mark.gpu

W, H, SPP, DEPTH = 32, 32, 256, 8
FAR = (3.0e3, -6.0e3, 5.0e3)
LIGHT_Y = 0.95


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def oracle_room():
    s = scenes.box_room(W, H, light_y=LIGHT_Y)
    acc, _ = oracle_render(s.desc(), rrhip.RenderParams.make(max_depth=DEPTH), 0, SPP)
    return acc


@pytest.fixture(scope="module")
def oracle_room_fixed_push():
    s = scenes.box_room(W, H, light_y=LIGHT_Y)
    acc, _ = oracle_render(s.desc(), rrhip.RenderParams.make(max_depth=DEPTH, ray_offset=1e-4), 0, SPP)
    return acc


def gpu_render(ctx, scene, ray_offset):
    ps = rrhip.PtScene(ctx, scene.desc())
    se = ps.session(rrhip.RenderParams.make(max_depth=DEPTH, ray_offset=ray_offset))
    se.render(0, SPP)
    acc = se.accum()
    se.close()
    ps.close()
    return acc


def compare(acc_gpu, acc_cpu, label):
    z = sigma_test(acc_gpu, acc_cpu)
    frac = float((z > 3).mean())
    zi = image_mean_z(acc_gpu, acc_cpu)
    mg, _, _ = mean_var(acc_gpu)
    mc, _, _ = mean_var(acc_cpu)
    say(f"{label}: outside 3 sigma {frac:.4%}, image-mean z {np.round(zi, 2)}, "
        f"means gpu {mg.mean(0).round(5)} oracle {mc.mean(0).round(5)}")
    return frac, zi, mg.mean(0) / mc.mean(0)


@pytest.mark.parametrize("instanced", [False, True], ids=["loose", "instanced"])
def test_far_room_with_bounded_origins_matches_the_oracle_at_the_origin(ctx, oracle_room, instanced):
    say(f"box room {W}x{H} moved to {FAR} ({'instanced walls' if instanced else 'loose'}), "
        f"error-bounded origins, {SPP} spp vs the oracle's room at the origin")
    acc = gpu_render(ctx, scenes.box_room(W, H, offset=FAR, instanced=instanced, light_y=LIGHT_Y), 0.0)
    frac, zi, _ = compare(acc, oracle_room, "far room, bounded origins")
    assert frac <= 0.01 and (zi < 3).all()


def test_both_origin_rules_at_the_origin_match_the_oracle(ctx, oracle_room, oracle_room_fixed_push):
    say("box room at the origin: bounded origins and the fixed push, each vs the oracle with its rule")
    for off, ref in ((0.0, oracle_room), (1e-4, oracle_room_fixed_push)):
        acc = gpu_render(ctx, scenes.box_room(W, H, light_y=LIGHT_Y), off)
        frac, zi, _ = compare(acc, ref, f"room at the origin, ray_offset {off}")
        assert frac <= 0.01 and (zi < 3).all()


def test_fixed_push_fails_far_from_the_origin(ctx, oracle_room_fixed_push):
    say(f"box room moved to {FAR} with render.rs's fixed 1e-4 push (expected to self-intersect)")
    acc = gpu_render(ctx, scenes.box_room(W, H, offset=FAR, light_y=LIGHT_Y), 1e-4)
    frac, zi, ratio = compare(acc, oracle_room_fixed_push, "far room, fixed push")
    assert zi.min() > 5 and ratio.max() < 0.95, "the fixed push was expected to self-intersect at 6e3"
