"""The GPU path tracer's next-event estimation and MIS against closed-form
direct illumination (tests/direct_light.py), the same bars the oracle
meets in tests/test_oracle_direct.py: a rect, a disk and a sphere light
(both strategies at max_depth 2, the light-sampling arm alone at max_depth
1 against its power-heuristic share), point and distant lights (exact per
pixel up to f32 rounding), and three lights of different kinds under the
light BVH's pick pmf. Each render must also reject a 2 % error and the
balance heuristic, so the bars can see those mistakes.

Sphere lights are rendered twice: with the default error-bounded ray
origins (expected: shadow and BSDF rays from the surface point itself) and
with render.rs's fixed 1e-4 push (expected: the share its shadow rays lose
to the light's own sphere near the silhouette, and BSDF rays from the
pushed point)."""
import sys

import numpy as np
import pytest

import direct_light as dl
import rrhip

pytestmark = pytest.mark.gpu

W, H = 24, 18
SPP = 2048


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def render(ctx, lights, depth, spp=SPP, ray_offset=0.0):
    s = dl.scene(lights, W, H)
    ps = rrhip.PtScene(ctx, s.desc())
    se = ps.session(rrhip.RenderParams.make(max_depth=depth, firefly_clamp=0.0, ray_offset=ray_offset))
    se.render(0, spp)
    acc = se.accum()
    se.close()
    ps.close()
    return acc


def expected(case, lights, pts, offset=0.0):
    if case == "sphere":
        return dl.sphere_total(lights[0], pts, offset=offset)
    return dl.expected_total(lights, pts)


@pytest.mark.parametrize("case", list(dl.CASES))
def test_gpu_both_strategies_equal_the_closed_form(ctx, case):
    say(f"direct light {case}: depth 2, {W}x{H}, {SPP} spp")
    lights = dl.CASES[case]()
    if case == "sphere":  # render.rs's fixed push first
        acc = render(ctx, lights, 2, ray_offset=dl.RAY_OFFSET)
        want = expected(case, lights, dl.pixel_points(W, H), offset=dl.RAY_OFFSET)
        dl.check(acc, want, "gpu sphere depth 2, fixed 1e-4 push", say=say)
    acc = render(ctx, lights, 2)
    want = expected(case, lights, dl.pixel_points(W, H))
    if case in ("point", "distant"):
        # exact per sample up to f32 rounding of the hit point and light
        m = acc[:, :3] / acc[:, 6:7]
        rel = np.abs(m / want - 1).max()
        say(f"{case}: max per-pixel relative error {rel:.2e}")
        assert rel < 2e-5
        return
    dl.check(acc, want, f"gpu {case} depth 2", say=say)
    assert dl.rejects(acc, want * 0.98) and dl.rejects(acc, want * 1.02)


@pytest.mark.parametrize("case", ["rect", "disk"])
def test_gpu_light_arm_equals_its_power_heuristic_share(ctx, case):
    say(f"direct light {case}: NEE arm (depth 1)")
    lights = dl.CASES[case]()
    acc = render(ctx, lights, 1)
    pts = dl.pixel_points(W, H)
    dl.check(acc, dl.nee_arm(lights[0], pts), f"gpu {case} NEE arm", say=say)
    assert dl.rejects(acc, dl.nee_arm(lights[0], pts, heuristic="balance"))


def test_gpu_sphere_light_arm(ctx):
    say("direct light sphere: NEE arm without and with the emitter sphere")
    pts = dl.pixel_points(W, H)
    bare = dl.sphere_light(geometry=False)[0]
    acc = render(ctx, [bare], 1)
    dl.check(acc, dl.sphere_nee_arm(bare, pts), "gpu sphere NEE arm", say=say)
    assert dl.rejects(acc, dl.sphere_nee_arm(bare, pts, heuristic="balance"))
    light = dl.sphere_light()[0]
    acc = render(ctx, [light], 1)
    dl.check(acc, dl.sphere_nee_arm(light, pts, clip=True, offset=0.0),
             "gpu sphere NEE arm, emitter sphere present", say=say)
    acc = render(ctx, [light], 1, ray_offset=dl.RAY_OFFSET)
    dl.check(acc, dl.sphere_nee_arm(light, pts, clip=True), "gpu sphere NEE arm, emitter sphere present, "
             "fixed 1e-4 push", say=say)
