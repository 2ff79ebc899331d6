"""Scenes built by the RIB front end render on the GPU as the oracle renders
them: the reference's cornell.rib and quadric_zoo.rib fixtures (at a
reduced resolution) and a RIB exercising PxrSurface, Pxr rect / sphere /
disk lights, PointsPolygons under ObjectInstance and an Interior medium.
Bar: at most 1 % of pixel-channels outside 3 sigma, image means within 3
sigma (the path-tracing parity bar)."""
import os
import sys

import numpy as np
import pytest

import rib
import rrhip
from pt_stats import image_mean_z, mean_var, oracle_render, sigma_test

pytestmark = pytest.mark.gpu

RIBS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rib")

PXR_RIB = '''
Format 64 48 1
Projection "perspective" "fov" [50]
Translate 0 -0.5 0
WorldBegin
  AttributeBegin
    Translate 0 2.4 5 Rotate 90 1 0 0 Scale 1.5 1.5 1
    Light "PxrRectLight" "key" "intensity" [6] "lightColor" [1 0.9 0.8]
  AttributeEnd
  AttributeBegin
    Translate -1.6 1.2 4 Scale 0.4 0.4 0.4
    Light "PxrSphereLight" "fill" "intensity" [8] "exposure" [-1]
  AttributeEnd
  AttributeBegin
    Translate 1.6 1.6 4.5 Rotate 60 1 0 0 Scale 0.6 0.6 0.6
    Light "PxrDiskLight" "rim" "intensity" [10] "lightColor" [0.6 0.8 1]
  AttributeEnd
  Bxdf "PxrSurface" "floor" "diffuseColor" [0.5 0.5 0.45] "specularIor" [1]
  Polygon "P" [-4 -0.5 1  4 -0.5 1  4 -0.5 9  -4 -0.5 9]
  ObjectBegin "tile"
    PointsPolygons [4 4] [0 1 2 3  4 5 6 7]
      "P" [-0.4 0 -0.4  0.4 0 -0.4  0.4 0 0.4  -0.4 0 0.4
           -0.4 0.3 -0.4  -0.4 0.3 0.4  0.4 0.3 0.4  0.4 0.3 -0.4]
  ObjectEnd
  AttributeBegin
    Bxdf "PxrSurface" "coat" "diffuseColor" [0.7 0.2 0.1] "clearcoatGain" [1] "specularRoughness" [0.3]
    Translate -0.9 -0.5 4.5 ObjectInstance "tile"
    Translate 1.8 0 0 Rotate 30 0 1 0 ObjectInstance "tile"
  AttributeEnd
  AttributeBegin
    Bxdf "PxrSurface" "murky" "diffuseGain" [0] "specularIor" [1] "glassRefractionGain" [1] "glassIor" [1.33]
    Interior "water" "sigma_a" [0.3 0.2 0.1] "sigma_s" [0.8 0.8 0.8] "g" [0.3]
    Translate 0 0.2 5.2 Sphere 0.6 -0.6 0.6 360
  AttributeEnd
WorldEnd
'''


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


def parity(g, c, label):
    z = sigma_test(g, c)
    frac = float((z > 3).mean())
    zi = image_mean_z(g, c)
    mg, _, _ = mean_var(g)
    mc, _, _ = mean_var(c)
    say(f"{label}: outside 3 sigma {frac:.4%}, image-mean z {np.round(zi, 2)}, means gpu {mg.mean(0).round(4)} "
        f"cpu {mc.mean(0).round(4)}")
    assert frac <= 0.01 and (zi < 3).all()


@pytest.mark.parametrize("name,w,h,spp", [("cornell", 48, 48, 64), ("quadric_zoo", 96, 54, 48)])
def test_rib_fixture_matches_oracle(ctx, name, w, h, spp):
    say(f"{name}.rib through the RIB front end at {w}x{h}, {spp} spp")
    s, _ = rib.load(os.path.join(RIBS, name + ".rib"))
    s.camera.width, s.camera.height = w, h
    p = rrhip.RenderParams.make(max_depth=6)
    parity(gpu_render(ctx, s, p, 0, spp), oracle_render(s.desc(), p, 0, spp)[0], name)


def test_pxr_lights_surfaces_instances_media_match_oracle(ctx):
    say("Pxr lights, PxrSurface, instanced PointsPolygons and an interior medium, 64x48, 64 spp")
    s, _ = rib.loads(PXR_RIB)
    p = rrhip.RenderParams.make(max_depth=8)
    parity(gpu_render(ctx, s, p, 0, 64), oracle_render(s.desc(), p, 0, 64)[0], "pxr")


MOTION_RIB = '''
Format 48 36 1
Projection "perspective" "fov" [45]
MotionBegin [0 1] Translate 0 0 0 Translate 0.15 0 0 MotionEnd
WorldBegin
  Light "PxrDomeLight" "sky" "intensity" [0.8]
  AttributeBegin
    Translate 0 3 4 Rotate 90 1 0 0 Scale 2 2 1
    Light "PxrRectLight" "key" "intensity" [5]
  AttributeEnd
  Bxdf "PxrSurface" "floor" "diffuseColor" [0.6 0.6 0.6] "specularIor" [1]
  Polygon "P" [-4 -1 1  4 -1 1  4 -1 9  -4 -1 9]
  ObjectBegin "box"
    Bxdf "PxrSurface" "red" "diffuseColor" [0.8 0.2 0.1]
    PointsPolygons [4 4 4 4] [0 1 2 3  4 5 6 7  0 4 7 3  1 2 6 5]
      "P" [-0.5 -0.5 -0.5  0.5 -0.5 -0.5  0.5 0.5 -0.5  -0.5 0.5 -0.5
           -0.5 -0.5 0.5  -0.5 0.5 0.5  0.5 0.5 0.5  0.5 -0.5 0.5]
  ObjectEnd
  AttributeBegin
    MotionBegin [0 1] Translate -0.6 -0.4 4 Translate 0.4 -0.2 4 MotionEnd
    Rotate 30 0 1 0
    ObjectInstance "box"
  AttributeEnd
WorldEnd
'''


def test_motion_blocks_match_oracle(ctx):
    say("RIB transform motion blocks: moving camera and a moving instance, 48x36, 64 spp")
    s, _ = rib.loads(MOTION_RIB)
    assert s.camera.has_motion_inv == 1 and s.instances[0].moving == 1
    p = rrhip.RenderParams.make(max_depth=6)
    parity(gpu_render(ctx, s, p, 0, 64), oracle_render(s.desc(), p, 0, 64)[0], "motion")


POLYGON_RIB = '''
Format 48 36 1
Projection "perspective" "fov" [45]
WorldBegin
  Light "PxrDomeLight" "sky" "intensity" [0.6]
  AttributeBegin
    Translate 0 3 3 Rotate 90 1 0 0 Scale 2 2 1
    Light "PxrRectLight" "key" "intensity" [6]
  AttributeEnd
  Bxdf "PxrSurface" "floor" "diffuseColor" [0.6 0.6 0.6] "specularIor" [1]
  Polygon "P" [-4 -1 1  4 -1 1  4 -1 9  -4 -1 9]
  AttributeBegin
    Bxdf "PxrSurface" "frame" "diffuseColor" [0.8 0.3 0.1] "specularIor" [1.5] "specularRoughness" [0.3]
    GeneralPolygon [6 4] "P" [-1.6 -1 5  1.6 -1 5  1.6 1.4 5  0.3 1.4 5  0.3 0.6 5  -1.6 0.6 5
                              -1 -0.6 5  0.8 -0.6 5  0.8 0.3 5  -1 0.3 5]
  AttributeEnd
  AttributeBegin
    Bxdf "PxrSurface" "plate" "diffuseColor" [0.2 0.5 0.8]
    Translate 0.9 -0.3 3.2 Rotate 35 0 1 0
    PointsGeneralPolygons [2] [4 4] [0 1 2 3  4 5 6 7]
      "P" [-0.5 -0.5 0  0.5 -0.5 0  0.5 0.5 0  -0.5 0.5 0  -0.2 -0.2 0  0.2 -0.2 0  0.2 0.2 0  -0.2 0.2 0]
  AttributeEnd
WorldEnd
'''


def test_polygons_with_holes_match_oracle(ctx):
    # a concave frame (GeneralPolygon) and a plate with a square window
    # (PointsGeneralPolygons): light reaches the floor through both holes
    say("RIB GeneralPolygon / PointsGeneralPolygons with holes, 48x36, 64 spp")
    s, _ = rib.loads(POLYGON_RIB)
    assert len(s.meshes) == 2
    p = rrhip.RenderParams.make(max_depth=6)
    parity(gpu_render(ctx, s, p, 0, 64), oracle_render(s.desc(), p, 0, 64)[0], "polygons with holes")


CLOUD_RIB = '''
Format 48 36 1
Projection "perspective" "fov" [40]
WorldBegin
  Light "PxrDomeLight" "sky" "intensity" [0.5]
  AttributeBegin
    Translate 0 3 4 Rotate 90 1 0 0 Scale 2 2 1
    Light "PxrRectLight" "key" "intensity" [6]
  AttributeEnd
  Bxdf "PxrSurface" "floor" "diffuseColor" [0.6 0.6 0.6] "specularIor" [1]
  Polygon "P" [-4 -1 1  4 -1 1  4 -1 9  -4 -1 9]
  AttributeBegin
    Bxdf "PxrSurface" "hull" "diffuseGain" [0] "specularIor" [1] "glassRefractionGain" [1] "glassIor" [1.33]
    Interior "cloud" "sigma_a" [0.1 0.1 0.1] "sigma_s" [3 2.5 2] "g" [0.3]
      "density" "fbm" "frequency" [1.6] "octaves" [4] "coverage" [0.55] "sharpness" [3]
    Translate 0 0.1 4.5 Sphere 1.0 -1.0 1.0 360
  AttributeEnd
  Atmosphere "haze" "sigma_a" [0 0 0] "sigma_s" [0.06 0.06 0.06] "density" "noise" "frequency" [0.5]
    "maxdistance" [30]
WorldEnd
'''


def test_procedural_cloud_media_match_oracle(ctx):
    # render.rs's procedural fBm density (Interior / Atmosphere "density"
    # "fbm" / "noise"), evaluated per lookup at the world point: a cloud
    # inside a water-glass sphere and a thin noisy haze over 30 units
    say("RIB procedural fBm cloud and haze, 48x36, 128 spp")
    s, _ = rib.loads(CLOUD_RIB)
    assert [m.procedural for m in s.media] == [rrhip.MEDIUM_FBM, rrhip.MEDIUM_FBM]
    p = rrhip.RenderParams.make(max_depth=32)
    parity(gpu_render(ctx, s, p, 0, 128), oracle_render(s.desc(), p, 0, 128)[0], "procedural cloud")
