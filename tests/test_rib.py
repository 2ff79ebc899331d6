"""The RIB front end (render.rs_amd/rib.py) against render.rs's parser and
scene-builder semantics:

* the scene builder's own unit tests (src/parser/scene_builder.rs:
  2026-2200) restated on the built scene description;
* the reference's RIB fixtures (tests/golden/rib, copied unchanged) built
  into the same scenes as the hand-written restatements the GPU parity
  tests already use (cornell.rib, quadric_zoo.rib): equal lights, equal
  quadrics and triangles, and identical closest hits for a fan of camera
  rays through the oracle's scene intersection;
* closed forms of the Pxr lights' placement (unit square, sphere and disk
  under the current transform), the PxrSurface parameter map and clamps
  (pbr.rs:133-185), media defaults and interior binding, and the object
  ids of the id AOV (identifier-named geometry shares an id).
"""
import os
import warnings

import numpy as np
import pytest

import orc
import rib
import rtex  # noqa: F401
import rrhip
import scenes

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RIBS = os.path.join(GOLD, "rib")


def build(text, strict=True):
    return rib.Builder(strict=strict).run(rib.parse(text))


def d3(v):
    return np.array(list(v), float)


def n_objects(b):
    return len(b.tris) + len(b.quadrics)


def mat_fields(m):
    """Every material field but the object id."""
    scal = [getattr(m, f) for f in ("diffuse_gain", "diffuse_roughness", "specular_roughness", "specular_ior",
                                    "clearcoat_gain", "clearcoat_roughness", "fuzz_gain", "glass_gain",
                                    "glass_ior", "glass_roughness", "area_light", "cutout", "interior")]
    vec = [d3(getattr(m, f)) for f in ("diffuse_color", "specular_face_color", "specular_edge_color",
                                       "fuzz_color", "refraction_color", "glow", "emission")]
    return tuple(np.round(np.concatenate([np.array(scal, float)] + vec), 12))


def light_fields(l):
    return np.r_[l.kind, l.intensity, d3(l.color), d3(l.p), d3(l.e1), d3(l.e2), l.radius]


# ---- tokenizer ---------------------------------------------------------------------

def test_tokenizer_values_comments_and_declarations():
    reqs = rib.parse('##RenderMan RIB\nversion 3.04 # trailing\n'
                     'Light "PxrRectLight" "key" "float intensity" [2.5e1] "color lightColor" [1 .5 -2E-1]\n'
                     'Attribute "identifier" "name" ["a \\"b\\""]\nPolygon "P" [\n# inside\n 0 0 0 1 0 0 0 1 0 ]\n')
    assert [r[0] for r in reqs] == ["version", "Light", "Attribute", "Polygon"]
    assert reqs[0][1] == [3.04]
    assert reqs[1][1][:3] == ["PxrRectLight", "key", "float intensity"]
    assert reqs[1][1][3] == (25.0,) and reqs[1][1][5] == (1.0, 0.5, -0.2)
    assert reqs[2][1][2] == ['a "b"']
    assert reqs[3][1][1] == (0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0)
    p = rib._Params(reqs[1][1], rib._pstart(reqs[1][1]))
    assert p.number("intensity") == 25.0 and p.color("lightColor", None) == (1.0, 0.5, -0.2)
    with pytest.raises(rib.RibError):
        rib.parse('Polygon "P" [0 0 "x"]')
    with pytest.raises(rib.RibError):
        rib.parse('Surface "matte')


# ---- scene_builder.rs unit tests ---------------------------------------------------------

def test_single_pass_build_with_attribute_blocks():
    b = build('''Format 320 240 1.0
        PixelSamples 2 2
        Projection "perspective" "fov" [45]
        WorldBegin
            LightSource "distantlight" 1 "from" [5 10 -10] "to" [0 0 0] "intensity" [0.8]
            AttributeBegin Color 0 0 1 Surface "plastic" Translate -2 0 5 Sphere 1 -1 1 360 AttributeEnd
            AttributeBegin Color 1 0 0 Surface "matte" Translate 2 0 5 Cylinder 0.5 -1 1 360 AttributeEnd
        WorldEnd''')
    assert n_objects(b) == 2 and len(b.lights) == 1 and len(b.materials) == 2
    assert b.pixel_samples == (2, 2) and b.spp == 4 and b.width == 320
    np.testing.assert_array_equal(d3(b.materials[0].diffuse_color), [0, 0, 1])
    np.testing.assert_array_equal(d3(b.materials[1].diffuse_color), [1, 0, 0])
    assert b.materials[0].diffuse_gain == 0.9 and b.materials[0].specular_ior == 1.5  # from_plastic
    d = np.array([-5.0, -10.0, 10.0])
    np.testing.assert_allclose(d3(b.lights[0].p), d / np.linalg.norm(d), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(b.quadrics[0][2][:, 3], [-2, 0, 5])


def test_attribute_end_restores_state():
    b = build('WorldBegin Color 1 1 1 AttributeBegin Color 0 1 0 Surface "metal" AttributeEnd '
              'Translate 0 0 5 Sphere 1 -1 1 360 WorldEnd')
    assert n_objects(b) == 1
    m = b.materials[0]
    np.testing.assert_array_equal(d3(m.diffuse_color), [1, 1, 1])
    assert m.diffuse_gain == 1.0 and m.specular_ior == 1.0  # matte, not metal


def test_new_quadrics_and_polygon():
    b = build('''WorldBegin Translate 0 0 5
        Torus 1 0.3 0 360 360  Disk 0 1 360  Paraboloid 1 0 1 360  Hyperboloid 1 0 -1 1 0 1 360
        Polygon "P" [-1 -1 0  1 -1 0  1 1 0  -1 1 0] WorldEnd''')
    assert n_objects(b) == 6
    assert [q[0] for q in b.quadrics] == [rrhip.QUADRIC_TORUS, rrhip.QUADRIC_DISK, rrhip.QUADRIC_PARABOLOID,
                                          rrhip.QUADRIC_HYPERBOLOID]
    assert b.tris[0][:3] == ((-1.0, -1.0, 5.0), (1.0, -1.0, 5.0), (1.0, 1.0, 5.0))


def test_concat_transform_matches_translate():
    a = build("WorldBegin ConcatTransform [1 0 0 0 0 1 0 0 0 0 1 0 3 4 5 1] Sphere 1 -1 1 360 WorldEnd")
    c = build("WorldBegin Translate 3 4 5 Sphere 1 -1 1 360 WorldEnd")
    np.testing.assert_array_equal(a.quadrics[0][2], c.quadrics[0][2])


def test_unknown_requests_strict_and_lenient():
    text = '''Option "searchpath" "string shader" ["/tmp"]
        Attribute "identifier" "name" ["ball"]
        GeometricApproximation "flatness" 0.5
        WorldBegin SubdivisionMesh "catmull-clark" [3] [0 1 2] [] [] [] [] Sphere 1 -1 1 360 WorldEnd'''
    with pytest.raises(rib.RibUnsupported):
        build(text)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        b = build(text, strict=False)
    assert n_objects(b) == 1 and len(w) == 2


def test_motion_block_takes_first_sample():
    """scene_builder.rs:2128-2143: the quadric keeps the block's first sample."""
    b = build("WorldBegin MotionBegin [0 1] Translate 0 0 5 Translate 0 0 9 MotionEnd Sphere 1 -1 1 360 WorldEnd")
    assert b.quadrics[0][2][2, 3] == 5.0
    with pytest.raises(rib.RibUnsupported):
        build('WorldBegin MotionBegin [0 1] Color 1 0 0 Color 0 1 0 MotionEnd WorldEnd')


def test_motion_block_moves_instances_and_camera():
    """Instances created after a transform block move from t0 to
    t1 * t0^-1 * placement (scene_builder.rs:1686-1692); a block before
    WorldBegin becomes the camera's motion, inverse(t1 * t0^-1)."""
    b = build('''MotionBegin [0 1] Translate 0 0 0 Translate 0.5 0 0 MotionEnd
        WorldBegin
        ObjectBegin "q" PointsPolygons [3] [0 1 2] "P" [0 0 0 1 0 0 0 1 0] ObjectEnd
        AttributeBegin MotionBegin [0 1] Translate 0 0 5 Translate 0 1 5 MotionEnd Scale 2 2 2
          PointsPolygons [3] [0 1 2] "P" [0 0 0 1 0 0 0 1 0]
          ObjectInstance "q" AttributeEnd
        PointsPolygons [3] [0 1 2] "P" [0 0 0 1 0 0 0 1 0] WorldEnd''')
    s = b.scene()
    assert s.camera.has_motion_inv == 1
    np.testing.assert_allclose(list(s.camera.motion_inv), [1, 0, 0, -0.5, 0, 1, 0, 0, 0, 0, 1, 0], atol=1e-15)
    (_, _, m0, m1), (_, _, i0, i1), (_, _, k0, k1) = b.instances
    np.testing.assert_allclose(m0[:, 3], [0, 0, 5])
    np.testing.assert_allclose(m0[0, 0], 2.0)
    np.testing.assert_allclose(m1[:, 3], [0, 1, 5])
    np.testing.assert_allclose(m1[:, :3], m0[:, :3])
    np.testing.assert_allclose(i1, m1)
    assert k1 is None  # the attribute block ended: static again
    assert [bool(x.moving) for x in s.instances] == [True, True, False]


def test_points_polygons_builds_mesh():
    b = build('''WorldBegin Translate 0 0 5
        PointsPolygons [4 4] [0 1 2 3  4 5 6 7]
          "P" [-1 -1 0  1 -1 0  1 1 0  -1 1 0  -1 -1 1  1 -1 1  1 1 1  -1 1 1] WorldEnd''')
    assert len(b.meshes) == 1 and len(b.instances) == 1
    assert b.meshes[0][1].shape == (4, 3)
    np.testing.assert_array_equal(b.meshes[0][1][:2], [[0, 1, 2], [0, 2, 3]])
    np.testing.assert_array_equal(b.instances[0][2][:, 3], [0, 0, 5])


def test_object_instance_shares_mesh():
    b = build('''WorldBegin
        ObjectBegin "quad" PointsPolygons [4] [0 1 2 3] "P" [-1 -1 0  1 -1 0  1 1 0  -1 1 0] ObjectEnd
        Translate -2 0 5 ObjectInstance "quad" Translate 4 0 0 ObjectInstance "quad" WorldEnd''')
    assert len(b.meshes) == 1 and len(b.instances) == 2
    assert b.instances[0][2][0, 3] == -2.0 and b.instances[1][2][0, 3] == 2.0
    assert b.instances[0][0] == b.instances[1][0] == 0


def test_inline_archive():
    b = build('''ArchiveBegin "ball" Sphere 1 -1 1 360 ArchiveEnd
        WorldBegin Translate 0 0 5 ReadArchive "ball" Translate 3 0 0 ReadArchive "ball" WorldEnd''')
    assert len(b.quadrics) == 2
    assert b.quadrics[0][2][0, 3] == 0.0 and b.quadrics[1][2][0, 3] == 3.0


# ---- the reference's fixtures ---------------------------------------------------------

def camera_fan(n=41):
    u = np.linspace(-0.32, 0.32, n)
    d = np.stack(np.meshgrid(u, u, indexing="ij"), -1).reshape(-1, 2)
    d = np.c_[d, np.ones(len(d))]
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def same_hits(a, b, label, tol=1e-9):
    """Closest hits of the two scenes' descriptions agree for every ray of
    a fan from the origin: hit / miss, distance, normal and the hit
    material's parameters."""
    da, db = a.desc(), b.desc()
    nhit = 0
    for d in camera_fan():
        ha = orc.scene_intersect(da, np.zeros(3), d)
        hb = orc.scene_intersect(db, np.zeros(3), d)
        assert (ha is None) == (hb is None), (label, d)
        if ha is None:
            continue
        nhit += 1
        assert ha[0] == pytest.approx(hb[0], abs=tol), (label, d)
        np.testing.assert_allclose(ha[2], hb[2], atol=1e-7)
        assert mat_fields(a.materials[ha[4]]) == mat_fields(b.materials[hb[4]]), (label, d)
    assert nhit > 100


def test_cornell_fixture_builds_the_restated_scene():
    s, spp = rib.load(os.path.join(RIBS, "cornell.rib"))
    ref = scenes.cornell_fixture(600, 600)
    assert spp == 64 and (s.camera.width, s.camera.height, s.camera.fov_deg) == (600, 600, 36.0)
    assert len(s.lights) == 1
    np.testing.assert_allclose(light_fields(s.lights[0]), light_fields(ref.lights[0]), rtol=0, atol=1e-15)
    # the ceiling emitter carries the light's index for MIS
    emit = [m for m in s.materials if d3(m.emission).max() > 0]
    assert len(emit) == 1 and emit[0].area_light == 0
    np.testing.assert_allclose(d3(emit[0].emission), 17.0 * np.array([1.0, 0.9, 0.75]))
    # every object has its own id (no identifiers in the file)
    assert sorted(m.id for m in s.materials) == list(range(1, len(s.materials) + 1))
    same_hits(s, ref, "cornell")


def test_quadric_zoo_fixture_builds_the_restated_scene():
    s, spp = rib.load(os.path.join(RIBS, "quadric_zoo.rib"))
    ref = scenes.quadric_zoo(960, 540)
    assert spp == 4
    for a, b in zip(s.lights, ref.lights):
        np.testing.assert_allclose(light_fields(a), light_fields(b), rtol=0, atol=1e-15)
    assert len(s.quadrics) == len(ref.quadrics) == 8
    for qa, qb in zip(s.quadrics, ref.quadrics):
        assert qa.kind == qb.kind
        for f in ("radius", "zmin", "zmax", "height", "thetamax"):
            assert getattr(qa, f) == getattr(qb, f), f
        np.testing.assert_array_equal(list(qa.p1), list(qb.p1))
        np.testing.assert_array_equal(list(qa.p2), list(qb.p2))
        np.testing.assert_allclose(list(qa.object_to_world), list(qb.object_to_world), rtol=0, atol=1e-15)
        assert mat_fields(s.materials[qa.material]) == mat_fields(ref.materials[qb.material])
    same_hits(s, ref, "zoo")


@pytest.mark.parametrize("name", ["minimal", "three_primitives", "transforms", "shadows", "materials", "bright",
                                  "simple_sphere"])
def test_small_fixtures_build_and_render(name):
    s, spp = rib.load(os.path.join(RIBS, name + ".rib"))
    s.camera.width, s.camera.height = 16, 12
    acc, _ = orc.render(s.desc(), rrhip.RenderParams.make(max_depth=4), 0, 1)
    assert np.isfinite(acc).all() and (acc[:, 6] == 1).all()


# ---- lights, materials, media, ids ---------------------------------------------------

def test_pxr_area_lights_placement():
    b = build('''WorldBegin
        AttributeBegin Translate 0 2 5 Rotate 90 1 0 0 Scale 2 3 1
          Light "PxrRectLight" "rect" "intensity" [4] "exposure" [1] "lightColor" [1 0.5 0.25] AttributeEnd
        AttributeBegin Translate 1 1 1 Scale 3 3 3 Light "PxrSphereLight" "s" "intensity" [2] AttributeEnd
        AttributeBegin Translate 0 0 4 Scale 2 2 2 Light "PxrDiskLight" "d" AttributeEnd
        AttributeBegin Rotate 90 0 1 0 Light "PxrDistantLight" "sun" "angleExtent" [2] AttributeEnd
        WorldEnd''')
    rect, sph, disk, sun = b.lights
    # Rotate 90 about x maps y to z: the unit square spans x in [-1, 1], z in [3.5, 6.5] at y = 2
    assert rect.kind == rrhip.LIGHT_RECT and rect.intensity == 8.0
    np.testing.assert_allclose(d3(rect.p), [-1, 2, 3.5], atol=1e-12)
    np.testing.assert_allclose(d3(rect.e1), [2, 0, 0], atol=1e-12)
    np.testing.assert_allclose(d3(rect.e2), [0, 0, 3], atol=1e-12)
    em = b.materials[b.tris[0][3]]
    assert em.area_light == 0 and em.diffuse_gain == 1.0 and d3(em.diffuse_color).max() == 0.0
    np.testing.assert_allclose(d3(em.emission), [8, 4, 2])
    corners = {tuple(np.round(v, 9)) for t in b.tris[:2] for v in t[:3]}
    assert corners == {(-1.0, 2.0, 3.5), (1.0, 2.0, 3.5), (1.0, 2.0, 6.5), (-1.0, 2.0, 6.5)}
    assert sph.kind == rrhip.LIGHT_SPHERE and sph.radius == pytest.approx(1.5)
    np.testing.assert_allclose(d3(sph.p), [1, 1, 1])
    assert disk.kind == rrhip.LIGHT_DISK
    np.testing.assert_allclose(d3(disk.e1), [1, 0, 0], atol=1e-12)
    kinds = [(q[0], b.materials[q[1]].area_light) for q in b.quadrics]
    assert kinds == [(rrhip.QUADRIC_SPHERE, 1), (rrhip.QUADRIC_DISK, 2)]
    # Rotate 90 about y maps -z to -x
    assert sun.kind == rrhip.LIGHT_DISTANT and sun.radius == pytest.approx(np.radians(1.0))
    np.testing.assert_allclose(d3(sun.p), [-1, 0, 0], atol=1e-12)


def test_pxr_surface_parameters_and_clamps():
    b = build('''WorldBegin Bxdf "PxrSurface" "m" "diffuseGain" [0.5] "diffuseColor" [0.1 0.2 0.3]
        "specularFaceColor" [0.9 0.8 0.7] "specularRoughness" [0.001] "specularIor" [1.8]
        "clearcoatGain" [0.4] "clearcoatRoughness" [3] "fuzzGain" [0.2] "glassRefractionGain" [0.6]
        "glassIor" [1.33] "glassRoughness" [0.3] "glowGain" [2] "glowColor" [0.5 1 0] "presence" [0.75]
        Sphere 1 -1 1 360
        Bxdf "PxrSurface" "plain" Sphere 1 -1 1 360 WorldEnd''')
    m, d = b.materials
    assert (m.diffuse_gain, m.specular_ior, m.clearcoat_gain, m.fuzz_gain) == (0.5, 1.8, 0.4, 0.2)
    assert (m.glass_gain, m.glass_ior, m.glass_roughness) == (0.6, 1.33, 0.3)
    assert m.specular_roughness == 0.005 and m.clearcoat_roughness == 1.0  # clamped to [0.005, 1]
    np.testing.assert_allclose(d3(m.glow), [1, 2, 0])
    assert m.cutout == pytest.approx(0.25)
    # PxrSurface defaults (pbr.rs:134-140)
    np.testing.assert_array_equal(d3(d.diffuse_color), [0.18, 0.18, 0.18])
    assert (d.diffuse_gain, d.specular_ior, d.specular_roughness) == (1.0, 1.5, 0.2)
    with pytest.raises(rib.RibUnsupported):
        build('WorldBegin Bxdf "PxrSurface" "s" "subsurfaceGain" [1] Sphere 1 -1 1 360 WorldEnd')


def test_media_and_object_ids():
    b = build('''WorldBegin Atmosphere "fog" "sigma_s" [0.2 0.2 0.2]
        AttributeBegin Attribute "identifier" "name" ["glass"] Interior "water" "sigma_a" [1 2 3] "g" [0.5]
          Sphere 1 -1 1 360 Translate 0 0 3 Sphere 1 -1 1 360 AttributeEnd
        Sphere 2 -2 2 360 WorldEnd''')
    fog, water = b.media
    assert b.atmosphere == 0 and fog["max_distance"] == 1200.0
    assert fog["sigma_a"] == (0.1, 0.1, 0.1) and fog["sigma_s"] == (0.2, 0.2, 0.2)
    assert water["max_distance"] == 0.0 and water["sigma_a"] == (1.0, 2.0, 3.0) and water["g"] == 0.5
    m0, m1, m2 = b.materials
    assert m0.interior == m1.interior == 2 and m2.interior == 0  # medium index + 1; 0 = none
    assert m0.id == m1.id == 1 and m2.id == 2  # identifier-named geometry shares an id


def test_dome_light_map_from_file():
    b = rib.Builder(base_dir=GOLD).run(rib.parse(
        'WorldBegin Light "PxrDomeLight" "sky" "intensity" [2] "lightColorMap" ["sky.hdr"] WorldEnd'))
    assert b.env.shape == (256, 512, 3) and b.lights[0].kind == rrhip.LIGHT_DOME
    assert b.lights[0].intensity == 2.0
    s = b.scene()
    d = s.desc()
    assert (d.env_width, d.env_height) == (512, 256)


# ---- curves, points and hair (scene_builder.rs:685-697, 1050-1170) ----------------------

def hair_ball_rib(n=400):
    """render.rs's hair parity scene as RIB text (tests/metal_pt_tests.rs:
    246-300), its control points from scenes.hair_strands."""
    strands = scenes.hair_strands(n)
    pts = " ".join("%.4f %.4f %.4f" % tuple(p) for s in strands for p in s)
    return f"""
        Format 96 96 1.0
        Projection "perspective" "fov" [35]
        Option "background" "color" [0.4 0.45 0.55]
        WorldBegin
            LightSource "distantlight" "sun" "from" [4 8 -6] "to" [0 0 6] "intensity" [1.4]
            AttributeBegin
                Color 0.6 0.6 0.6
                Polygon "P" [-30 -1.3 -5  30 -1.3 -5  30 -1.3 40  -30 -1.3 40]
            AttributeEnd
            AttributeBegin
                Bxdf "PxrMarschnerHair" "fur" "color" [0.5 0.3 0.12]
                    "roughness" [0.3] "azimuthalRoughness" [0.35]
                Translate 0 0.4 7
                Curves "linear" [{" ".join(["4"] * n)}] "nonperiodic"
                    "P" [{pts}]
                    "width" [0.03 0.008]
            AttributeEnd
        WorldEnd
    """


def test_hair_ball_through_rib_equals_scene():
    b = build(hair_ball_rib())
    s = b.scene()
    want = scenes.hair_ball()
    assert len(s.curve_sets) == 1 and s.curve_sets[0].num_segments == 1200  # metal_pt_tests.rs:303: > 1000
    got_c, want_c = s.curve_sets[0], want.curve_sets[0]
    for f in ("p0", "p1", "v0", "v1"):
        n = 4 if f in ("p0", "p1") else 1
        a = np.ctypeslib.as_array(ctypes_ptr(getattr(got_c, f)), (1200 * n,))
        w = np.ctypeslib.as_array(ctypes_ptr(getattr(want_c, f)), (1200 * n,))
        np.testing.assert_array_equal(a, w)
    (inst,) = [i for i in s.instances if i.kind == rrhip.INSTANCE_CURVES]
    np.testing.assert_allclose(np.array(inst.object_to_world).reshape(3, 4)[:, 3], [0, 0.4, 7])
    m = s.materials[inst.material]
    assert m.hair == 1 and m.hair_beta_m == 0.3 and m.hair_beta_n == 0.35 and m.hair_eta == 1.55
    np.testing.assert_allclose(list(m.hair_sigma_a), scenes.sigma_a_from_color((0.5, 0.3, 0.12), 0.35))
    np.testing.assert_allclose(list(m.diffuse_color), [0.35, 0.22, 0.12])  # the matte render.rs keeps
    np.testing.assert_allclose(list(s.lights[0].p), list(want.lights[0].p), atol=1e-15)
    assert tuple(s.background) == (0.4, 0.45, 0.55)


def ctypes_ptr(addr):
    import ctypes
    return ctypes.cast(addr, ctypes.POINTER(ctypes.c_float))


def test_cubic_curves_basis_shading_rate_and_points():
    b = build("""
        WorldBegin
        Basis "bezier" 3 "b-spline" 1
        ShadingRate 2
        Curves "cubic" [5] "nonperiodic" "P" [0 0 0  1 1 0  2 0 0  3 1 0  4 0 0] "constantwidth" [0.2]
        Bxdf "PxrMarschnerHair" "h" "absorption" [0.5] "ior" [1.6]
        Translate 1 2 3
        Points "P" [0 0 0  1 0 0  2 0 0] "width" [0.1 0.3]
        WorldEnd""")
    s = b.scene()
    assert len(s.curve_sets) == 2
    # b-spline step 1 over 5 points: 2 spans of 8 / ShadingRate = 4 segments
    want = scenes.dice_curve([[0, 0, 0], [1, 1, 0], [2, 0, 0], [3, 1, 0], [4, 0, 0]],
                             (rib._BASES["b-spline"][0], 1), 0.2, 0.2, 4)
    c = s.curve_sets[0]
    assert c.num_segments == 8 == len(want[0])
    np.testing.assert_array_equal(np.ctypeslib.as_array(ctypes_ptr(c.p0), (32,)), want[0].ravel())
    # points: zero-length segments, radius = width / 2 (constantwidth 0.05 past the widths), v = 0.5
    pt = s.curve_sets[1]
    p0 = np.ctypeslib.as_array(ctypes_ptr(pt.p0), (12,)).reshape(3, 4)
    np.testing.assert_array_equal(p0, np.ctypeslib.as_array(ctypes_ptr(pt.p1), (12,)).reshape(3, 4))
    np.testing.assert_allclose(p0[:, 3], [0.05, 0.15, 0.025])
    np.testing.assert_array_equal(np.ctypeslib.as_array(ctypes_ptr(pt.v0), (3,)), 0.5)
    insts = [i for i in s.instances if i.kind == rrhip.INSTANCE_CURVES]
    np.testing.assert_allclose(np.array(insts[1].object_to_world).reshape(3, 4)[:, 3], [1, 2, 3])
    m = s.materials[insts[1].material]
    assert m.hair == 1 and m.hair_eta == 1.6 and list(m.hair_sigma_a) == [0.5, 0.5, 0.5]
    assert s.materials[insts[0].material].hair == 0  # before the Bxdf: the default matte


def test_curve_requests_edge_cases():
    # periodic is diced open with a warning (scene_builder.rs:1063-1065)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        b = build('WorldBegin Curves "linear" [3] "periodic" "P" [0 0 0 1 0 0 2 0 0] WorldEnd')
    assert any("periodic" in str(x.message) for x in w)
    assert b.scene().curve_sets[0].num_segments == 2
    # inside an object definition curves are not instanced (reject_in_object_def)
    with pytest.raises(rib.RibUnsupported):
        build('ObjectBegin "a" Curves "linear" [2] "nonperiodic" "P" [0 0 0 1 0 0] ObjectEnd')
    # a cubic curve with fewer than 4 points dices to nothing
    assert len(build('WorldBegin Curves "cubic" [3] "nonperiodic" "P" [0 0 0 1 0 0 2 0 0] WorldEnd')
               .scene().curve_sets) == 0
    # a later PxrSurface clears the hair state
    s = build('WorldBegin Bxdf "PxrMarschnerHair" "h" Bxdf "PxrSurface" "s" '
              'Curves "linear" [2] "nonperiodic" "P" [0 0 0 1 0 0] WorldEnd').scene()
    assert s.materials[s.instances[0].material].hair == 0


# ---- patterns and textures ---------------------------------------------------------

def test_pattern_graph_binds_to_bxdf():
    # scene_builder.rs:1993-2027: a checker mixed 25 % towards green drives
    # diffuseColor; the checker's cell (0, 0) is red
    s, _ = rib.loads("""
        Format 64 64 1.0
        WorldBegin
            Pattern "PxrChecker" "check" "colorA" [1 0 0] "colorB" [0 0 1] "sScale" [4] "tScale" [4]
            Pattern "PxrMix" "blend" "reference color color1" ["check:resultRGB"] "color2" [0 1 0] "mix" [0.25]
            Bxdf "PxrSurface" "mat" "reference color diffuseColor" ["blend:resultRGB"] "specularIor" [1]
            Sphere 1 -1 1 360
        WorldEnd""")
    assert len(s.patterns) == 2
    assert [(b.material, b.field, b.pattern) for b in s.bindings] == [(len(s.materials) - 1, 0, 1)]
    c = orc.pattern_eval(s, 1, [0.1, 0.1])[0]
    assert abs(c[0] - 0.75) < 1e-9 and abs(c[1] - 0.25) < 1e-9


def test_pattern_request_defaults_and_inputs():
    # pattern_request (scene_builder.rs:1760-1848): checker scales default
    # to 8, fractal layers -> octaves, colour-correct input defaults to
    # grey 0.5, a single number is a grey constant, ramp axis "t"
    b = build("""
        WorldBegin
            Pattern "PxrChecker" "c"
            Pattern "PxrFractal" "f" "frequency" [2] "layers" [6] "gain" [0.4]
            Pattern "PxrColorCorrect" "cc" "gamma" [2]
            Pattern "PxrMix" "m" "color1" [0.2] "reference color mix" ["f:resultRGB"]
            Pattern "PxrRamp" "r" "axis" ["t"]
        WorldEnd""")
    (kc, c), (kf, f), (kcc, cc), (km, m), (kr, r) = b.patterns
    assert kc == rrhip.PATTERN_CHECKER and c["scale"] == (8.0, 8.0)
    assert c["color_a"] == (1.0, 1.0, 1.0) and c["color_b"] == (0.0, 0.0, 0.0)
    assert kf == rrhip.PATTERN_FRACTAL and f["octaves"] == 6 and f["gain"] == 0.4 and f["frequency"] == 2
    assert kcc == rrhip.PATTERN_COLOR_CORRECT and cc["input"][0] == -1 and cc["constant"][0] == (0.5, 0.5, 0.5)
    assert km == rrhip.PATTERN_MIX and list(m["input"]) == [-1, -1, 1]
    assert m["constant"][0] == (0.2, 0.2, 0.2) and m["constant"][1] == (1.0, 1.0, 1.0)
    assert kr == rrhip.PATTERN_RAMP and r["use_t"] == 1 and len(r["stops"]) == 2
    s = b.scene()
    # colour correct of the grey default: 0.5^(1/2)
    np.testing.assert_allclose(orc.pattern_eval(s, 2, [0, 0])[0], np.sqrt(0.5), rtol=1e-12)


def test_pattern_references_that_do_not_resolve():
    text = """
        WorldBegin
            Pattern "PxrChecker" "c"
            Pattern "PxrWorley" "w"
            Pattern "PxrTexture" "t"
            Bxdf "PxrSurface" "m" "reference color diffuseColor" ["nope:resultRGB"]
                "reference float bumpAmount" ["c:resultR"] "reference color specularFaceColor" ["t:resultRGB"]
            Sphere 1 -1 1 360
        WorldEnd"""
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        b = build(text)
    msgs = " ".join(str(x.message) for x in w)
    assert "PxrWorley" in msgs and "nope" in msgs and "bumpAmount" in msgs and "filename" in msgs
    assert len(b.patterns) == 2  # the texture node stays, magenta
    assert b.patterns[1][1]["texture"] == -1
    assert [(f, n) for _, f, n in b.bindings] == [("specularFaceColor", 1)]
    s = b.scene()
    np.testing.assert_array_equal(orc.pattern_eval(s, 1, [0.3, 0.3])[0], [1, 0, 1])
    # a bound presence is a binding like the others
    b = build("""WorldBegin Pattern "PxrChecker" "c"
                 Bxdf "PxrSurface" "m" "reference float presence" ["c:resultR"] Sphere 1 -1 1 360 WorldEnd""")
    assert [(f, n) for _, f, n in b.bindings] == [("presence", 0)]


def test_points_polygons_st():
    b = build("""
        WorldBegin
            PointsPolygons [3] [0 1 2] "P" [0 0 0 1 0 0 0 1 0] "st" [0 0 1 0 0 1]
            PointsPolygons [3] [0 1 2] "P" [0 0 0 1 0 0 0 1 0] "st" [0 0 1 0]
        WorldEnd""")
    np.testing.assert_array_equal(b.meshes[0][3], [[0, 0], [1, 0], [0, 1]])
    assert b.meshes[1][3] is None  # not one pair per point: ignored


def test_textured_still_life_builds(tmp_path):
    import rib_scenes
    path = rib_scenes.write_still_life(str(tmp_path))
    s, _ = rib.load(path)
    kinds = [p.kind for p in s.patterns]
    assert kinds == [rrhip.PATTERN_TEXTURE, rrhip.PATTERN_TEXTURE, rrhip.PATTERN_COLOR_CORRECT,
                     rrhip.PATTERN_TRIPLANAR, rrhip.PATTERN_FRACTAL, rrhip.PATTERN_CHECKER, rrhip.PATTERN_MIX,
                     rrhip.PATTERN_RAMP]
    # the four wood tiles and the label, each read once (the triplanar
    # ball reuses tile 1001)
    assert len(s.textures) == 5
    udim = np.ctypeslib.as_array(s.patterns[0].udim, (100,))
    assert [int(udim[t - 1001]) for t in (1001, 1002, 1011, 1012)] == [0, 1, 2, 3]
    assert (np.delete(udim, [0, 1, 10, 11]) == -1).all()
    assert s.patterns[3].texture == 0
    fields = sorted((b.material, rrhip.BIND_FIELDS[b.field]) for b in s.bindings)
    assert [f for _, f in fields] == ["diffuseColor"] * 3 + ["specularRoughness", "diffuseColor",
                                                             "diffuseColor"]
    a, _ = orc.render(s.desc(), rrhip.RenderParams.make(max_depth=4), 0, 2)
    assert np.isfinite(a).all() and a[:, :3].mean() > 0


REF_FIXTURES = "/root/reference/tests/fixtures"


@pytest.mark.skipif(not os.path.isdir(REF_FIXTURES), reason="the reference tree is not present")
@pytest.mark.parametrize("name", ["label.tex", "wood_1001.tex"])
def test_tex_writer_reproduces_reference_files(tmp_path, name):
    # level 0 of the reference's own .tex file, written again: every mip
    # level, tile and header byte as render.rs's tex.rs wrote them
    import rtex
    src = os.path.join(REF_FIXTURES, name)
    dst = str(tmp_path / name)
    rtex.write_tex(dst, rtex.read_tex(src))
    assert open(dst, "rb").read() == open(src, "rb").read()


def test_motion_block_deforms_points_polygons():
    # scene_builder.rs:1661-1672 and its transform / deformation motion test
    # (1945-1987): a block of PointsPolygons gives the mesh its last sample's
    # "P" as the shutter-close endpoint; a block whose samples differ in
    # vertex count keeps the first sample static
    b = build('''WorldBegin
        AttributeBegin Translate 0 0 8
          MotionBegin [0 1]
            PointsPolygons [4] [0 1 2 3] "P" [-1 -1 0  1 -1 0  1 1 0  -1 1 0]
            PointsPolygons [4] [0 1 2 3] "P" [-1 -1 2  1 -1 2  1 1 2  -1 1 2]
          MotionEnd
        AttributeEnd
        MotionBegin [0 1]
          PointsPolygons [3] [0 1 2] "P" [0 0 0 1 0 0 0 1 0]
          PointsPolygons [3] [0 1 2] "P" [0 0 0 1 0 0 0 1 0 5 5 5]
        MotionEnd
        WorldEnd''')
    pos, _, _, _, pos1 = b.meshes[0]
    np.testing.assert_allclose(pos1[:, 2] - pos[:, 2], 2.0)
    np.testing.assert_allclose(pos1[:, :2], pos[:, :2])
    assert b.meshes[1][4] is None
    s = b.scene()
    # the oracle sees the quad at z = 8 at shutter open and z = 10 at close
    t, *_ = orc.scene_intersect_n(s.desc(), [[0, 0, 0, 0, 0, 1]])
    assert abs(t[0] - 8.0) < 1e-9


def test_procedural_requests(tmp_path):
    # Procedural "DelayedReadArchive" reads its archive at once; "RunProgram"
    # replays what the program prints; "DynamicLoad" has no counterpart
    import os
    import stat
    (tmp_path / "part.rib").write_text('Sphere 0.5 -0.5 0.5 360\n')
    prog = tmp_path / "gen.sh"
    prog.write_text('#!/bin/sh\necho "Translate $1 0 0 Sphere 0.25 -0.25 0.25 360"\n')
    prog.chmod(prog.stat().st_mode | stat.S_IEXEC)
    (tmp_path / "scene.rib").write_text('''WorldBegin
        Procedural "DelayedReadArchive" ["part.rib"] [-1 1 -1 1 -1 1]
        AttributeBegin Procedural "RunProgram" ["gen.sh" "3"] [-1 1 -1 1 -1 1] AttributeEnd
        WorldEnd''')
    s, _ = rib.load(str(tmp_path / "scene.rib"), strict=False)
    assert len(s.quadrics) == 2
    assert sorted(q.radius for q in s.quadrics) == [0.25, 0.5]
    assert any(abs(q.object_to_world[3] - 3.0) < 1e-12 for q in s.quadrics)
    with pytest.raises(rib.RibUnsupported):
        rib.loads('WorldBegin Procedural "DynamicLoad" ["x.so" ""] [-1 1 -1 1 -1 1] WorldEnd')
    assert os.path.exists(tmp_path / "part.rib")


def test_media_density_parameters():
    # Interior / Atmosphere "density" "fbm" | "noise" (scene_builder.rs:
    # 328-373): render.rs's defaults, overrides, another name stays
    # homogeneous with a warning
    s, _ = rib.loads('''WorldBegin
        Atmosphere "haze" "density" "noise"
        AttributeBegin Interior "c" "density" "fbm" "frequency" [1.5] "octaves" [3] "gain" [0.6]
          "lacunarity" [2.5] "coverage" [0.7] "sharpness" [2] Sphere 1 -1 1 360 AttributeEnd
        WorldEnd''')
    a, c = s.media
    assert (a.procedural, a.octaves, a.frequency, a.gain, a.lacunarity, a.coverage, a.sharpness) == \
        (rrhip.MEDIUM_FBM, 5, 0.3, 0.55, 2.0, 0.5, 4.0)
    assert (c.procedural, c.octaves, c.frequency, c.gain, c.lacunarity, c.coverage, c.sharpness) == \
        (rrhip.MEDIUM_FBM, 3, 1.5, 0.6, 2.5, 0.7, 2.0)
    with pytest.warns(UserWarning):
        s2, _ = rib.loads('WorldBegin Atmosphere "h" "density" "voxels" WorldEnd')
    assert s2.media[0].procedural == 0 and s2.media[0].nx == 0
