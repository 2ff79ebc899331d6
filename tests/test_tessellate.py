"""Patch tessellation of the RIB front end (render.rs_amd/tessellate.py)
against render.rs src/geometry/patches.rs:

* patches.rs's own tests: a Bézier sheet interpolates its corner control
  points, bulges below its raised cage and dices to 81 vertices; a 7 x 4
  B-spline grid (4 x 1 patches) is one welded sheet of 17 x 5 vertices
  and 128 triangles; a flat quadratic NURBS net evaluates to its plane
  within its hull and interpolates its clamped corners;
* closed forms: a bicubic Bézier patch against de Casteljau's algorithm
  in both directions, bilinear patches reproducing their corner lerps;
  render.rs's NURBS basis recursion (patches.rs:228-260) worked by hand
  for order 2 and 3 on clamped knots — it updates its values in place
  from the span down, so it is not the textbook basis (order 2 gives
  [t (1 - t), t] instead of [1 - t, t]) and the tessellation keeps that
  to produce render.rs's surfaces;
* the RIB requests: Basis u / v steps, Patch, periodic PatchMesh seams,
  NuPatch "Pw", and the reference's stilllife.rib fixture building its
  cloth and band from PatchMesh / NuPatch;
* subdivision (subdiv.rs): one round against the Catmull-Clark rules,
  infinite creases and corners, holes, the pan.rib crate;
* polygons with holes (earclip.rs's own L and square-with-hole cases by
  closed-form area; random concave outlines with holes on tilted planes
  covered exactly once per point inside, brute-force even-odd, nothing
  outside), GeneralPolygon / PointsGeneralPolygons requests and their
  deformation blocks.
"""
import math
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "render.rs_amd"))

import rib  # noqa: E402
import tessellate as T  # noqa: E402

BEZIER = rib._BASES["bezier"][0]
BSPLINE = rib._BASES["b-spline"][0]


def test_bezier_patch_interpolates_corners():
    # patches.rs:341-361
    pts = []
    for j in range(4):
        for i in range(4):
            pts += [i, j, 1.0 if (1 <= i <= 2 and 1 <= j <= 2) else 0.0]
    pos, tris, nrm, st = T.tessellate_bicubic(pts, 4, 4, False, False, BEZIER, 3, BEZIER, 3, 8)
    assert abs(pos[0][0]) < 1e-5 and abs(pos[0][1]) < 1e-5
    assert abs(pos[-1][0] - 3) < 1e-5 and abs(pos[-1][1] - 3) < 1e-5
    assert 0.3 < pos[:, 2].max() < 1.0
    assert len(pos) == 81


def test_bspline_mesh_multiple_patches_no_cracks():
    # patches.rs:363-377
    pts = []
    for j in range(4):
        for i in range(7):
            pts += [i, j, ((i * 3 + j) % 2) * 0.3]
    pos, tris, nrm, st = T.tessellate_bicubic(pts, 7, 4, False, False, BSPLINE, 1, BSPLINE, 1, 4)
    assert len(pos) == (4 * 4 + 1) * (1 * 4 + 1)
    assert len(tris) == (4 * 4) * 4 * 2
    # welded: every interior grid vertex is shared by six triangles
    use = np.bincount(tris.ravel(), minlength=len(pos)).reshape(5, 17)
    assert (use[1:-1, 1:-1] == 6).all()


def test_nurbs_flat_sheet():
    # patches.rs:379-416
    uknot = [0, 0, 0, 0.5, 1, 1, 1]
    vknot = [0, 0, 0, 1, 1, 1]
    pts = [c for j in range(3) for i in range(4) for c in (i, j, 0.0)]
    pos, tris, nrm, st = T.tessellate_nurbs(4, 3, uknot, 0, 1, 3, 3, vknot, 0, 1, pts, False, 8, 8)
    assert np.abs(pos[:, 2]).max() < 1e-5
    assert pos[:, 0].min() >= -1e-5 and pos[:, 0].max() <= 3 + 1e-4
    assert abs(pos[0][0]) < 1e-5 and abs(pos[0][1]) < 1e-5


def de_casteljau(ctrl, t):
    p = [np.asarray(c, float) for c in ctrl]
    while len(p) > 1:
        p = [p[i] * (1 - t) + p[i + 1] * t for i in range(len(p) - 1)]
    return p[0]


def test_bicubic_bezier_against_de_casteljau():
    rng = np.random.default_rng(3)
    ctrl = rng.normal(size=(4, 4, 3))
    segs = 6
    pos, _, _, st = T.tessellate_bicubic(ctrl.ravel(), 4, 4, False, False, BEZIER, 3, BEZIER, 3, segs)
    for k, (u, v) in enumerate(st.astype(float)):
        rows = [de_casteljau(ctrl[j], u) for j in range(4)]
        want = de_casteljau(rows, v)
        np.testing.assert_allclose(pos[k], want, atol=2e-6)


def test_bilinear_reproduces_corner_lerps():
    ctrl = np.array([[0, 0, 0], [2, 0, 1], [0, 3, 0], [2, 3, -1]], float)
    pos, tris, _, st = T.tessellate_bilinear(ctrl.ravel(), 2, 2, False, False, 4)
    u, v = st[:, 0:1].astype(float), st[:, 1:2].astype(float)
    want = (ctrl[0] * (1 - u) + ctrl[1] * u) * (1 - v) + (ctrl[2] * (1 - u) + ctrl[3] * u) * v
    np.testing.assert_allclose(pos, want, atol=1e-6)
    assert len(tris) == 4 * 4 * 2


def test_nurbs_basis_follows_render_rs_recursion():
    # order 2, knots 0 0 1 1: span 1; level 1 sets b1 = t, then b0 reads the
    # already-raised b1: b0 = b1 (1 - t) = t (1 - t)
    for t in (0.0, 0.2, 0.5, 0.9):
        np.testing.assert_allclose(T.bspline_basis([0, 0, 1, 1], 2, 2, t), [t * (1 - t), t], atol=1e-15)
    # order 3, knots 0 0 0 1 1 1: span 2; level 1: b2 = t, b1 = t (1 - t);
    # level 2: b2 = t * t, b1 = b1 t + b2 (1 - t) = t (1 - t) t + t^2 (1 - t),
    # b0 = b1 (1 - t) with the new b1
    for t in (0.0, 0.3, 0.7, 1.0):
        b2 = t * t
        b1 = t * (1 - t) * t + b2 * (1 - t)
        b0 = b1 * (1 - t)
        np.testing.assert_allclose(T.bspline_basis([0, 0, 0, 1, 1, 1], 3, 3, t), [b0, b1, b2], atol=1e-15)
    # hence a rational quadratic arc is not the textbook circle: points come
    # out as the weighted sums of this basis
    w = math.sqrt(0.5)
    arc = [(1, 0, 1), (w, w, w), (0, 1, 1)]  # homogeneous (x w, y w, w)
    pts = [c for j in range(2) for (x, y, ww) in arc for c in (x, y, j * ww, ww)]
    pos, _, _, _ = T.tessellate_nurbs(3, 3, [0, 0, 0, 1, 1, 1], 0, 1, 2, 2, [0, 0, 1, 1], 0, 1, pts, True, 4, 2)
    t = 0.5
    bu = T.bspline_basis([0, 0, 0, 1, 1, 1], 3, 3, t)
    bv = T.bspline_basis([0, 0, 1, 1], 2, 2, 0.5)
    hom = np.array([[x, y, j * ww, ww] for j in range(2) for (x, y, ww) in arc]).reshape(2, 3, 4)
    acc = sum(bv[j] * bu[i] * hom[j, i] for j in range(2) for i in range(3))
    np.testing.assert_allclose(pos[1 * 5 + 2], acc[:3] / acc[3], atol=1e-6)


def test_rib_patch_requests():
    b = rib.Builder().run(rib.parse('''WorldBegin
        ShadingRate 4
        Basis "b-spline" 1 "catmull-rom" 1
        PatchMesh "bicubic" 6 "periodic" 4 "nonperiodic"
          "P" [''' + " ".join(f"{math.cos(i * math.pi / 3):.6f} {math.sin(i * math.pi / 3):.6f} {j}"
                               for j in range(4) for i in range(6)) + ''']
        Patch "bilinear" "P" [0 0 0  1 0 0  0 1 0  1 1 0]
        NuPatch 3 3 [0 0 0 1 1 1] 0 1 2 2 [0 0 1 1] 0 1
          "Pw" [1 0 0 1  0.7071068 0.7071068 0 0.7071068  0 1 0 1  1 0 1 1  0.7071068 0.7071068 0.7071068 0.7071068  0 1 1 1]
        WorldEnd'''))
    assert b.state.basis_u[1] == 1 and b.state.basis_v[1] == 1
    (pos, tris, nrm, st, _), (pb, tb, _, _, _), (pn, tn, _, _, _) = b.meshes
    segs = 4  # 16 / ShadingRate 4
    # periodic in u: 6 patches x 4 spans around, no seam column; 1 patch in v
    assert len(pos) == 6 * segs * (1 * segs + 1)
    assert len(tris) == 6 * segs * segs * 2
    r = np.sqrt(pos[:, 0] ** 2 + pos[:, 1] ** 2)
    assert r.std() < 0.05  # a closed tube around the axis
    assert len(pb) == (segs + 1) ** 2
    assert len(pn) == 5 * 5  # order 3 / 2, one span each: 4 segments (the minimum) per direction
    assert len(b.instances) == 3


REF = "/root/reference/tests/fixtures/stilllife.rib"


@pytest.mark.skipif(not os.path.exists(REF), reason="the reference tree is not present")
def test_stilllife_fixture_builds(tmp_path):
    # render.rs's UDIM still life: its cloth is a b-spline PatchMesh and its
    # band a NuPatch; with the fixture's textures beside it
    for f in ("stilllife.rib", "sky.hdr", "label.tex", "wood_1001.tex", "wood_1002.tex", "wood_1011.tex",
              "wood_1012.tex"):
        shutil.copy(os.path.join(os.path.dirname(REF), f), tmp_path / f)
    s, spp = rib.load(str(tmp_path / "stilllife.rib"))
    assert len(s.textures) == 5 and len(s.patterns) == 8
    assert len(s.meshes) == 4  # deck, crate, cloth, band
    cloth, band = s.meshes[2], s.meshes[3]
    assert cloth.num_vertices == (4 * 32 + 1) ** 2  # 7x7 b-spline grid: 4x4 patches, 32 spans (ShadingRate 0.5)
    # NuPatch 7 x 3 of order 3: 32 x (7 - 3 + 1) segments clamped to 128 in u, 32 x 1 in v
    assert band.num_vertices == (128 + 1) * (32 + 1)


# ---- Catmull-Clark (subdiv.rs) -------------------------------------------------------

CUBE = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]],
                float)
CUBE_FACES = [[0, 1, 2, 3], [5, 4, 7, 6], [4, 0, 3, 7], [1, 5, 6, 2], [3, 2, 6, 7], [4, 5, 1, 0]]
RING = [(0, 1, 1e9), (1, 2, 1e9), (2, 3, 1e9), (3, 0, 1e9)]


def test_cube_rounds_toward_sphere():
    # subdiv.rs:393-410
    pos, tris, _, _ = T.tessellate_subdiv(CUBE, CUBE_FACES, interpolate_boundary=True, levels=3)
    assert len(tris) > 500
    r = np.linalg.norm(pos.astype(float), axis=1)
    assert r.max() < 1.5 and r.max() / r.min() < 1.35


def test_infinite_crease_keeps_edge_sharp():
    # subdiv.rs:412-440
    pos, _, _, _ = T.tessellate_subdiv(CUBE, CUBE_FACES, RING, interpolate_boundary=True, levels=3)
    on = np.abs(pos[:, 2] + 1) < 1e-4
    assert on.sum() > 20
    assert np.maximum(np.abs(pos[on, 0]), np.abs(pos[on, 1])).max() > 0.9


def test_corner_tags_pin_vertices():
    # subdiv.rs:442-460
    pos, _, _, _ = T.tessellate_subdiv(CUBE, CUBE_FACES, RING, [(v, 1e9) for v in range(4)], interpolate_boundary=True,
                                       levels=3)
    pinned = ((np.abs(np.abs(pos[:, 0]) - 1) < 1e-5) & (np.abs(np.abs(pos[:, 1]) - 1) < 1e-5) &
              (np.abs(pos[:, 2] + 1) < 1e-5))
    assert pinned.sum() >= 4


def test_hole_faces_removed():
    # subdiv.rs:462-469
    with_hole = len(T.tessellate_subdiv(CUBE, CUBE_FACES, holes=[0], interpolate_boundary=True, levels=1)[1])
    without = len(T.tessellate_subdiv(CUBE, CUBE_FACES, interpolate_boundary=True, levels=1)[1])
    assert with_hole < without


def test_one_round_against_the_catmull_clark_rules():
    # a closed cube, one round: face points are face centroids, edge points
    # (a + b + f1 + f2) / 4, vertex points (F + 2M + (n - 3) P) / n with
    # n = 3; the result has 24 quads over 26 vertices
    pos, tris, _, _ = T.tessellate_subdiv(CUBE, CUBE_FACES, levels=1)
    assert len(pos) == 6 + 12 + 8 and len(tris) == 48
    p = pos.astype(float)
    for c in ([0, 0, -1], [1, 0, 0], [0, 1, 0]):  # face points
        assert np.min(np.linalg.norm(p - c, axis=1)) < 1e-6
    # edge (0, 1): (-1,-1,-1) and (1,-1,-1) with faces z=-1 and y=-1
    want_e = (np.array([-1, -1, -1]) + [1, -1, -1] + [0, 0, -1] + [0, -1, 0]) / 4
    assert np.min(np.linalg.norm(p - want_e, axis=1)) < 1e-6
    # vertex 0: F = average of 3 face points, M = average of 3 edge midpoints
    F = (np.array([0, 0, -1]) + [-1, 0, 0] + [0, -1, 0]) / 3
    M = (np.array([0, -1, -1]) + [-1, 0, -1] + [-1, -1, 0]) / 3
    want_v = (F + 2 * M + 0 * CUBE[0]) / 3
    assert np.min(np.linalg.norm(p - want_v, axis=1)) < 1e-6


def test_rib_subdivision_mesh_tags():
    b = rib.Builder().run(rib.parse('''WorldBegin ShadingRate 8
        SubdivisionMesh "catmull-clark" [4 4 4 4 4 4] [0 1 2 3  5 4 7 6  4 0 3 7  1 5 6 2  3 2 6 7  4 5 1 0]
          ["crease" "corner" "hole" "interpolateboundary"] [5 1  1 1  1 0  0 0] [0 1 2 3 0  0  5] [20 20]
          "P" [-1 -1 -1  1 -1 -1  1 1 -1  -1 1 -1  -1 -1 1  1 -1 1  1 1 1  -1 1 1]
        WorldEnd'''))
    pos, tris, nrm, st, _ = b.meshes[0]
    assert st is None and nrm is not None
    # ShadingRate 8 -> one level; the hole face (5) drops its 4 quads
    assert len(tris) == (6 - 1) * 4 * 2
    pinned = np.all(np.abs(pos - [-1, -1, -1]) < 1e-6, axis=1)
    assert pinned.any()  # vertex 0 is a corner of sharpness 20 (infinite)


@pytest.mark.skipif(not os.path.exists("/root/reference/tests/fixtures/pan.rib"),
                    reason="the reference tree is not present")
def test_pan_fixture_builds(tmp_path):
    # render.rs's pan.rib: a subdivided crate among moving and deforming
    # geometry, UDIM wood and a label
    src = os.path.dirname(REF)
    for f in ("pan.rib", "sky.hdr", "label.tex", "wood_1001.tex", "wood_1002.tex", "wood_1011.tex",
              "wood_1012.tex"):
        if os.path.exists(os.path.join(src, f)):
            shutil.copy(os.path.join(src, f), tmp_path / f)
    s, _ = rib.load(str(tmp_path / "pan.rib"))
    assert any(m.positions1 for m in s.meshes)  # the flapping banner deforms
    assert any(i.moving for i in s.instances)


def test_gen_fbm_equals_the_oracle():
    # the host fBm the RIB Displace request uses equals the oracle's
    # restatement of displace.rs (itself pinned by displace.rs's tests)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    import rrhip
    p = np.random.default_rng(4).uniform(-5, 5, (500, 3))
    got = rrhip.gen_fbm(p, 1.6, 5, 0.55, 2.0, (0.3, -1.2, 7.0))
    want = orc.fbm(p, 1.6, 5, 0.55, 2.0, (0.3, -1.2, 7.0))
    np.testing.assert_array_equal(got, want)


def test_rib_displace_moves_vertices_along_normals():
    # displace_mesh (displace.rs:101-123): each vertex moves by amplitude *
    # fbm(p) along its smooth normal; the attribute block scopes it
    import rrhip
    text = '''WorldBegin ShadingRate 8
        AttributeBegin
        Displace "noise" "amplitude" [0.3] "frequency" [1.5] "octaves" [3]
        SubdivisionMesh "catmull-clark" [4 4 4 4 4 4] [0 1 2 3  5 4 7 6  4 0 3 7  1 5 6 2  3 2 6 7  4 5 1 0]
          "P" [-1 -1 -1  1 -1 -1  1 1 -1  -1 1 -1  -1 -1 1  1 -1 1  1 1 1  -1 1 1]
        AttributeEnd
        SubdivisionMesh "catmull-clark" [4 4 4 4 4 4] [0 1 2 3  5 4 7 6  4 0 3 7  1 5 6 2  3 2 6 7  4 5 1 0]
          "P" [-1 -1 -1  1 -1 -1  1 1 -1  -1 1 -1  -1 -1 1  1 -1 1  1 1 1  -1 1 1]
        WorldEnd'''
    b = rib.Builder().run(rib.parse(text))
    (pd, td, nd, _, _), (p0, t0, n0, _, _) = b.meshes
    d = rrhip.gen_fbm(p0.astype(float), 1.5, 3, 0.5, 2.0) * 0.3
    np.testing.assert_allclose(pd, (p0.astype(float) + n0.astype(float) * d[:, None]).astype(np.float32), atol=1e-6)
    np.testing.assert_array_equal(td, t0)
    assert np.abs(pd - p0).max() > 1e-3


def _tri_areas(p, t):
    p = np.asarray(p, np.float64)
    a, b, c = p[t[:, 0]], p[t[:, 1]], p[t[:, 2]]
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)


def _even_odd(xy, loops, pts):
    """Brute-force point-in-polygon over all loops (even-odd rule)."""
    inside = np.zeros(len(pts), bool)
    for lp in loops:
        v = xy[lp]
        w = np.roll(v, -1, axis=0)
        for (x0, y0), (x1, y1) in zip(v, w):
            cross = (y0 > pts[:, 1]) != (y1 > pts[:, 1])
            xint = x0 + (pts[:, 1] - y0) * (x1 - x0) / np.where(y1 != y0, y1 - y0, 1.0)
            inside ^= cross & (pts[:, 0] < xint)
    return inside


def _covered(xy, tris, pts):
    """Number of triangles (2D) holding each point."""
    n = np.zeros(len(pts), int)
    for a, b, c in xy[tris]:
        d1 = (b[0] - a[0]) * (pts[:, 1] - a[1]) - (b[1] - a[1]) * (pts[:, 0] - a[0])
        d2 = (c[0] - b[0]) * (pts[:, 1] - b[1]) - (c[1] - b[1]) * (pts[:, 0] - b[0])
        d3 = (a[0] - c[0]) * (pts[:, 1] - c[1]) - (a[1] - c[1]) * (pts[:, 0] - c[0])
        n += ((d1 > 0) & (d2 > 0) & (d3 > 0)) | ((d1 < 0) & (d2 < 0) & (d3 < 0))
    return n


def test_concave_outline_closed_form():
    # earclip.rs's own case: an L of area 3 cut into 4 triangles, all
    # counter-clockwise in the plane
    p = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [1, 1, 0], [1, 2, 0], [0, 2, 0]], float)
    t = T.triangulate_loops(p, [list(range(6))])
    assert t.shape == (4, 3)
    assert abs(_tri_areas(p, t).sum() - 3.0) < 1e-12
    e = p[t]
    assert (np.cross(e[:, 1] - e[:, 0], e[:, 2] - e[:, 0])[:, 2] > 0).all()


def test_square_with_hole_closed_form():
    # earclip.rs's own case: a 4x4 square less a centred 2x2 hole, area 12;
    # no triangle lies in the hole and bridges reuse the loops' vertices
    p = np.array([[0, 0, 0], [4, 0, 0], [4, 4, 0], [0, 4, 0], [1, 1, 0], [3, 1, 0], [3, 3, 0], [1, 3, 0]], float)
    t = T.triangulate_loops(p, [[0, 1, 2, 3], [4, 5, 6, 7]])
    assert abs(_tri_areas(p, t).sum() - 12.0) < 1e-9
    assert len(t) == 8  # n + 2h - 2 triangles for 8 vertices, one hole
    c = p[t].mean(1)
    assert not ((c[:, 0] > 1) & (c[:, 0] < 3) & (c[:, 1] > 1) & (c[:, 1] < 3)).any()


@pytest.mark.parametrize("seed", range(6))
def test_random_polygons_with_holes_cover_exactly(seed):
    # star-shaped concave outlines with up to three disjoint square holes,
    # on a tilted plane: the triangles cover each sample point inside the
    # polygon (even-odd over all loops) exactly once, and nothing outside;
    # areas add up to outline minus holes
    rng = np.random.default_rng(seed)
    k = 7 + seed * 3
    ang = np.sort(rng.uniform(0, 2 * np.pi, k))
    rad = rng.uniform(4.0, 8.0, k)
    xy = np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)
    loops = [list(range(k))]
    holes = [(-1.5, -1.5), (1.2, 0.3), (-0.8, 1.6)][:seed % 4]
    for cx, cy in holes:
        s = 0.45
        base = len(xy)
        xy = np.vstack([xy, [[cx - s, cy - s], [cx + s, cy - s], [cx + s, cy + s], [cx - s, cy + s]]])
        loops.append(list(range(base, base + 4)))
    if seed % 2:
        loops[0] = loops[0][::-1]  # clockwise outline
    # embed in a plane whose normal has a negative dominant component
    rot = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    p3 = np.column_stack([xy, np.zeros(len(xy))]) @ rot.T + rng.normal(size=3)
    t = T.triangulate_loops(p3, loops)
    assert t is not None and len(t) == sum(len(lp) for lp in loops) + 2 * (len(loops) - 1) - 2
    a = sum(abs(0.5 * np.sum(xy[lp, 0] * np.roll(xy[lp, 1], -1) - np.roll(xy[lp, 0], -1) * xy[lp, 1]))
            for lp in loops[:1]) - len(holes) * 0.9 ** 2
    assert abs(_tri_areas(p3, t).sum() - a) < 1e-9
    pts = rng.uniform(-8.5, 8.5, (4000, 2))
    inside = _even_odd(xy, loops, pts)
    np.testing.assert_array_equal(_covered(xy, t, pts), inside.astype(int))


def test_rib_general_polygon_requests():
    b = rib.Builder().run(rib.parse('''WorldBegin
        GeneralPolygon [4 4] "P" [0 0 0  4 0 0  4 4 0  0 4 0   1 1 0  3 1 0  3 3 0  1 3 0]
        PointsGeneralPolygons [1 2] [4 4 4] [0 1 2 3  4 5 6 7  8 9 10 11]
          "P" [0 0 1  1 0 1  1 1 1  0 1 1   0 0 2  4 0 2  4 4 2  0 4 2   1 1 2  3 1 2  3 3 2  1 3 2]
          "st" [0 0  1 0  1 1  0 1   0 0  1 0  1 1  0 1   .25 .25  .75 .25  .75 .75  .25 .75]
        PointsPolygons [4] [0 1 2 3] "P" [0 0 1  1 0 1  1 1 1  0 1 1]
        WorldEnd'''))
    (pg, tg, ng, sg, _), (pp, tp, npg, sp, _), (p1, t1, _, _, _) = b.meshes
    assert sg is None and ng is not None
    assert abs(_tri_areas(pg, tg).sum() - 12.0) < 1e-6
    assert sp is not None and sp.shape == (12, 2)
    # a one-loop polygon is fanned exactly as PointsPolygons fans it
    np.testing.assert_array_equal(tp[:2], t1)
    assert abs(_tri_areas(pp, tp[2:]).sum() - 12.0) < 1e-6
    assert len(b.instances) == 3


def test_rib_points_general_polygons_deformation_block():
    b = rib.Builder().run(rib.parse('''WorldBegin
        MotionBegin [0 1]
          PointsGeneralPolygons [1] [3] [0 1 2] "P" [0 0 0  1 0 0  0 1 0]
          PointsGeneralPolygons [1] [3] [0 1 2] "P" [0 0 1  1 0 1  0 1 1]
        MotionEnd
        WorldEnd'''))
    pos, tri, _, _, pos1 = b.meshes[0]
    np.testing.assert_array_equal(pos1, pos + np.float32([0, 0, 1]))
