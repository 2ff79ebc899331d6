"""Pins the oracle's BVH traversal and triangle test (render.rs
src/accel/bvh.rs, src/geometry/mesh.rs) with the reference's own KATs:
brute force agreement (bvh.rs:403-453), any-hit limits (bvh.rs:455-472),
the cube instance closed forms (mesh.rs:473-527)."""
import numpy as np

import orc

RAY = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("tmin", "<f4"),
                ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("tmax", "<f4")])


class RefLcg:
    """The LCG of bvh.rs:406-410."""

    def __init__(self, state=0x12345678):
        self.s = state

    def rnd(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & ((1 << 64) - 1)
        return ((self.s >> 11) / float(1 << 53)) * 8.0 - 4.0


def lcg_soup(ntris=500, nrays=2000):
    g = RefLcg()
    tris = []
    for _ in range(ntris):
        b = [g.rnd(), g.rnd(), g.rnd()]
        p1 = [b[0] + g.rnd() * 0.2, b[1] + g.rnd() * 0.2, b[2] + g.rnd() * 0.2]
        p2 = [b[0] + g.rnd() * 0.2, b[1] + g.rnd() * 0.2, b[2] + g.rnd() * 0.2]
        tris.append([b, p1, p2])
    pos = np.array(tris, np.float32).reshape(-1, 3)
    idx = np.arange(3 * ntris, dtype=np.uint32).reshape(-1, 3)
    rays = []
    for _ in range(nrays):
        o = [g.rnd(), g.rnd(), g.rnd()]
        d = np.array([g.rnd(), g.rnd(), g.rnd()])
        ln = np.linalg.norm(d)
        if ln < 1e-6:
            continue
        d = d / ln  # Ray::new normalises (ray.rs:11-17)
        rays.append((o[0], o[1], o[2], 1e-9, d[0], d[1], d[2], np.inf))
    return pos, idx, np.array(rays, RAY)


def cube(scale=1.0, dz=5.0):
    p = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], np.float64)
    p = p * scale + np.array([0, 0, dz])
    quads = [[0, 1, 2, 3], [5, 4, 7, 6], [4, 0, 3, 7], [1, 5, 6, 2], [3, 2, 6, 7], [4, 5, 1, 0]]
    idx = []
    for q in quads:
        idx += [[q[0], q[1], q[2]], [q[0], q[2], q[3]]]
    return p.astype(np.float32), np.array(idx, np.uint32)


def z_ray(tmax=np.inf):
    return np.array([(0, 0, 0, 1e-9, 0, 0, 1, tmax)], RAY)


def test_bvh_matches_brute_force_reference_params():
    pos, idx, rays = lcg_soup()
    m = orc.Mesh(pos, idx, bins=16, max_leaf=4, sah_leaf_mult=4)
    bp, bt = orc.brute_closest(pos, idx, rays)
    p, t, _, _ = m.trace_closest(rays)
    assert (bp != 0xFFFFFFFF).sum() > 100
    assert np.array_equal(bp, p)
    hit = bp != 0xFFFFFFFF
    assert np.abs(bt[hit] - t[hit]).max() < 1e-9


def test_bvh_matches_brute_force_query_params():
    pos, idx, rays = lcg_soup()
    m = orc.Mesh(pos, idx, bins=32, max_leaf=4, sah_leaf_mult=1)
    bp, bt = orc.brute_closest(pos, idx, rays)
    p, t, _, _ = m.trace_closest(rays)
    assert np.array_equal(bp, p)
    assert np.array_equal(bt[bp != 0xFFFFFFFF], t[bp != 0xFFFFFFFF])


def test_any_hit_agrees_with_closest():
    pos, idx, rays = lcg_soup()
    m = orc.Mesh(pos, idx, bins=32, max_leaf=4, sah_leaf_mult=1)
    p, _, _, _ = m.trace_closest(rays)
    assert np.array_equal(m.trace_any(rays).astype(bool), p != 0xFFFFFFFF)


def test_any_hit_limit_single_triangle():
    # bvh.rs:455-472
    pos = np.array([[-1, -1, 5], [1, -1, 5], [0, 1, 5]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    m = orc.Mesh(pos, idx)
    assert m.trace_any(z_ray(10.0))[0] == 1
    assert m.trace_any(z_ray(4.0))[0] == 0


def test_cube_closed_forms():
    # mesh.rs:473-496: t = 4 for the unit cube at z = 5, t = 3 when scaled 2x.
    pos, idx = cube()
    p, t, _, _ = orc.Mesh(pos, idx).trace_closest(z_ray())
    assert p[0] != 0xFFFFFFFF and abs(t[0] - 4.0) < 1e-9
    pos, idx = cube(scale=2.0)
    p, t, _, _ = orc.Mesh(pos, idx).trace_closest(z_ray())
    assert abs(t[0] - 3.0) < 1e-9


def test_occlusion_respects_limit():
    # mesh.rs:519-527
    pos, idx = cube()
    m = orc.Mesh(pos, idx)
    assert m.trace_any(z_ray(10.0))[0] == 1
    assert m.trace_any(z_ray(3.5))[0] == 0


def test_barycentrics_of_known_hit():
    # Ray through (x, y) = (0.25, 0.5) on the unit right triangle (0,0)-(1,0)-(0,1) at z = 2:
    # u = 0.25, v = 0.5 exactly.
    pos = np.array([[0, 0, 2], [1, 0, 2], [0, 1, 2]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    r = np.array([(0.25, 0.5, 0, 1e-9, 0, 0, 1, np.inf)], RAY)
    p, t, u, v = orc.Mesh(pos, idx).trace_closest(r)
    assert p[0] == 0 and t[0] == 2.0 and u[0] == 0.25 and v[0] == 0.5


def test_heightfield_traversal_matches_brute_force():
    import rrhip
    pos, idx = rrhip.gen_heightfield(48, 24, 7)
    lo, hi = pos.min(0), pos.max(0)
    rays = rrhip.gen_query_rays(lo, hi, 0.1, 7, 3000)
    bp, bt = orc.brute_closest(pos, idx, rays)
    p, t, _, _ = orc.Mesh(pos, idx, bins=32, max_leaf=4, sah_leaf_mult=1).trace_closest(rays)
    assert np.array_equal(bp, p)
    assert np.array_equal(bt[bp != 0xFFFFFFFF], t[bp != 0xFFFFFFFF])
    assert 0.2 < (bp != 0xFFFFFFFF).mean() < 0.8
