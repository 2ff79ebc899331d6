"""GPU parity of the ray-query kernels against the f64 oracle.

Bar: bit-exact. The kernel re-tests every candidate the f32 prefilter
cannot rule out in f64 with render.rs's operation order
(src/geometry/mesh.rs:117-142), so hit ids must equal the f64 CPU result
and the returned f32 t / u / v must equal the f64 values rounded to f32.
"""
import sys

import numpy as np
import pytest

import orc
import rrhip
from test_oracle_bvh import RAY, cube, lcg_soup, z_ray

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def assert_closest_equal(hits, oprim, ot, ou=None, ov=None):
    assert np.array_equal(hits["prim"], oprim), f"{(hits['prim'] != oprim).sum()} id mismatches"
    hit = oprim != rrhip.NO_HIT
    assert np.array_equal(hits["t"][hit], ot[hit].astype(np.float32))
    assert np.isinf(hits["t"][~hit]).all()
    if ou is not None:
        assert np.array_equal(hits["u"][hit], ou[hit].astype(np.float32))
        assert np.array_equal(hits["v"][hit], ov[hit].astype(np.float32))


def test_device_is_gfx950(ctx):
    name, arch = ctx.device_name()
    say(f"device {name} {arch}")
    assert arch.startswith("gfx950")


def test_cube_closed_forms(ctx):
    say("cube closed forms")
    for scale, want in ((1.0, 4.0), (2.0, 3.0)):
        pos, idx = cube(scale=scale)
        sc = ctx.upload_mesh(pos, idx)
        h = sc.trace_closest(z_ray())
        assert h["prim"][0] != rrhip.NO_HIT and h["t"][0] == np.float32(want)
        assert sc.trace_any(z_ray(10.0))[0] == 1
        assert sc.trace_any(z_ray(want - 0.5))[0] == 0
        sc.close()


def test_single_triangle_barycentrics(ctx):
    say("single triangle")
    pos = np.array([[0, 0, 2], [1, 0, 2], [0, 1, 2]], np.float32)
    sc = ctx.upload_mesh(pos, np.array([[0, 1, 2]], np.uint32))
    r = np.array([(0.25, 0.5, 0, 1e-9, 0, 0, 1, np.inf), (0.75, 0.75, 0, 1e-9, 0, 0, 1, np.inf),
                  (0.25, 0.5, 0, 1e-9, 0, 0, 1, 2.0), (0.25, 0.5, 3, 1e-9, 0, 0, 1, np.inf)], RAY)
    h = sc.trace_closest(r)
    assert h["prim"].tolist() == [0, rrhip.NO_HIT, rrhip.NO_HIT, rrhip.NO_HIT]
    assert (h["t"][0], h["u"][0], h["v"][0]) == (2.0, 0.25, 0.5)
    assert sc.trace_any(r).tolist() == [1, 0, 0, 0]
    sc.close()


def test_empty_and_degenerate_scenes(ctx):
    say("empty / degenerate scenes")
    pos = np.array([[0, 0, 2], [1, 0, 2], [0, 1, 2]], np.float32)
    sc = ctx.upload_mesh(pos, np.zeros((0, 3), np.uint32))
    r = np.array([(0.25, 0.5, 0, 1e-9, 0, 0, 1, np.inf)] * 5, RAY)
    assert (sc.trace_closest(r)["prim"] == rrhip.NO_HIT).all()
    assert (sc.trace_any(r) == 0).all()
    sc.close()
    # 100 copies of one triangle: equal t everywhere, the lowest id wins.
    sc = ctx.upload_mesh(pos, np.tile(np.array([[0, 1, 2]], np.uint32), (100, 1)))
    h = sc.trace_closest(r)
    assert (h["prim"] == 0).all() and (h["t"] == 2.0).all()
    sc.close()


def test_lcg_soup_matches_brute_force(ctx):
    say("LCG soup vs brute force (bvh.rs:403-453 pattern)")
    pos, idx, rays = lcg_soup(ntris=500, nrays=2000)
    bp, bt = orc.brute_closest(pos, idx, rays)
    sc = ctx.upload_mesh(pos, idx)
    assert_closest_equal(sc.trace_closest(rays), bp, bt)
    assert np.array_equal(sc.trace_any(rays).astype(bool), bp != rrhip.NO_HIT)
    sc.close()


def test_lcg_soup_bounded_rays(ctx):
    say("LCG soup with finite tmax / tmin")
    pos, idx, rays = lcg_soup(ntris=500, nrays=2000)
    rng = np.random.default_rng(1)
    rays["tmax"] = rng.uniform(0.5, 6.0, rays.shape[0]).astype(np.float32)
    rays["tmin"] = rng.uniform(0.0, 0.5, rays.shape[0]).astype(np.float32)
    m = orc.Mesh(pos, idx, bins=32, max_leaf=4, sah_leaf_mult=1)
    op, ot, ou, ov = m.trace_closest(rays)
    sc = ctx.upload_mesh(pos, idx)
    assert_closest_equal(sc.trace_closest(rays), op, ot, ou, ov)
    assert np.array_equal(sc.trace_any(rays), m.trace_any(rays))
    sc.close()


@pytest.mark.parametrize("nx,ny,nrays,seed", [(16, 8, 5000, 3), (128, 64, 50000, 4)])
def test_small_heightfields(ctx, nx, ny, nrays, seed):
    say(f"heightfield {nx}x{ny} rays={nrays} seed={seed}")
    pos, idx = rrhip.gen_heightfield(nx, ny, seed)
    rays = rrhip.gen_query_rays(pos.min(0), pos.max(0), 0.1, seed, nrays)
    m = orc.Mesh(pos, idx, bins=32, max_leaf=4, sah_leaf_mult=1)
    op, ot, ou, ov = m.trace_closest(rays)
    sc = ctx.upload_mesh(pos, idx)
    assert_closest_equal(sc.trace_closest(rays), op, ot, ou, ov)
    assert np.array_equal(sc.trace_any(rays), m.trace_any(rays))
    sc.close()


def test_device_resident_entry_points(ctx):
    say("device-resident entry points")
    pos, idx = rrhip.gen_heightfield(64, 32, 9)
    rays = rrhip.gen_query_rays(pos.min(0), pos.max(0), 0.1, 9, 30000)
    sc = ctx.upload_mesh(pos, idx)
    host = sc.trace_closest(rays)
    d_rays = ctx.to_device(rays)
    d_hits = ctx.alloc(rays.shape[0] * 16)
    d_occ = ctx.alloc(rays.shape[0])
    sc.trace_closest_device(d_rays, rays.shape[0], d_hits)
    sc.trace_any_device(d_rays, rays.shape[0], d_occ)
    ctx.synchronize()
    dev = ctx.to_host(d_hits, np.empty(rays.shape[0], rrhip.HIT_DTYPE))
    occ = ctx.to_host(d_occ, np.empty(rays.shape[0], np.uint8))
    assert dev.tobytes() == host.tobytes()
    assert np.array_equal(occ.astype(bool), host["prim"] != rrhip.NO_HIT)
    ms = sc.time_trace_device(d_rays, rays.shape[0], d_hits, 3)
    assert ms > 0
    again = ctx.to_host(d_hits, np.empty(rays.shape[0], rrhip.HIT_DTYPE))
    assert again.tobytes() == host.tobytes()
    for p in (d_rays, d_hits, d_occ):
        ctx.free(p)
    sc.close()


def test_trace_cost_counts(ctx):
    say("trace cost counters")
    pos, idx = rrhip.gen_heightfield(64, 32, 2)
    rays = rrhip.gen_query_rays(pos.min(0), pos.max(0), 0.1, 2, 10000)
    sc = ctx.upload_mesh(pos, idx)
    nodes, tris = sc.trace_cost(rays)
    anodes, atris = sc.trace_cost(rays, any_hit=True)
    assert nodes >= rays.shape[0] and tris > 0
    assert anodes <= nodes * 1.5  # any-hit stops at the first hit
    # Counting must not change results.
    assert sc.trace_closest(rays).tobytes() == sc.trace_closest(rays).tobytes()
    sc.close()


def test_query_benchmark_full_size(ctx):
    """The 1,048,576-triangle heightfield with 4,194,304 rays: every hit id,
    t, u, v and any-hit flag against the f64 oracle."""
    say("full-size ray query: 1024x512 heightfield, 4194304 rays, seed 1")
    pos, idx = rrhip.gen_heightfield(1024, 512, 1)
    rays = rrhip.gen_query_rays(pos.min(0), pos.max(0), 0.1, 1, 4194304)
    sc = ctx.upload_mesh(pos, idx)
    hits = sc.trace_closest(rays)
    occ = sc.trace_any(rays)
    say("gpu done; oracle build + trace")
    m = orc.Mesh(pos, idx, bins=32, max_leaf=4, sah_leaf_mult=1)
    op, ot, ou, ov = m.trace_closest(rays)
    say(f"hit fraction {np.mean(op != rrhip.NO_HIT):.4f}")
    assert_closest_equal(hits, op, ot, ou, ov)
    assert np.array_equal(occ, m.trace_any(rays))
    # size-independent: any-hit == (closest hit exists)
    assert np.array_equal(occ.astype(bool), hits["prim"] != rrhip.NO_HIT)
    sc.close()


def test_every_launch_configuration_agrees(ctx):
    say("launch configurations: width x block x stack x rays_per_wave")
    pos, idx = rrhip.gen_heightfield(256, 128, 5)
    rays = rrhip.gen_query_rays(pos.min(0), pos.max(0), 0.1, 5, 40000)
    m = orc.Mesh(pos, idx, bins=32, max_leaf=4, sah_leaf_mult=1)
    op, ot, ou, ov = m.trace_closest(rays)
    oany = m.trace_any(rays)
    sc = ctx.upload_mesh(pos, idx)
    info = sc.info()
    depth, depth4 = info["depth"], info["depth4"]
    assert 0 < depth4 < depth
    for width, need in ((2, depth), (4, depth4)):
        stacks = [s for s in (16, 24, 32, 48, 64) if s >= need]
        for block in (64, 128, 256):
            for stack in stacks[:2]:
                for rpw in (0, 64, 192, 1024):  # 0: one round of resident waves
                    sc.set_tuning(block, stack, rpw, width)
                    assert sc.get_tuning() == {"block": block, "stack": stack, "rays_per_wave": rpw, "width": width}
                    assert_closest_equal(sc.trace_closest(rays), op, ot, ou, ov)
                    assert np.array_equal(sc.trace_any(rays), oany)
    with pytest.raises(rrhip.RrhipError):
        sc.set_tuning(128, 16 if depth > 16 else 8, 256)
    with pytest.raises(rrhip.RrhipError):
        sc.set_tuning(128, 64, 256, 3)
    with pytest.raises(rrhip.RrhipError):
        sc.set_tuning(96, stacks[0], 256)
    with pytest.raises(rrhip.RrhipError):
        sc.set_tuning(128, stacks[0], 100)
    d_rays = ctx.to_device(rays)
    d_hits = ctx.alloc(rays.shape[0] * 16)
    best, ms = sc.autotune(d_rays, rays.shape[0], d_hits)
    say(f"autotune chose {best} at {ms:.3f} ms")
    assert sc.get_tuning() == best and ms > 0
    assert_closest_equal(sc.trace_closest(rays), op, ot, ou, ov)
    # any hit is tuned apart; the closest-hit configuration stays
    d_occ = ctx.alloc(rays.shape[0])
    best_any, ms_any = sc.autotune(d_rays, rays.shape[0], d_occ, any_hit=True)
    say(f"any-hit autotune chose {best_any} at {ms_any:.3f} ms")
    assert sc.get_tuning() == best and ms_any > 0
    assert np.array_equal(sc.trace_any(rays), oany)
    assert_closest_equal(sc.trace_closest(rays), op, ot, ou, ov)
    ctx.free(d_occ)
    ctx.free(d_rays)
    ctx.free(d_hits)
    sc.close()
