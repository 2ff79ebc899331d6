"""The C-ABI library: it loads, exports every symbol include/rrhip.h
declares, and its host-only entry points (generators, BVH build, errors)
behave. No GPU compute here."""
import ctypes
import os
import re

import numpy as np
import pytest

import orc
import rrhip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rrhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rrhip_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(rrhip.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in rrhip.h but not exported"
    assert sorted(rrhip.EXPORTED) == syms


def test_abi_version_and_error_string():
    lib = rrhip.load()
    assert lib.rrhip_abi_version() == 10
    h = ctypes.c_void_p()
    st = lib.rrhip_init(-1, ctypes.byref(h))
    assert st != 0 and not h
    assert len(lib.rrhip_last_error()) > 0


def test_init_without_device_is_a_clean_error():
    lib = rrhip.load()
    h = ctypes.c_void_p()
    st = lib.rrhip_init(0, ctypes.byref(h))
    if st == 0:  # a GPU is visible (GPU box): the context must be usable
        assert lib.rrhip_destroy(h) == 0
    else:
        assert st in (2, 5), lib.rrhip_last_error()


def test_null_arguments_rejected():
    lib = rrhip.load()
    assert lib.rrhip_scene_upload(None, None, None) == 1
    assert lib.rrhip_trace_closest(None, None, 0, None) == 1
    assert lib.rrhip_build_info(None, None) == 1


def test_heightfield_generator_matches_oracle_fbm():
    pos, idx = rrhip.gen_heightfield(64, 32, 11, amplitude=0.1, frequency=4.0, octaves=6)
    assert pos.shape == (65 * 33, 3) and idx.shape == (64 * 32 * 2, 3)
    # Offset drawn from PCG32(seed, 0x5eed) as 1000 * u.
    u = orc.pcg32_seq(11, 0x5EED, 3).astype(np.float64) / 2.0**32
    off = 1000.0 * u
    xs = np.arange(65) / 64.0
    ys = np.arange(33) / 32.0
    X, Y = np.meshgrid(xs, ys)
    pts = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    amp = float(np.float32(0.1))  # the ABI takes the amplitude as f32
    z = amp * orc.fbm(pts, frequency=4.0, octaves=6, gain=0.5, lacunarity=2.0, offset=off)
    assert np.array_equal(pos[:, 2], z.astype(np.float32))
    assert np.array_equal(pos[:, 0], X.ravel().astype(np.float32))
    # Quad (i, j) -> triangles (v00, v10, v11), (v00, v11, v01).
    assert idx[0].tolist() == [0, 1, 66] and idx[1].tolist() == [0, 66, 65]


def test_query_rays_deterministic_and_on_sphere():
    lo, hi = np.array([0, 0, -0.1], np.float32), np.array([1, 1, 0.1], np.float32)
    a = rrhip.gen_query_rays(lo, hi, 0.1, 5, 5000)
    b = rrhip.gen_query_rays(lo, hi, 0.1, 5, 5000)
    assert a.tobytes() == b.tobytes()
    d = np.stack([a["dx"], a["dy"], a["dz"]], 1).astype(np.float64)
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-6
    assert abs(a["dz"].mean()) < 0.05
    ox = a["ox"]
    assert ox.min() >= -0.05 - 1e-6 and ox.max() <= 1.05 + 1e-6
    assert (a["tmin"] == np.float32(1e-9)).all() and np.isinf(a["tmax"]).all()
    # Ray i uses PCG32::for_pixel_sample(i, seed): its first draw sets ox.
    u0 = orc.pcg32_pixel_sample(17, 5, 1)[0] / 2.0**32
    assert a["ox"][17] == np.float32(0.5 + 0.55 * (2 * u0 - 1))


def test_build_info_reports_a_valid_tree():
    pos, idx = rrhip.gen_heightfield(128, 64, 1)
    info = rrhip.build_info(pos, idx)
    assert info["num_triangles"] == idx.shape[0]
    # leaves <= 4 triangles each, child-pair nodes = leaves - 1
    assert info["num_leaves"] >= idx.shape[0] / 4
    assert info["num_nodes"] == info["num_leaves"] - 1
    assert info["device_bytes"] == 64 * info["num_nodes"] + 48 * idx.shape[0]
    assert 1 <= info["depth"] < 48
    assert np.allclose(info["bounds_lo"], pos.min(0)) and np.allclose(info["bounds_hi"], pos.max(0))


def test_build_rejects_bad_meshes():
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    with pytest.raises(rrhip.RrhipError):
        rrhip.build_info(pos, np.array([[0, 1, 3]], np.uint32))  # index out of range
    bad = pos.copy()
    bad[1, 0] = np.nan
    with pytest.raises(rrhip.RrhipError):
        rrhip.build_info(bad, np.array([[0, 1, 2]], np.uint32))
    with pytest.raises(rrhip.RrhipError):
        rrhip.build_info(pos, np.array([[0, 1, 2]], np.uint32), bins=1)


def test_build_edge_cases():
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    one = rrhip.build_info(pos, np.array([[0, 1, 2]], np.uint32))
    assert one["num_nodes"] == 1 and one["depth"] == 1
    empty = rrhip.build_info(pos, np.zeros((0, 3), np.uint32))
    assert empty["num_nodes"] == 0
    # All centroids coincide: the builder must still terminate with small leaves.
    many = np.tile(np.array([[0, 1, 2]], np.uint32), (100, 1))
    info = rrhip.build_info(pos, many)
    assert info["num_leaves"] >= 25
