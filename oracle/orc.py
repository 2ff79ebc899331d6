"""ORACLE — test infrastructure only.

ctypes loader for liboracle.so, the CPU f64 restatement of render.rs's
hot path (see orc_math.hpp). Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg may import this module; it is the checker,
never the thing measured or shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    """Compiles liboracle.so with g++ (always recompiles)."""
    r = subprocess.run(["make", "-B", "-C", HERE, "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built; run make -C oracle")
    lib = ctypes.CDLL(LIB_PATH)
    vp, u64, u32, i32, f64 = (ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                              ctypes.c_double)
    sig = {
        "orc_pcg32_seq": (None, [u64, u64, u64, vp]),
        "orc_pcg32_pixel_sample": (None, [u64, u64, u64, vp]),
        "orc_perlin": (None, [vp, u64, vp]),
        "orc_fbm": (None, [vp, u64, f64, u32, f64, f64, vp, vp]),
        "orc_mesh_create": (vp, [vp, u64, vp, u64, i32, i32, i32]),
        "orc_mesh_destroy": (None, [vp]),
        "orc_mesh_node_count": (u64, [vp]),
        "orc_mesh_trace_closest": (None, [vp, vp, u64, vp, vp, vp, vp, i32, vp]),
        "orc_mesh_trace_any": (None, [vp, vp, u64, vp, i32]),
        "orc_brute_closest": (None, [vp, vp, u64, vp, u64, vp, vp, i32]),
        "orc_edge_distance": (None, [vp, vp, vp, vp, u64, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def default_threads():
    return max(1, min(16, os.cpu_count() or 1))


def pcg32_seq(seed, stream, n):
    out = np.empty(n, np.uint32)
    load().orc_pcg32_seq(seed, stream, n, _p(out))
    return out


def pcg32_pixel_sample(pixel, sample, n):
    out = np.empty(n, np.uint32)
    load().orc_pcg32_pixel_sample(pixel, sample, n, _p(out))
    return out


def perlin(xyz):
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    out = np.empty(xyz.shape[0], np.float64)
    load().orc_perlin(_p(xyz), xyz.shape[0], _p(out))
    return out


def fbm(xyz, frequency=1.0, octaves=4, gain=0.5, lacunarity=2.0, offset=(0.0, 0.0, 0.0)):
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    off = np.ascontiguousarray(offset, np.float64)
    out = np.empty(xyz.shape[0], np.float64)
    load().orc_fbm(_p(xyz), xyz.shape[0], frequency, octaves, gain, lacunarity, _p(off), _p(out))
    return out


class Mesh:
    """f64 BVH over an f32 triangle mesh (bvh.rs + mesh.rs restated)."""

    def __init__(self, positions, indices, bins=16, max_leaf=4, sah_leaf_mult=4):
        self.pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        self.idx = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        self.h = load().orc_mesh_create(_p(self.pos), self.pos.shape[0], _p(self.idx),
                                        self.idx.shape[0], bins, max_leaf, sah_leaf_mult)

    def __del__(self):
        if getattr(self, "h", None):
            load().orc_mesh_destroy(self.h)
            self.h = None

    def node_count(self):
        return load().orc_mesh_node_count(self.h)

    def trace_closest(self, rays, threads=None, with_stats=False):
        """rays: structured RAY_DTYPE array or (n, 8) f32. Returns prim, t, u, v."""
        r = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
        n = r.shape[0]
        prim = np.empty(n, np.uint32)
        t = np.empty(n, np.float64)
        u = np.empty(n, np.float64)
        v = np.empty(n, np.float64)
        stats = np.zeros(2, np.uint64)
        load().orc_mesh_trace_closest(self.h, _p(r), n, _p(prim), _p(t), _p(u), _p(v),
                                      threads or default_threads(), _p(stats) if with_stats else None)
        if with_stats:
            return prim, t, u, v, stats
        return prim, t, u, v

    def trace_any(self, rays, threads=None):
        r = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
        occ = np.empty(r.shape[0], np.uint8)
        load().orc_mesh_trace_any(self.h, _p(r), r.shape[0], _p(occ), threads or default_threads())
        return occ


def brute_closest(positions, indices, rays, threads=None):
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
    r = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
    prim = np.empty(r.shape[0], np.uint32)
    t = np.empty(r.shape[0], np.float64)
    load().orc_brute_closest(_p(pos), _p(idx), idx.shape[0], _p(r), r.shape[0], _p(prim), _p(t),
                             threads or default_threads())
    return prim, t


def edge_distance(positions, indices, rays, prims):
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
    r = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
    pr = np.ascontiguousarray(prims, np.uint32)
    out = np.empty(r.shape[0], np.float64)
    load().orc_edge_distance(_p(pos), _p(idx), _p(r), _p(pr), r.shape[0], _p(out))
    return out
