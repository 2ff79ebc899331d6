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


# ---- path tracing ------------------------------------------------------------

def _pt_sigs():
    lib = load()
    if getattr(lib, "_pt_ready", False):
        return lib
    vp, u64, u32, i32, f64 = (ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                              ctypes.c_double)
    sig = {
        "orc_render": (i32, [vp, vp, u32, u32, vp, i32, vp]),
        "orc_render_pixels": (i32, [vp, vp, vp, u64, u32, u32, vp, i32, vp]),
        "orc_render_adaptive": (i32, [vp, vp, u32, f64, vp, i32]),
        "orc_render_film": (i32, [vp, vp, u32, u32, vp, i32]),
        "orc_env_pdf_eval_n": (None, [u32, u32, vp, vp, u64, vp, vp]),
        "orc_env_sample_n": (None, [u32, u32, vp, vp, u64, vp, vp, vp, ctypes.POINTER(f64)]),
        "orc_camera_rays": (None, [vp, u64, vp, vp, vp, vp]),
        "orc_bsdf_eval": (None, [vp, vp, vp, f64, vp, vp]),
        "orc_bsdf_eval_n": (None, [vp, vp, vp, f64, u64, vp, vp]),
        "orc_bsdf_sample_n": (None, [vp, vp, f64, u64, u64, u64, vp, vp, vp, vp]),
        "orc_light_sampler_create": (vp, [vp, u32]),
        "orc_light_sampler_destroy": (None, [vp]),
        "orc_light_sampler_sample": (i32, [vp, vp, f64, ctypes.POINTER(u32), ctypes.POINTER(f64)]),
        "orc_light_sampler_pmf": (f64, [vp, vp, u32]),
        "orc_scene_intersect": (i32, [vp, vp, vp, ctypes.POINTER(f64), vp, vp, ctypes.POINTER(i32),
                                      ctypes.POINTER(u32)]),
        "orc_scene_intersect_n": (None, [vp, vp, u64, vp, vp, vp, vp, vp]),
        "orc_scene_st_n": (None, [vp, vp, u64, vp, vp, vp, vp]),
        "orc_texture_level": (u64, [u32, u32, vp, u32, vp]),
        "orc_texture_sample": (None, [u32, u32, vp, u32, vp, vp, vp, vp, vp, u64, vp]),
        "orc_pattern_eval": (None, [u32, vp, u32, vp, i32, u64, vp, vp, vp, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._pt_ready = True
    return lib


def render(desc, params, start, count, accum=None, threads=None):
    """Adds samples [start, start+count) of every pixel into accum
    (npix x 8 doubles, the RRHIP_ACCUM layout). Returns (accum, rays)."""
    lib = _pt_sigs()
    npix = desc.camera.width * desc.camera.height
    if accum is None:
        accum = np.zeros((npix, 8), np.float64)
    rays = ctypes.c_uint64()
    lib.orc_render(ctypes.addressof(desc), ctypes.addressof(params), start, count, _p(accum),
                   threads or default_threads(), ctypes.byref(rays))
    return accum, rays.value


def render_pixels(desc, params, pixels, start, count, threads=None):
    """render over the listed pixels only (full-image indices y * width +
    x, so each keeps its whole-image seeds): (len(pixels) x 8 doubles,
    rays)."""
    lib = _pt_sigs()
    px = np.ascontiguousarray(pixels, np.uint32).ravel()
    accum = np.zeros((px.size, 8), np.float64)
    rays = ctypes.c_uint64()
    lib.orc_render_pixels(ctypes.addressof(desc), ctypes.addressof(params), _p(px), px.size, start, count,
                          _p(accum), threads or default_threads(), ctypes.byref(rays))
    return accum, rays.value


def render_film(desc, params, start, count, threads=None):
    """pt::render_film over samples [start, start+count): (npix, 18) f64 in
    the RRHIP_FILM layout (beauty, diffuse, specular, albedo, normal rgb /
    xyz, depth, dominant id, its coverage)."""
    lib = _pt_sigs()
    npix = desc.camera.width * desc.camera.height
    film = np.zeros((npix, 18), np.float64)
    lib.orc_render_film(ctypes.addressof(desc), ctypes.addressof(params), start, count, _p(film),
                        threads or default_threads())
    return film


def env_pdf_eval(rgb, dirs):
    """EnvMap::pdf and ::eval (envmap.rs:84-148) of the (h, w, 3) map at
    directions dirs (n x 3): (pdf (n,), radiance (n, 3))."""
    lib = _pt_sigs()
    rgb = np.ascontiguousarray(rgb, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    n = dirs.shape[0]
    pdf, rad = np.zeros(n), np.zeros((n, 3))
    lib.orc_env_pdf_eval_n(rgb.shape[1], rgb.shape[0], _p(rgb), _p(dirs), n, _p(pdf), _p(rad))
    return pdf, rad


def env_sample(rgb, u):
    """EnvMap::sample (envmap.rs:99-131) for uniform pairs u (n x 2):
    (dirs (n, 3), radiance (n, 3), pdf (n,), mean luminance)."""
    lib = _pt_sigs()
    rgb = np.ascontiguousarray(rgb, np.float32)
    u = np.ascontiguousarray(u, np.float64).reshape(-1, 2)
    n = u.shape[0]
    dirs, rad, pdf = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    ml = ctypes.c_double()
    lib.orc_env_sample_n(rgb.shape[1], rgb.shape[0], _p(rgb), _p(u), n, _p(dirs), _p(rad), _p(pdf),
                         ctypes.byref(ml))
    return dirs, rad, pdf, ml.value


def camera_rays(desc, xy, samples):
    """Camera rays (origins, unit directions) of pixels xy (n x 2) at the
    given sample indices."""
    lib = _pt_sigs()
    xy = np.ascontiguousarray(xy, np.uint32).reshape(-1, 2)
    samples = np.ascontiguousarray(samples, np.uint64)
    n = xy.shape[0]
    o = np.zeros((n, 3))
    d = np.zeros((n, 3))
    lib.orc_camera_rays(ctypes.addressof(desc), n, _p(xy), _p(samples), _p(o), _p(d))
    return o, d


def render_adaptive(desc, params, max_spp, tol, threads=None):
    """pt::render_adaptive (pt/mod.rs:285-336) into a fresh accumulator
    (npix x 8 doubles; column 6 is each pixel's sample count)."""
    lib = _pt_sigs()
    npix = desc.camera.width * desc.camera.height
    accum = np.zeros((npix, 8), np.float64)
    lib.orc_render_adaptive(ctypes.addressof(desc), ctypes.addressof(params), max_spp, float(tol), _p(accum),
                            threads or default_threads())
    return accum


def bsdf_eval(mat, wo, wi, eta):
    lib = _pt_sigs()
    wo = np.ascontiguousarray(wo, np.float64)
    wi = np.ascontiguousarray(wi, np.float64)
    f = np.zeros(3)
    pdf = np.zeros(1)
    lib.orc_bsdf_eval(ctypes.addressof(mat), _p(wo), _p(wi), eta, _p(f), _p(pdf))
    return f, pdf[0]


def bsdf_eval_n(mat, wo, wi, eta):
    lib = _pt_sigs()
    wo = np.ascontiguousarray(wo, np.float64)
    wi = np.ascontiguousarray(wi, np.float64).reshape(-1, 3)
    f = np.zeros_like(wi)
    pdf = np.zeros(wi.shape[0])
    lib.orc_bsdf_eval_n(ctypes.addressof(mat), _p(wo), _p(wi), eta, wi.shape[0], _p(f), _p(pdf))
    return f, pdf


def bsdf_sample_n(mat, wo, eta, seed, stream, n):
    lib = _pt_sigs()
    wo = np.ascontiguousarray(wo, np.float64)
    wi = np.zeros((n, 3))
    f = np.zeros((n, 3))
    pdf = np.zeros(n)
    ok = np.zeros(n, np.uint8)
    lib.orc_bsdf_sample_n(ctypes.addressof(mat), _p(wo), eta, seed, stream, n, _p(wi), _p(f), _p(pdf), _p(ok))
    return wi, f, pdf, ok.astype(bool)


class LightSampler:
    def __init__(self, lights):
        lib = _pt_sigs()
        self._arr = (type(lights[0]) * len(lights))(*lights)
        self.h = lib.orc_light_sampler_create(ctypes.addressof(self._arr), len(lights))

    def __del__(self):
        if getattr(self, "h", None):
            _pt_sigs().orc_light_sampler_destroy(self.h)
            self.h = None

    def sample(self, p, u):
        lib = _pt_sigs()
        pp = np.ascontiguousarray(p, np.float64)
        li = ctypes.c_uint32()
        pmf = ctypes.c_double()
        if not lib.orc_light_sampler_sample(self.h, _p(pp), u, ctypes.byref(li), ctypes.byref(pmf)):
            return None
        return li.value, pmf.value

    def pmf(self, p, light):
        pp = np.ascontiguousarray(p, np.float64)
        return _pt_sigs().orc_light_sampler_pmf(self.h, _p(pp), light)


def scene_intersect(desc, o, d):
    lib = _pt_sigs()
    o = np.ascontiguousarray(o, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    t = ctypes.c_double()
    point = np.zeros(3)
    normal = np.zeros(3)
    front = ctypes.c_int()
    mat = ctypes.c_uint32()
    if not lib.orc_scene_intersect(ctypes.addressof(desc), _p(o), _p(d), ctypes.byref(t), _p(point),
                                   _p(normal), ctypes.byref(front), ctypes.byref(mat)):
        return None
    return t.value, point, normal, bool(front.value), mat.value


def scene_intersect_n(desc, rays):
    """Closest hits of rays (n x 6: origin, direction) in one scene:
    (t (-1 on a miss), normal, tangent, front, material)."""
    lib = _pt_sigs()
    r = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
    n = r.shape[0]
    t, nrm, tan = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 3))
    front, mat = np.zeros(n, np.int32), np.zeros(n, np.uint32)
    lib.orc_scene_intersect_n(ctypes.addressof(desc), _p(r), n, _p(t), _p(nrm), _p(tan), _p(front), _p(mat))
    return t, nrm, tan, front.astype(bool), mat


def scene_st_n(desc, rays):
    """Texture coordinates at the closest hits of rays (n x 6): (t (-1 on
    a miss), st (n x 2), st density (n), gradients of s and t (n x 2 x 3))."""
    lib = _pt_sigs()
    r = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
    n = r.shape[0]
    t, st, dens, grad = np.zeros(n), np.zeros((n, 2)), np.zeros(n), np.zeros((n, 2, 3))
    lib.orc_scene_st_n(ctypes.addressof(desc), _p(r), n, _p(t), _p(st), _p(dens), _p(grad))
    return t, st, dens, grad


# ---- textures and patterns ---------------------------------------------------

def texture_chain(rgb):
    """Every mip level of a texture (tex.rs's 2x2 box chain) as h x w x 3."""
    lib = _pt_sigs()
    a = np.ascontiguousarray(rgb, np.float32)
    h, w = a.shape[:2]
    out, lvl = [], 0
    while True:
        wh = lib.orc_texture_level(w, h, _p(a), lvl, None)
        if wh == 0:
            return out
        lw, lh = wh >> 32, wh & 0xFFFFFFFF
        b = np.zeros((lh, lw, 3), np.float32)
        lib.orc_texture_level(w, h, _p(a), lvl, _p(b))
        out.append(b)
        lvl += 1


def texture_sample(rgb, s, t, footprint=None, du=None, dv=None, wrap=0):
    """cache.rs lookups: trilinear from the footprint, or EWA from the
    texture-space axes du, dv (n x 2 each)."""
    lib = _pt_sigs()
    a = np.ascontiguousarray(rgb, np.float32)
    h, w = a.shape[:2]
    s = np.ascontiguousarray(np.atleast_1d(s), np.float64)
    n = s.size
    t = np.ascontiguousarray(np.broadcast_to(t, (n,)), np.float64)
    fp = np.ascontiguousarray(np.broadcast_to(0.0 if footprint is None else footprint, (n,)), np.float64)
    out = np.zeros((n, 3), np.float32)
    if du is not None:
        du = np.ascontiguousarray(np.broadcast_to(du, (n, 2)), np.float64)
        dv = np.ascontiguousarray(np.broadcast_to(dv, (n, 2)), np.float64)
    lib.orc_texture_sample(w, h, _p(a), wrap, _p(s), _p(t), _p(fp), None if du is None else _p(du),
                           None if dv is None else _p(dv), n, _p(out))
    return out


def pattern_eval(scene, index, st, p=(0, 0, 0), n=(0, 0, 1), footprint=0.0, dst_major=(0, 0), dst_minor=(0, 0)):
    """pattern::eval of node `index` of a scenes.Scene's graph at shading
    points (st n x 2; the others broadcast): n x 3 rgb."""
    lib = _pt_sigs()
    st = np.ascontiguousarray(np.atleast_2d(st), np.float64)
    m = st.shape[0]
    b = lambda v, k: np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (m, k) if k > 1 else (m,)),
                                          np.float64)
    P, N, F, A, B = b(p, 3), b(n, 3), b(footprint, 1), b(dst_major, 2), b(dst_minor, 2)
    d = scene.desc()
    out = np.zeros((m, 3))
    lib.orc_pattern_eval(d.num_textures, ctypes.cast(d.textures, ctypes.c_void_p), d.num_patterns,
                         ctypes.cast(d.patterns, ctypes.c_void_p), index, m, _p(st), _p(P), _p(N), _p(F), _p(A),
                         _p(B), _p(out))
    return out


# ---- neural radiance cache ---------------------------------------------------

MLP_PARAMS = 19648
N_IN = 48


def _nrc_sigs():
    lib = load()
    if getattr(lib, "_nrc_ready", False):
        return lib
    vp, u64, u32, i32, f64 = (ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                              ctypes.c_double)
    sig = {
        "orc_nrc_level_res": (None, [vp, vp]),
        "orc_nrc_corner_index": (u32, [vp, i32, u32, u32, u32]),
        "orc_nrc_inputs": (None, [vp, vp, vp, u64, vp]),
        "orc_nrc_forward": (None, [vp, vp, u64, i32, vp]),
        "orc_nrc_infer": (None, [vp, vp, vp, vp, u64, i32, vp]),
        "orc_nrc_gradients": (f64, [vp, vp, vp, vp, vp, u64, i32, vp, vp]),
        "orc_nrc_mlp_gradients": (f64, [vp, vp, vp, u64, i32, vp, vp]),
        "orc_nrc_adam": (None, [vp, vp, vp, vp, u64, u64, f64, f64, f64, f64]),
        "orc_bf16_round": (f64, [f64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._nrc_ready = True
    return lib


class NrcConfig:
    """Hash-grid geometry: levels, log2 entries per level, coarsest and
    finest resolution (cells per axis), position bounds."""

    def __init__(self, levels=16, log2_table=19, base_res=16.0, max_res=2048.0, lo=(0, 0, 0), hi=(1, 1, 1)):
        self.levels, self.log2_table = levels, log2_table
        self.base_res, self.max_res = base_res, max_res
        self.lo, self.hi = tuple(lo), tuple(hi)

    def arr(self):
        return np.array([self.levels, self.log2_table, self.base_res, self.max_res, *self.lo, *self.hi],
                        np.float64)

    @property
    def table_floats(self):
        return self.levels * (1 << self.log2_table) * 2


def nrc_level_res(cfg):
    out = np.zeros(cfg.levels, np.int32)
    _nrc_sigs().orc_nrc_level_res(_p(cfg.arr()), _p(out))
    return out


def nrc_corner_index(cfg, level, x, y, z):
    return _nrc_sigs().orc_nrc_corner_index(_p(cfg.arr()), level, x, y, z)


def _q(queries):
    q = np.ascontiguousarray(queries, np.float32).reshape(-1, 16)
    return q


def nrc_inputs(cfg, table, queries):
    q = _q(queries)
    t = np.ascontiguousarray(table, np.float32)
    out = np.zeros((q.shape[0], N_IN))
    _nrc_sigs().orc_nrc_inputs(_p(cfg.arr()), _p(t), _p(q), q.shape[0], _p(out))
    return out


def nrc_forward(w, x, mimic=False):
    w = np.ascontiguousarray(w, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, N_IN)
    out = np.zeros((x.shape[0], 3))
    _nrc_sigs().orc_nrc_forward(_p(w), _p(x), x.shape[0], int(mimic), _p(out))
    return out


def nrc_infer(cfg, table, w, queries, mimic=False):
    q = _q(queries)
    t = np.ascontiguousarray(table, np.float32)
    w = np.ascontiguousarray(w, np.float64)
    out = np.zeros((q.shape[0], 3))
    _nrc_sigs().orc_nrc_infer(_p(cfg.arr()), _p(t), _p(w), _p(q), q.shape[0], int(mimic), _p(out))
    return out


def nrc_gradients(cfg, table, w, queries, targets, mimic=False, want_table=True):
    """Returns (loss, grad_w, grad_table or None)."""
    q = _q(queries)
    t = np.ascontiguousarray(table, np.float32)
    w = np.ascontiguousarray(w, np.float64)
    tg = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
    gw = np.zeros(MLP_PARAMS)
    gt = np.zeros(cfg.table_floats) if want_table else None
    loss = _nrc_sigs().orc_nrc_gradients(_p(cfg.arr()), _p(t), _p(w), _p(q), _p(tg), q.shape[0], int(mimic),
                                         _p(gw), _p(gt) if want_table else None)
    return loss, gw, gt


def nrc_mlp_gradients(w, x, targets, mimic=False):
    """MLP alone on raw inputs: returns (loss, grad_w, grad_x)."""
    w = np.ascontiguousarray(w, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, N_IN)
    tg = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
    gw = np.zeros(MLP_PARAMS)
    gx = np.zeros_like(x)
    loss = _nrc_sigs().orc_nrc_mlp_gradients(_p(w), _p(x), _p(tg), x.shape[0], int(mimic), _p(gw), _p(gx))
    return loss, gw, gx


def nrc_adam(p, g, m, v, t, lr=1e-2, b1=0.9, b2=0.99, eps=1e-15):
    """In place on f64 arrays p, m, v."""
    g = np.ascontiguousarray(g, np.float64)
    _nrc_sigs().orc_nrc_adam(_p(p), _p(g), _p(m), _p(v), p.size, t, lr, b1, b2, eps)


def bf16_round(x):
    return _nrc_sigs().orc_bf16_round(float(x))


# ---- participating media -------------------------------------------------------

def _med_sigs():
    lib = load()
    if getattr(lib, "_med_ready", False):
        return lib
    vp, u64, f64 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double
    sig = {
        "orc_medium_density": (None, [vp, vp, u64, vp]),
        "orc_medium_sample_n": (None, [vp, vp, vp, f64, vp, u64, u64, u64, vp, vp, vp]),
        "orc_medium_transmittance_n": (None, [vp, vp, vp, f64, u64, u64, u64, vp]),
        "orc_hg_phase": (f64, [f64, f64]),
        "orc_equiangular_n": (None, [vp, vp, f64, vp, vp, u64, vp, vp, vp]),
        "orc_distance_pdf_approx": (f64, [vp, f64, f64]),
        "orc_hg_sample_n": (None, [f64, vp, u64, u64, u64, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._med_ready = True
    return lib


def medium_density(medium, pts):
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    out = np.zeros(pts.shape[0])
    _med_sigs().orc_medium_density(ctypes.addressof(medium), _p(pts), pts.shape[0], _p(out))
    return out


def medium_sample(medium, o, d, t_max, beta, seed, stream, n):
    """Returns (scatter bool[n], t[n], weight[n, 3])."""
    o = np.ascontiguousarray(o, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    b = np.ascontiguousarray(beta, np.float64)
    sc = np.zeros(n, np.uint8)
    t = np.zeros(n)
    w = np.zeros((n, 3))
    _med_sigs().orc_medium_sample_n(ctypes.addressof(medium), _p(o), _p(d), t_max, _p(b), seed, stream, n,
                                    _p(sc), _p(t), _p(w))
    return sc.astype(bool), t, w


def medium_transmittance(medium, o, d, dist, seed, stream, n):
    o = np.ascontiguousarray(o, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    out = np.zeros((n, 3))
    _med_sigs().orc_medium_transmittance_n(ctypes.addressof(medium), _p(o), _p(d), dist, seed, stream, n, _p(out))
    return out


def hg_phase(g, cos_theta):
    return _med_sigs().orc_hg_phase(g, cos_theta)


def equiangular(o, d, t_max, light, u):
    """volume.rs:195-232 for each u: (t, returned pdf, equiangular_pdf(t))."""
    lib = _med_sigs()
    o, d, light = (np.ascontiguousarray(x, np.float64) for x in (o, d, light))
    u = np.ascontiguousarray(u, np.float64)
    t, pdf, pdf_at = np.zeros_like(u), np.zeros_like(u), np.zeros_like(u)
    lib.orc_equiangular_n(_p(o), _p(d), float(t_max), _p(light), _p(u), u.size, _p(t), _p(pdf), _p(pdf_at))
    return t, pdf, pdf_at


def distance_pdf_approx(medium, t, t_max):
    return _med_sigs().orc_distance_pdf_approx(ctypes.addressof(medium), float(t), float(t_max))


def hg_sample(g, wo, seed, stream, n):
    wo = np.ascontiguousarray(wo, np.float64)
    wi = np.zeros((n, 3))
    pdf = np.zeros(n)
    _med_sigs().orc_hg_sample_n(g, _p(wo), seed, stream, n, _p(wi), _p(pdf))
    return wi, pdf


# ---- curves and hair (curves.rs, hair.rs) ---------------------------------------------

def _hair_sigs():
    lib = load()
    if getattr(lib, "_hair_ready", False):
        return lib
    vp, u64, i32, f64 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    sig = {
        "orc_curve_intersect": (None, [vp, vp, ctypes.c_float, ctypes.c_float, vp, u64, f64, vp, vp, vp, vp]),
        "orc_dice_curve": (u64, [vp, u64, vp, u64, f64, f64, u64, vp, vp, vp, vp, u64]),
        "orc_hair_params": (None, [vp, vp, ctypes.POINTER(f64)]),
        "orc_hair_sigma_a_from_color": (None, [vp, f64, vp]),
        "orc_hair_scalar": (None, [i32, vp, u64, f64, vp]),
        "orc_hair_mp": (None, [vp, u64, vp]),
        "orc_hair_ap": (None, [f64, f64, f64, vp, vp]),
        "orc_hair_fresnel": (f64, [f64, f64, f64]),
        "orc_hair_eval": (None, [vp, vp, vp, vp, u64, vp, vp]),
        "orc_hair_sample": (None, [vp, vp, vp, vp, u64, vp, vp, vp, vp]),
        "orc_fiber_frame": (None, [vp, vp, vp, vp, ctypes.POINTER(f64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._hair_ready = True
    return lib


def curve_intersect(p0, p1, rays, v0=0.0, v1=1.0, t_max=np.inf):
    """One rounded-cone segment (p0, p1: xyz + radius) against rays
    (n x 6: origin, direction). Returns (t (-1 on a miss), normal, tangent, v)."""
    a = np.ascontiguousarray(p0, np.float32)
    b = np.ascontiguousarray(p1, np.float32)
    r = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
    n = r.shape[0]
    t, nrm, tan, v = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    _hair_sigs().orc_curve_intersect(_p(a), _p(b), v0, v1, _p(r), n, float(t_max), _p(t), _p(nrm), _p(tan), _p(v))
    return t, nrm, tan, v


def dice_curve(ctrl, basis=None, step=3, width_root=0.01, width_tip=0.01, segs=8):
    """curves.rs:191-258. Returns (p0, p1 (n x 4 float32), v0, v1)."""
    c = np.ascontiguousarray(ctrl, np.float64).reshape(-1, 3)
    bas = None if basis is None else np.ascontiguousarray(basis, np.float64).reshape(16)
    cap = max(1, c.shape[0] * (segs + 1))
    p0, p1 = np.zeros((cap, 4), np.float32), np.zeros((cap, 4), np.float32)
    v0, v1 = np.zeros(cap, np.float32), np.zeros(cap, np.float32)
    m = _hair_sigs().orc_dice_curve(_p(c), c.shape[0], None if bas is None else _p(bas), step, width_root,
                                    width_tip, segs, _p(p0), _p(p1), _p(v0), _p(v1), cap)
    return p0[:m], p1[:m], v0[:m], v1[:m]


def _hp(sigma_a, beta_m, beta_n, eta):
    return np.array([*sigma_a, beta_m, beta_n, eta], np.float64)


def hair_params(sigma_a=(0, 0, 0), beta_m=0.3, beta_n=0.3, eta=1.55):
    """(v[4], s) of HairParams::new."""
    v = np.zeros(4)
    s = ctypes.c_double()
    _hair_sigs().orc_hair_params(_p(_hp(sigma_a, beta_m, beta_n, eta)), _p(v), ctypes.byref(s))
    return v, s.value


def hair_sigma_a_from_color(c, beta_n):
    out = np.zeros(3)
    _hair_sigs().orc_hair_sigma_a_from_color(_p(np.ascontiguousarray(c, np.float64)), beta_n, _p(out))
    return out


def hair_scalar(which, x, a=0.0):
    """which: 'i0', 'log_i0', 'logistic', 'logistic_cdf', 'trimmed_logistic',
    'sample_trimmed_logistic' (a = the logistic scale)."""
    k = ["i0", "log_i0", "logistic", "logistic_cdf", "trimmed_logistic", "sample_trimmed_logistic"].index(which)
    x = np.ascontiguousarray(x, np.float64).ravel()
    out = np.zeros_like(x)
    _hair_sigs().orc_hair_scalar(k, _p(x), x.size, float(a), _p(out))
    return out


def hair_mp(cos_ti, cos_to, sin_ti, sin_to, v):
    rows = np.ascontiguousarray(np.broadcast_arrays(cos_ti, cos_to, sin_ti, sin_to, v), np.float64)
    rows = np.ascontiguousarray(rows.reshape(5, -1).T)
    out = np.zeros(rows.shape[0])
    _hair_sigs().orc_hair_mp(_p(rows), rows.shape[0], _p(out))
    return out


def hair_ap(cos_to, eta, h, t):
    out = np.zeros((4, 3))
    _hair_sigs().orc_hair_ap(cos_to, eta, h, _p(np.ascontiguousarray(t, np.float64)), _p(out))
    return out


def hair_fresnel(cos_i, eta_i, eta_t):
    return _hair_sigs().orc_hair_fresnel(cos_i, eta_i, eta_t)


def hair_eval(wo, wi, h, sigma_a=(0, 0, 0), beta_m=0.3, beta_n=0.3, eta=1.55):
    """f (n x 3) and pdf (n) in the fiber frame."""
    wo = np.ascontiguousarray(wo, np.float64).reshape(-1, 3)
    wi = np.ascontiguousarray(wi, np.float64).reshape(-1, 3)
    h = np.ascontiguousarray(np.broadcast_to(h, (wo.shape[0],)), np.float64)
    f, pdf = np.zeros(wo.shape), np.zeros(wo.shape[0])
    _hair_sigs().orc_hair_eval(_p(_hp(sigma_a, beta_m, beta_n, eta)), _p(wo), _p(wi), _p(h), wo.shape[0], _p(f),
                               _p(pdf))
    return f, pdf


def hair_sample(wo, h, u, sigma_a=(0, 0, 0), beta_m=0.3, beta_n=0.3, eta=1.55):
    """hair::sample from uniforms u (n x 4): (wi, f, pdf, ok)."""
    u = np.ascontiguousarray(u, np.float64).reshape(-1, 4)
    n = u.shape[0]
    wo = np.ascontiguousarray(np.broadcast_to(wo, (n, 3)), np.float64)
    h = np.ascontiguousarray(np.broadcast_to(h, (n,)), np.float64)
    wi, f, pdf, ok = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n, np.uint8)
    _hair_sigs().orc_hair_sample(_p(_hp(sigma_a, beta_m, beta_n, eta)), _p(wo), _p(h), _p(u), n, _p(wi), _p(f),
                                 _p(pdf), _p(ok))
    return wi, f, pdf, ok.astype(bool)


def fiber_frame(tangent, n, wo):
    """(x, y, z rows, h) of FiberFrame::new."""
    xyz = np.zeros((3, 3))
    h = ctypes.c_double()
    _hair_sigs().orc_fiber_frame(_p(np.ascontiguousarray(tangent, np.float64)), _p(np.ascontiguousarray(n, np.float64)),
                                 _p(np.ascontiguousarray(wo, np.float64)), _p(xyz), ctypes.byref(h))
    return xyz, h.value
