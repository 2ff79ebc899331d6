"""ctypes binding of librrhip.so (the C ABI in include/rrhip.h).

This is the harness the tests and bench.py drive the HIP backend through,
the way a Rust `extern "C"` block would. It holds no compute: every call
goes to the native library, and importing fails loudly when the library
has not been built (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librrhip.so")

NO_HIT = 0xFFFFFFFF

RAY_DTYPE = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("tmin", "<f4"),
                      ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim", "<u4"), ("u", "<f4"), ("v", "<f4")])

# Rays one rrhip_trace_closest_noalloc call may stage (RRHIP_NOALLOC_MAX_RAYS).
NOALLOC_MAX_RAYS = 1 << 20

STATUS_NAMES = {
    0: "OK", 1: "INVALID_ARGUMENT", 2: "NO_DEVICE", 3: "DEVICE",
    4: "OUT_OF_MEMORY", 5: "UNSUPPORTED", 6: "INTERNAL",
}

# Every symbol include/rrhip.h declares (checked by the CPU test suite).
EXPORTED = [
    "rrhip_abi_version", "rrhip_last_error", "rrhip_init", "rrhip_destroy",
    "rrhip_device_name", "rrhip_stream", "rrhip_synchronize",
    "rrhip_buffer_alloc", "rrhip_buffer_free", "rrhip_copy_to_device", "rrhip_copy_to_host",
    "rrhip_scene_upload", "rrhip_scene_destroy", "rrhip_scene_get_info", "rrhip_build_info",
    "rrhip_trace_closest", "rrhip_trace_any", "rrhip_trace_closest_device",
    "rrhip_trace_any_device", "rrhip_trace_cost", "rrhip_time_trace_device",
    "rrhip_gen_heightfield", "rrhip_gen_query_rays",
    "rrhip_scene_get_tuning", "rrhip_scene_set_tuning", "rrhip_scene_autotune",
    "rrhip_device_copy",
    "rrhip_session_create", "rrhip_session_destroy", "rrhip_render_samples", "rrhip_read_accum",
    "rrhip_load_accum", "rrhip_session_get_stats", "rrhip_gen_icosphere", "rrhip_gen_fbm", "rrhip_relmse",
    "rrhip_nrc_create", "rrhip_nrc_destroy", "rrhip_nrc_get_info", "rrhip_nrc_get_params",
    "rrhip_nrc_set_params", "rrhip_nrc_infer", "rrhip_nrc_train_step", "rrhip_nrc_gradients",
    "rrhip_session_set_nrc", "rrhip_gen_density_grid",
    "rrhip_batch_create", "rrhip_batch_destroy", "rrhip_batch_render", "rrhip_batch_read_accum",
    "rrhip_batch_get_stats", "rrhip_session_set_launch_mode", "rrhip_session_autotune",
    "rrhip_session_set_adaptive", "rrhip_session_average_spp",
    "rrhip_session_set_film", "rrhip_read_film", "rrhip_pcg32_draws",
    "rrhip_trace_closest_noalloc", "rrhip_session_set_profiling", "rrhip_session_get_profile",
    "rrhip_nrc_time_steps", "rrhip_batch_render_persistent", "rrhip_hair_bsdf",
]


class RrhipError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"rrhip status {status} ({STATUS_NAMES.get(status, '?')}): {msg}")
        self.status = status


class BuildParams(ctypes.Structure):
    _fields_ = [("bins", ctypes.c_uint32), ("max_leaf", ctypes.c_uint32),
                ("reserved0", ctypes.c_uint32), ("reserved1", ctypes.c_uint32)]


class MeshDesc(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("num_vertices", ctypes.c_uint64),
                ("indices", ctypes.c_void_p), ("num_triangles", ctypes.c_uint64),
                ("normals", ctypes.c_void_p), ("st", ctypes.c_void_p), ("positions1", ctypes.c_void_p)]


D3 = ctypes.c_double * 3


# rrhip_instance_desc kinds
INSTANCE_MESH, INSTANCE_CURVES = 0, 1


class InstanceDesc(ctypes.Structure):
    _fields_ = [("mesh", ctypes.c_uint32), ("material", ctypes.c_uint32),
                ("object_to_world", ctypes.c_double * 12), ("object_to_world1", ctypes.c_double * 12),
                ("moving", ctypes.c_uint32), ("kind", ctypes.c_uint32)]


class CurvesDesc(ctypes.Structure):
    _fields_ = [("p0", ctypes.c_void_p), ("p1", ctypes.c_void_p), ("v0", ctypes.c_void_p),
                ("v1", ctypes.c_void_p), ("num_segments", ctypes.c_uint64)]


class TriangleDesc(ctypes.Structure):
    _fields_ = [("v0", D3), ("v1", D3), ("v2", D3), ("material", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class SphereDesc(ctypes.Structure):
    _fields_ = [("center", D3), ("radius", ctypes.c_double), ("material", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class MaterialDesc(ctypes.Structure):
    _fields_ = [("diffuse_gain", ctypes.c_double), ("diffuse_color", D3),
                ("diffuse_roughness", ctypes.c_double),
                ("specular_face_color", D3), ("specular_edge_color", D3),
                ("specular_roughness", ctypes.c_double), ("specular_ior", ctypes.c_double),
                ("clearcoat_gain", ctypes.c_double), ("clearcoat_roughness", ctypes.c_double),
                ("fuzz_gain", ctypes.c_double), ("fuzz_color", D3),
                ("glass_gain", ctypes.c_double), ("glass_ior", ctypes.c_double),
                ("glass_roughness", ctypes.c_double), ("refraction_color", D3),
                ("glow", D3), ("emission", D3), ("area_light", ctypes.c_int32),
                ("interior", ctypes.c_uint32), ("id", ctypes.c_uint32), ("cutout", ctypes.c_double),
                ("hair", ctypes.c_uint32), ("reserved_hair", ctypes.c_uint32), ("hair_sigma_a", D3),
                ("hair_beta_m", ctypes.c_double), ("hair_beta_n", ctypes.c_double),
                ("hair_eta", ctypes.c_double)]


class LightDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("intensity", ctypes.c_double), ("color", D3), ("p", D3), ("e1", D3), ("e2", D3),
                ("radius", ctypes.c_double)]


class CameraDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("filter", ctypes.c_uint32), ("projection", ctypes.c_uint32),
                ("fov_deg", ctypes.c_double), ("filter_width", ctypes.c_double),
                ("eye", D3), ("look_at", D3), ("up", D3),
                ("lens_radius", ctypes.c_double), ("focal_distance", ctypes.c_double),
                ("ortho_half", ctypes.c_double * 2), ("motion_inv", ctypes.c_double * 12),
                ("has_motion_inv", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class MediumDesc(ctypes.Structure):
    _fields_ = [("sigma_a", D3), ("sigma_s", D3), ("g", ctypes.c_double), ("emission", D3),
                ("max_distance", ctypes.c_double), ("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32),
                ("nz", ctypes.c_uint32), ("procedural", ctypes.c_uint32), ("density", ctypes.c_void_p),
                ("grid_lo", D3), ("grid_hi", D3), ("octaves", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("frequency", ctypes.c_double), ("gain", ctypes.c_double), ("lacunarity", ctypes.c_double),
                ("coverage", ctypes.c_double), ("sharpness", ctypes.c_double)]


MEDIUM_FBM = 1  # rrhip_medium_desc.procedural: render.rs's fBm cloud density


# rrhip_quadric_desc kinds
QUADRIC_SPHERE, QUADRIC_CYLINDER, QUADRIC_CONE, QUADRIC_DISK, QUADRIC_PARABOLOID = 0, 1, 2, 3, 4
QUADRIC_HYPERBOLOID, QUADRIC_TORUS = 5, 6


class QuadricDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("material", ctypes.c_uint32), ("radius", ctypes.c_double),
                ("zmin", ctypes.c_double), ("zmax", ctypes.c_double), ("height", ctypes.c_double),
                ("thetamax", ctypes.c_double), ("p1", ctypes.c_double * 3), ("p2", ctypes.c_double * 3),
                ("object_to_world", ctypes.c_double * 12)]


class TextureDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("texels", ctypes.POINTER(ctypes.c_float))]


# rrhip_pattern_desc kinds, wrap modes, bound fields
(PATTERN_TEXTURE, PATTERN_CHECKER, PATTERN_FRACTAL, PATTERN_MIX, PATTERN_COLOR_CORRECT, PATTERN_RAMP,
 PATTERN_TRIPLANAR) = range(7)
WRAP_PERIODIC, WRAP_CLAMP, WRAP_BLACK = range(3)
BIND_FIELDS = ("diffuseColor", "diffuseGain", "diffuseRoughness", "specularFaceColor", "specularEdgeColor",
               "specularRoughness", "clearcoatGain", "fuzzGain", "fuzzColor", "glassRefractionGain",
               "glassRoughness", "refractionColor", "glowColor", "presence")
BIND_PRESENCE = 13


class PatternDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("wrap", ctypes.c_uint32), ("texture", ctypes.c_int32),
                ("use_t", ctypes.c_uint32), ("udim", ctypes.POINTER(ctypes.c_int32)),
                ("input", ctypes.c_int32 * 3), ("octaves", ctypes.c_uint32),
                ("constant", D3 * 3), ("color_a", D3), ("color_b", D3), ("cc_gain", D3), ("cc_offset", D3),
                ("scale", ctypes.c_double * 2), ("frequency", ctypes.c_double), ("gain", ctypes.c_double),
                ("lacunarity", ctypes.c_double), ("gamma", ctypes.c_double), ("saturation", ctypes.c_double),
                ("num_stops", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("positions", ctypes.POINTER(ctypes.c_double)), ("colors", ctypes.POINTER(ctypes.c_double))]

    @classmethod
    def make(cls, kind, **kw):
        """A node with render.rs's defaults (pattern.rs, rib/mod.rs Pattern
        parsing): unit scales, checker white/black, fBm of 4 octaves, gain
        0.5, lacunarity 2, identity colour correction, inputs on their
        constants."""
        p = cls()
        p.kind = kind
        p.texture = -1
        p.input[:] = [-1, -1, -1]
        p.octaves = 4
        p.scale[:] = [1.0, 1.0]
        p.color_a[:] = [1.0, 1.0, 1.0]
        p.color_b[:] = [0.0, 0.0, 0.0]
        p.cc_gain[:] = [1.0, 1.0, 1.0]
        p.frequency = 1.0
        p.gain = 0.5
        p.lacunarity = 2.0
        p.gamma = 1.0
        p.saturation = 1.0
        for k, v in kw.items():
            if k == "input":
                p.input[:] = list(v)
            elif k == "constant":
                for i, c in enumerate(v):
                    p.constant[i][:] = list(c)
            elif k in ("color_a", "color_b", "cc_gain", "cc_offset", "scale"):
                getattr(p, k)[:] = list(v)
            else:
                setattr(p, k, v)
        return p


class BindingDesc(ctypes.Structure):
    _fields_ = [("material", ctypes.c_uint32), ("field", ctypes.c_uint32), ("pattern", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("num_meshes", ctypes.c_uint32),
                ("meshes", ctypes.POINTER(MeshDesc)), ("build", BuildParams),
                ("num_instances", ctypes.c_uint32), ("num_triangles", ctypes.c_uint32),
                ("instances", ctypes.POINTER(InstanceDesc)),
                ("triangles", ctypes.POINTER(TriangleDesc)),
                ("num_spheres", ctypes.c_uint32), ("num_materials", ctypes.c_uint32),
                ("spheres", ctypes.POINTER(SphereDesc)),
                ("materials", ctypes.POINTER(MaterialDesc)),
                ("num_lights", ctypes.c_uint32), ("reserved0", ctypes.c_uint32),
                ("lights", ctypes.POINTER(LightDesc)),
                ("camera", CameraDesc), ("background", D3),
                ("num_media", ctypes.c_uint32), ("atmosphere", ctypes.c_int32),
                ("media", ctypes.POINTER(MediumDesc)),
                ("env_width", ctypes.c_uint32), ("env_height", ctypes.c_uint32),
                ("env_rgb", ctypes.POINTER(ctypes.c_float)),
                ("num_quadrics", ctypes.c_uint32), ("reserved_q", ctypes.c_uint32),
                ("quadrics", ctypes.POINTER(QuadricDesc)),
                ("num_curve_sets", ctypes.c_uint32), ("reserved_c", ctypes.c_uint32),
                ("curve_sets", ctypes.POINTER(CurvesDesc)),
                ("num_textures", ctypes.c_uint32), ("num_patterns", ctypes.c_uint32),
                ("textures", ctypes.POINTER(TextureDesc)), ("patterns", ctypes.POINTER(PatternDesc)),
                ("num_bindings", ctypes.c_uint32), ("reserved_p", ctypes.c_uint32),
                ("bindings", ctypes.POINTER(BindingDesc))]


class RenderParams(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("max_depth", ctypes.c_uint32),
                ("rr_start", ctypes.c_uint32), ("samples_per_pass", ctypes.c_uint32),
                ("firefly_clamp", ctypes.c_double), ("sample_base", ctypes.c_uint64),
                ("ray_offset", ctypes.c_double)]

    @staticmethod
    def make(max_depth=10, rr_start=3, samples_per_pass=0, firefly_clamp=60.0, sample_base=0, ray_offset=0.0):
        """ray_offset 0: error-bounded ray origins (the default); > 0:
        render.rs's fixed push along the shading normal (1e-4). The oracle
        always follows render.rs (1e-4)."""
        return RenderParams(ctypes.sizeof(RenderParams), max_depth, rr_start, samples_per_pass,
                            firefly_clamp, sample_base, ray_offset)


class BatchJob(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("spp", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("params", RenderParams)]


class SessionStats(ctypes.Structure):
    _fields_ = [("extension_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
                ("samples_rendered", ctypes.c_uint64), ("path_slots", ctypes.c_uint64),
                ("samples_per_pass", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("last_render_ms", ctypes.c_double), ("cache_queries", ctypes.c_uint64),
                ("cache_samples", ctypes.c_uint64)]


class PtProfile(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in (
        "extend_rays", "extend_nodes", "extend_triangles", "extend_instances",
        "shadow_rays", "shadow_nodes", "shadow_triangles", "shadow_instances",
        "cache_queries", "cache_samples", "passes", "bounce_rounds")] + \
        [(k, ctypes.c_double) for k in ("extend_ms", "shade_ms", "shadow_ms", "cache_infer_ms", "cache_steps_ms")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class NrcConfig(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("levels", ctypes.c_uint32),
                ("log2_table", ctypes.c_uint32), ("seed", ctypes.c_uint32),
                ("base_res", ctypes.c_float), ("max_res", ctypes.c_float),
                ("lo", ctypes.c_float * 3), ("hi", ctypes.c_float * 3),
                ("learning_rate", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("epsilon", ctypes.c_float),
                ("batch", ctypes.c_uint32), ("steps_per_pass", ctypes.c_uint32),
                ("train_paths", ctypes.c_uint32), ("spread", ctypes.c_float), ("ema", ctypes.c_float)]

    @staticmethod
    def make(levels=16, log2_table=19, seed=1, base_res=16.0, max_res=2048.0, lo=(0, 0, 0), hi=(0, 0, 0),
             learning_rate=1e-2, beta1=0.9, beta2=0.99, epsilon=1e-15, batch=4096, steps_per_pass=16,
             train_paths=65536, spread=1e-4, ema=0.0):
        return NrcConfig(ctypes.sizeof(NrcConfig), levels, log2_table, seed, base_res, max_res,
                         (ctypes.c_float * 3)(*lo), (ctypes.c_float * 3)(*hi), learning_rate, beta1, beta2,
                         epsilon, batch, steps_per_pass, train_paths, spread, ema)


class NrcInfo(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_uint32), ("log2_table", ctypes.c_uint32),
                ("res", ctypes.c_uint32 * 16), ("mlp_params", ctypes.c_uint64),
                ("grid_params", ctypes.c_uint64), ("steps", ctypes.c_uint64),
                ("last_loss", ctypes.c_float), ("lo", ctypes.c_float * 3), ("hi", ctypes.c_float * 3),
                ("reserved", ctypes.c_uint32)]


LIGHT_POINT, LIGHT_DISTANT, LIGHT_RECT, LIGHT_SPHERE, LIGHT_DISK, LIGHT_DOME = range(6)
ACCUM_STRIDE = 8
# rrhip_read_film columns (RRHIP_FILM_CHANNELS)
FILM_CHANNELS = 18
FILM_COLUMNS = {"beauty": slice(0, 3), "diffuse": slice(3, 6), "specular": slice(6, 9),
                "albedo": slice(9, 12), "normal": slice(12, 15), "depth": 15, "id": 16, "coverage": 17}


class SceneInfo(ctypes.Structure):
    _fields_ = [("num_triangles", ctypes.c_uint64), ("num_nodes", ctypes.c_uint64),
                ("num_leaves", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64),
                ("depth", ctypes.c_uint32), ("node_bytes", ctypes.c_uint32),
                ("tri_bytes", ctypes.c_uint32), ("depth4", ctypes.c_uint32),
                ("build_ms", ctypes.c_double),
                ("bounds_lo", ctypes.c_float * 3), ("bounds_hi", ctypes.c_float * 3)]


class TraceTuning(ctypes.Structure):
    _fields_ = [("block", ctypes.c_uint32), ("stack", ctypes.c_uint32),
                ("rays_per_wave", ctypes.c_uint32), ("width", ctypes.c_uint32)]

    def as_dict(self):
        return {"block": self.block, "stack": self.stack, "rays_per_wave": self.rays_per_wave,
                "width": self.width or 2}


_lib = None


def load():
    """Loads librrhip.so (raises if it is missing: build it first)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                           "or python render.rs_amd/build_lib.py")
    lib = ctypes.CDLL(LIB_PATH)
    vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
    sig = {
        "rrhip_abi_version": (u32, []),
        "rrhip_last_error": (ctypes.c_char_p, []),
        "rrhip_init": (i32, [i32, ctypes.POINTER(vp)]),
        "rrhip_destroy": (i32, [vp]),
        "rrhip_device_name": (i32, [vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]),
        "rrhip_stream": (vp, [vp]),
        "rrhip_synchronize": (i32, [vp]),
        "rrhip_buffer_alloc": (i32, [vp, u64, ctypes.POINTER(vp)]),
        "rrhip_buffer_free": (i32, [vp, vp]),
        "rrhip_copy_to_device": (i32, [vp, vp, vp, u64]),
        "rrhip_copy_to_host": (i32, [vp, vp, vp, u64]),
        "rrhip_scene_upload": (i32, [vp, ctypes.POINTER(SceneDesc), ctypes.POINTER(vp)]),
        "rrhip_scene_destroy": (i32, [vp]),
        "rrhip_scene_get_info": (i32, [vp, ctypes.POINTER(SceneInfo)]),
        "rrhip_build_info": (i32, [ctypes.POINTER(SceneDesc), ctypes.POINTER(SceneInfo)]),
        "rrhip_trace_closest": (i32, [vp, vp, u64, vp]),
        "rrhip_trace_any": (i32, [vp, vp, u64, vp]),
        "rrhip_trace_closest_device": (i32, [vp, vp, u64, vp, vp]),
        "rrhip_trace_closest_noalloc": (i32, [vp, vp, u64, vp]),
        "rrhip_trace_any_device": (i32, [vp, vp, u64, vp, vp]),
        "rrhip_trace_cost": (i32, [vp, vp, u64, i32, ctypes.POINTER(u64)]),
        "rrhip_time_trace_device": (i32, [vp, vp, u64, i32, vp, u32, ctypes.POINTER(ctypes.c_float)]),
        "rrhip_device_copy": (i32, [vp, vp, vp, u64]),
        "rrhip_session_create": (i32, [vp, ctypes.POINTER(RenderParams), ctypes.POINTER(vp)]),
        "rrhip_session_destroy": (i32, [vp]),
        "rrhip_render_samples": (i32, [vp, u32, u32]),
        "rrhip_read_accum": (i32, [vp, vp]),
        "rrhip_load_accum": (i32, [vp, vp]),
        "rrhip_session_get_stats": (i32, [vp, ctypes.POINTER(SessionStats)]),
        "rrhip_session_set_profiling": (i32, [vp, i32]),
        "rrhip_nrc_time_steps": (i32, [vp, vp, vp, u64, u32, ctypes.POINTER(ctypes.c_float)]),
        "rrhip_session_get_profile": (i32, [vp, ctypes.POINTER(PtProfile)]),
        "rrhip_relmse": (i32, [vp, vp, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]),
        "rrhip_pcg32_draws": (i32, [vp, vp, vp, u32, u32, vp]),
        "rrhip_hair_bsdf": (i32, [vp, ctypes.POINTER(MaterialDesc), u32, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "rrhip_nrc_create": (i32, [vp, ctypes.POINTER(NrcConfig), ctypes.POINTER(vp)]),
        "rrhip_nrc_destroy": (i32, [vp]),
        "rrhip_nrc_get_info": (i32, [vp, ctypes.POINTER(NrcInfo)]),
        "rrhip_nrc_get_params": (i32, [vp, vp, vp]),
        "rrhip_nrc_set_params": (i32, [vp, vp, vp]),
        "rrhip_nrc_infer": (i32, [vp, vp, u64, vp]),
        "rrhip_nrc_train_step": (i32, [vp, vp, vp, u64, ctypes.POINTER(ctypes.c_float)]),
        "rrhip_nrc_gradients": (i32, [vp, vp, vp, u64, ctypes.POINTER(ctypes.c_float), vp, vp]),
        "rrhip_session_set_nrc": (i32, [vp, vp]),
        "rrhip_scene_get_tuning": (i32, [vp, ctypes.POINTER(TraceTuning)]),
        "rrhip_scene_set_tuning": (i32, [vp, ctypes.POINTER(TraceTuning)]),
        "rrhip_scene_autotune": (i32, [vp, vp, u64, i32, vp, ctypes.POINTER(TraceTuning),
                                       ctypes.POINTER(ctypes.c_float)]),
        "rrhip_session_set_launch_mode": (i32, [vp, u32]),
        "rrhip_session_autotune": (i32, [vp, u32, ctypes.POINTER(u32), ctypes.POINTER(ctypes.c_float)]),
        "rrhip_session_set_adaptive": (i32, [vp, ctypes.c_double]),
        "rrhip_session_set_film": (i32, [vp, i32]),
        "rrhip_read_film": (i32, [vp, vp]),
        "rrhip_session_average_spp": (i32, [vp, ctypes.POINTER(ctypes.c_double)]),
        "rrhip_batch_create": (i32, [vp, u32, vp, u32, ctypes.POINTER(vp)]),
        "rrhip_batch_destroy": (i32, [vp]),
        "rrhip_batch_render": (i32, [vp, vp, ctypes.POINTER(ctypes.c_double)]),
        "rrhip_batch_render_persistent": (i32, [vp, vp, ctypes.POINTER(ctypes.c_double)]),
        "rrhip_batch_read_accum": (i32, [vp, u32, vp]),
        "rrhip_batch_get_stats": (i32, [vp, ctypes.POINTER(SessionStats)]),
        "rrhip_gen_density_grid": (i32, [u32, u64, ctypes.c_float, u32, ctypes.c_float, ctypes.c_float,
                                         ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), vp]),
        "rrhip_gen_icosphere": (i32, [u32, u64, ctypes.c_float, ctypes.c_float, u32, vp, vp, vp]),
        "rrhip_gen_fbm": (i32, [u64, vp, ctypes.c_double, u32, ctypes.c_double, ctypes.c_double, vp, vp]),
        "rrhip_gen_heightfield": (i32, [u32, u32, u64, ctypes.c_float, ctypes.c_float, u32, vp, vp]),
        "rrhip_gen_query_rays": (i32, [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                       ctypes.c_float, u64, u64, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(status):
    if status != 0:
        msg = load().rrhip_last_error().decode(errors="replace")
        raise RrhipError(status, msg)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---- workload generators -------------------------------------------------

def gen_heightfield(nx, ny, seed, amplitude=0.1, frequency=4.0, octaves=6):
    """Displaced heightfield grid (positions (V,3) f32, indices (T,3) u32)."""
    lib = load()
    pos = np.empty(((nx + 1) * (ny + 1), 3), np.float32)
    idx = np.empty((nx * ny * 2, 3), np.uint32)
    _check(lib.rrhip_gen_heightfield(nx, ny, seed, amplitude, frequency, octaves, _ptr(pos), _ptr(idx)))
    return pos, idx


def gen_fbm(points, frequency=1.0, octaves=4, gain=0.5, lacunarity=2.0, offset=(0.0, 0.0, 0.0)):
    """render.rs's fBm (displace.rs) at points (n x 3), on the host."""
    lib = load()
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    out = np.empty(p.shape[0], np.float64)
    off = np.ascontiguousarray(offset, np.float64)
    _check(lib.rrhip_gen_fbm(p.shape[0], _ptr(p), float(frequency), int(octaves), float(gain), float(lacunarity),
                             _ptr(off), _ptr(out)))
    return out


def gen_icosphere(levels, seed, amplitude=0.15, frequency=3.0, octaves=4):
    """Displaced icosphere: positions (V,3), indices (T,3), normals (V,3)."""
    lib = load()
    nv = 10 * 4 ** levels + 2
    nt = 20 * 4 ** levels
    pos = np.empty((nv, 3), np.float32)
    idx = np.empty((nt, 3), np.uint32)
    nrm = np.empty((nv, 3), np.float32)
    _check(lib.rrhip_gen_icosphere(levels, seed, amplitude, frequency, octaves, _ptr(pos), _ptr(idx), _ptr(nrm)))
    return pos, idx, nrm


def gen_density_grid(n, seed, frequency=2.0, octaves=4, coverage=0.5, sharpness=3.0, lo=(-1, -1, -1),
                     hi=(1, 1, 1)):
    """fBm cloud density (medium.rs DensityField::Fbm) on an n^3 vertex grid (z, y, x order)."""
    out = np.empty((n, n, n), np.float32)
    lo_c = (ctypes.c_float * 3)(*[float(x) for x in lo])
    hi_c = (ctypes.c_float * 3)(*[float(x) for x in hi])
    _check(load().rrhip_gen_density_grid(n, seed, frequency, octaves, coverage, sharpness, lo_c, hi_c, _ptr(out)))
    return out


def gen_query_rays(lo, hi, inflate, seed, n):
    lib = load()
    rays = np.empty(n, RAY_DTYPE)
    lo_c = (ctypes.c_float * 3)(*[float(x) for x in lo])
    hi_c = (ctypes.c_float * 3)(*[float(x) for x in hi])
    _check(lib.rrhip_gen_query_rays(lo_c, hi_c, inflate, seed, n, _ptr(rays)))
    return rays


def mesh_bounds(positions):
    return positions.min(axis=0), positions.max(axis=0)


def _info_dict(i):
    return {f: (list(getattr(i, f)) if f.startswith("bounds") else getattr(i, f))
            for f, _ in SceneInfo._fields_ if not f.startswith("reserved")}


def _mesh_desc(positions, indices, bins, max_leaf):
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
    mesh = MeshDesc(_ptr(pos).value, pos.shape[0], _ptr(idx).value, idx.shape[0], None, None, None)
    desc = SceneDesc()
    desc.struct_size = ctypes.sizeof(SceneDesc)
    desc.num_meshes = 1
    desc.meshes = ctypes.pointer(mesh)
    desc.build = BuildParams(bins, max_leaf, 0, 0)
    return pos, idx, mesh, desc


def build_info(positions, indices, bins=32, max_leaf=4):
    """Host-only BVH build statistics (no device needed)."""
    keep = _mesh_desc(positions, indices, bins, max_leaf)
    i = SceneInfo()
    _check(load().rrhip_build_info(ctypes.byref(keep[3]), ctypes.byref(i)))
    return _info_dict(i)


# ---- context / scene -----------------------------------------------------

class Context:
    def __init__(self, device=0):
        lib = load()
        h = ctypes.c_void_p()
        _check(lib.rrhip_init(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            load().rrhip_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_name(self):
        name = ctypes.create_string_buffer(256)
        arch = ctypes.create_string_buffer(256)
        _check(load().rrhip_device_name(self.h, name, 256, arch, 256))
        return name.value.decode(), arch.value.decode()

    def synchronize(self):
        _check(load().rrhip_synchronize(self.h))

    def upload_mesh(self, positions, indices, bins=32, max_leaf=4):
        return Scene(self, positions, indices, bins, max_leaf)

    # device buffers
    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        _check(load().rrhip_buffer_alloc(self.h, nbytes, ctypes.byref(p)))
        return p

    def free(self, p):
        _check(load().rrhip_buffer_free(self.h, p))

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        _check(load().rrhip_copy_to_device(self.h, p, _ptr(arr), arr.nbytes))
        return p

    def stream(self):
        return load().rrhip_stream(self.h)

    def device_copy(self, dst, src, nbytes):
        _check(load().rrhip_device_copy(self.stream(), dst, src, nbytes))

    def to_host(self, p, out):
        _check(load().rrhip_copy_to_host(self.h, _ptr(out), p, out.nbytes))
        return out

    def pcg32_draws(self, pixels, samples, count):
        """(n, count) u32: the first draws of each (pixel, sample) stream,
        generated on the device."""
        px = np.ascontiguousarray(pixels, np.uint64).ravel()
        sm = np.ascontiguousarray(samples, np.uint64).ravel()
        if px.shape != sm.shape:
            raise ValueError("pcg32_draws: pixels and samples differ in length")
        out = np.empty((px.size, count), np.uint32)
        _check(load().rrhip_pcg32_draws(self.h, _ptr(px), _ptr(sm), px.size, count, _ptr(out)))
        return out

    def hair_bsdf(self, material, wo, wi, h, u):
        """The device fiber BSDF of a hair MaterialDesc on n fiber-frame
        queries: (f, pdf) towards wi, and a sample (ws, fs, pdfs) from the
        uniforms u (n x 4); pdfs = 0 where sampling fails."""
        wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
        n = wo.shape[0]
        wi = np.ascontiguousarray(wi, np.float32).reshape(n, 3)
        h = np.ascontiguousarray(np.broadcast_to(h, (n,)), np.float32)
        u = np.ascontiguousarray(u, np.float32).reshape(n, 4)
        f, pdf = np.empty((n, 3), np.float32), np.empty(n, np.float32)
        ws, fs, pdfs = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32), np.empty(n, np.float32)
        _check(load().rrhip_hair_bsdf(self.h, ctypes.byref(material), n, _ptr(wo), _ptr(wi), _ptr(h), _ptr(u),
                                      _ptr(f), _ptr(pdf), _ptr(ws), _ptr(fs), _ptr(pdfs)))
        return f, pdf, ws, fs, pdfs


class PtScene:
    """A path-tracing scene uploaded from a scenes.Scene description."""

    def __init__(self, ctx, desc):
        self.ctx = ctx
        self._desc = desc
        h = ctypes.c_void_p()
        _check(load().rrhip_scene_upload(ctx.h, ctypes.byref(desc), ctypes.byref(h)))
        self.h = h
        self.npix = desc.camera.width * desc.camera.height

    def close(self):
        if self.h:
            load().rrhip_scene_destroy(self.h)
            self.h = None

    def session(self, params):
        return Session(self, params)


class Session:
    def __init__(self, scene, params):
        self.scene = scene
        self.params = params
        h = ctypes.c_void_p()
        _check(load().rrhip_session_create(scene.h, ctypes.byref(params), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            load().rrhip_session_destroy(self.h)
            self.h = None

    def render(self, start, count):
        _check(load().rrhip_render_samples(self.h, start, count))

    def accum(self):
        out = np.empty((self.scene.npix, ACCUM_STRIDE), np.float64)
        _check(load().rrhip_read_accum(self.h, _ptr(out)))
        return out

    def load_accum(self, acc=None):
        """Replaces the accumulator with `acc`, an (npix, ACCUM_STRIDE)
        array (None clears it). Any other shape is refused: the C call
        reads exactly npix * ACCUM_STRIDE doubles."""
        if acc is None:
            _check(load().rrhip_load_accum(self.h, None))
            return
        acc = np.ascontiguousarray(acc, np.float64)
        if acc.shape != (self.scene.npix, ACCUM_STRIDE):
            raise ValueError(f"load_accum: expected an accumulator of shape ({self.scene.npix}, {ACCUM_STRIDE}), "
                             f"got {acc.shape}")
        _check(load().rrhip_load_accum(self.h, _ptr(acc)))

    def set_launch_mode(self, mode):
        """Bit 0: one hipGraph per pass size; bit 1: queue sorted by shading
        class; bit 2: cache optimiser steps on the rendering stream."""
        _check(load().rrhip_session_set_launch_mode(self.h, mode))

    def autotune(self, samples):
        """Times every launch mode, keeps the fastest: returns (mode, ms)."""
        mode = ctypes.c_uint32()
        ms = ctypes.c_float()
        _check(load().rrhip_session_autotune(self.h, samples, ctypes.byref(mode), ctypes.byref(ms)))
        return mode.value, ms.value

    def set_adaptive(self, tolerance):
        """Adaptive sampling (WfSession::set_adaptive, renderer.rs:580-584):
        pixels stop once their 95 % confidence half-width of the mean
        luminance is below tolerance * max(mean, 0.05), checked before
        samples 32, 48, 64, ...; 0 turns it off."""
        _check(load().rrhip_session_set_adaptive(self.h, float(tolerance)))

    def set_film(self, enable=True):
        """Film mode (pt::render_film): later samples also accumulate the
        AOVs; the film is cleared now and with the accumulator."""
        _check(load().rrhip_session_set_film(self.h, 1 if enable else 0))

    def film(self):
        """(npix, FILM_CHANNELS) f64: beauty, diffuse, specular, albedo,
        normal, depth, dominant id, its coverage (FILM_COLUMNS)."""
        out = np.empty((self.scene.npix, FILM_CHANNELS), np.float64)
        _check(load().rrhip_read_film(self.h, _ptr(out)))
        return out

    def average_spp(self):
        """Mean samples per pixel (WfSession::average_spp, renderer.rs:587-591)."""
        out = ctypes.c_double()
        _check(load().rrhip_session_average_spp(self.h, ctypes.byref(out)))
        return out.value

    def render_adaptive(self, max_spp, tolerance):
        """pt::render_adaptive (src/raytracer/pt/mod.rs:285-336) on this
        session: clears the accumulator, renders samples [0, max_spp) with
        adaptive stopping, returns the average samples per pixel; the image
        is accum()[:, :3] / accum()[:, 6:7]."""
        self.load_accum(None)
        self.set_adaptive(tolerance)
        self.render(0, max(int(max_spp), 1))
        return self.average_spp()

    def set_nrc(self, nrc):
        """Attaches a radiance cache (Nrc) to the session, or detaches it (None)."""
        _check(load().rrhip_session_set_nrc(self.h, None if nrc is None else nrc.h))
        self._nrc = nrc  # the cache must outlive its use by the session

    def relmse(self, reference, eps=0.01):
        out = ctypes.c_double()
        _check(load().rrhip_relmse(self.h, reference.h, eps, ctypes.byref(out)))
        return out.value

    def stats(self):
        st = SessionStats()
        _check(load().rrhip_session_get_stats(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in SessionStats._fields_ if f != "reserved"}

    def set_profiling(self, enable=True):
        """Counting trace kernels and per-kernel device times from now on
        (the profile starts empty); False turns it off."""
        _check(load().rrhip_session_set_profiling(self.h, int(bool(enable))))

    def profile(self):
        p = PtProfile()
        _check(load().rrhip_session_get_profile(self.h, ctypes.byref(p)))
        return p.as_dict()


class Scene:
    def __init__(self, ctx, positions, indices, bins=32, max_leaf=4):
        lib = load()
        self.ctx = ctx
        self._pos, self._idx, self._mesh, desc = _mesh_desc(positions, indices, bins, max_leaf)
        h = ctypes.c_void_p()
        _check(lib.rrhip_scene_upload(ctx.h, ctypes.byref(desc), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            load().rrhip_scene_destroy(self.h)
            self.h = None

    def info(self):
        i = SceneInfo()
        _check(load().rrhip_scene_get_info(self.h, ctypes.byref(i)))
        return _info_dict(i)

    def get_tuning(self):
        t = TraceTuning()
        _check(load().rrhip_scene_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, block, stack, rays_per_wave, width=2):
        t = TraceTuning(block, stack, rays_per_wave, width)
        _check(load().rrhip_scene_set_tuning(self.h, ctypes.byref(t)))

    def autotune(self, d_rays, n, d_out, any_hit=False):
        t = TraceTuning()
        ms = ctypes.c_float()
        _check(load().rrhip_scene_autotune(self.h, d_rays, n, int(any_hit), d_out, ctypes.byref(t),
                                           ctypes.byref(ms)))
        return t.as_dict(), ms.value

    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.empty(rays.shape[0], HIT_DTYPE)
        _check(load().rrhip_trace_closest(self.h, _ptr(rays), rays.shape[0], _ptr(hits)))
        return hits

    def trace_closest_noalloc(self, rays):
        """Closest hit staged through the library's fixed per-device array
        (no device allocation; at most NOALLOC_MAX_RAYS rays)."""
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.empty(rays.shape[0], HIT_DTYPE)
        _check(load().rrhip_trace_closest_noalloc(self.h, _ptr(rays), rays.shape[0], _ptr(hits)))
        return hits

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.empty(rays.shape[0], np.uint8)
        _check(load().rrhip_trace_any(self.h, _ptr(rays), rays.shape[0], _ptr(occ)))
        return occ

    def trace_cost(self, rays, any_hit=False):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = (ctypes.c_uint64 * 2)()
        _check(load().rrhip_trace_cost(self.h, _ptr(rays), rays.shape[0], int(any_hit), out))
        return int(out[0]), int(out[1])

    def trace_closest_device(self, d_rays, n, d_hits, stream=None):
        _check(load().rrhip_trace_closest_device(self.h, d_rays, n, d_hits, stream))

    def trace_any_device(self, d_rays, n, d_occ, stream=None):
        _check(load().rrhip_trace_any_device(self.h, d_rays, n, d_occ, stream))

    def time_trace_device(self, d_rays, n, d_out, reps, any_hit=False):
        ms = ctypes.c_float()
        _check(load().rrhip_time_trace_device(self.h, d_rays, n, int(any_hit), d_out, reps, ctypes.byref(ms)))
        return ms.value


# ---- neural radiance cache -------------------------------------------------

NRC_MLP_PARAMS = 19648


def nrc_queries(pos, dir, normal, roughness, albedo):
    """Packs per-query arrays into the (n, 16) f32 rrhip_nrc_query layout."""
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    q = np.zeros((pos.shape[0], 16), np.float32)
    q[:, 0:3] = pos
    q[:, 3] = np.asarray(roughness, np.float32).reshape(-1)
    q[:, 4:7] = np.asarray(dir, np.float32).reshape(-1, 3)
    q[:, 8:11] = np.asarray(normal, np.float32).reshape(-1, 3)
    q[:, 12:15] = np.asarray(albedo, np.float32).reshape(-1, 3)
    return q


class Nrc:
    """Neural radiance cache (rrhip_nrc_*): hash grid + MLP on the device."""

    def __init__(self, ctx, config=None):
        h = ctypes.c_void_p()
        cfg = config if config is not None else NrcConfig.make()
        _check(load().rrhip_nrc_create(ctx.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.ctx = ctx
        self.config = cfg

    def close(self):
        if getattr(self, "h", None):
            load().rrhip_nrc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def info(self):
        i = NrcInfo()
        _check(load().rrhip_nrc_get_info(self.h, ctypes.byref(i)))
        d = {f: getattr(i, f) for f, _ in NrcInfo._fields_ if f != "reserved"}
        d["res"] = list(i.res)[: i.levels]
        d["lo"], d["hi"] = list(i.lo), list(i.hi)
        return d

    def get_params(self):
        info = self.info()
        mlp = np.empty(NRC_MLP_PARAMS, np.float32)
        grid = np.empty(info["grid_params"], np.float32)
        _check(load().rrhip_nrc_get_params(self.h, _ptr(mlp), _ptr(grid)))
        return mlp, grid

    def set_params(self, mlp=None, grid=None):
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        g = None if grid is None else np.ascontiguousarray(grid, np.float32)
        if m is not None and m.size != NRC_MLP_PARAMS:
            raise ValueError("mlp must hold 19648 floats")
        if g is not None and g.size != self.info()["grid_params"]:
            raise ValueError("grid size mismatch")
        _check(load().rrhip_nrc_set_params(self.h, None if m is None else _ptr(m), None if g is None else _ptr(g)))

    def infer(self, queries):
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 16)
        out = np.empty((q.shape[0], 3), np.float32)
        _check(load().rrhip_nrc_infer(self.h, _ptr(q), q.shape[0], _ptr(out)))
        return out

    def train_step(self, queries, targets):
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        if t.shape[0] != q.shape[0]:
            raise ValueError("one target per query")
        loss = ctypes.c_float()
        _check(load().rrhip_nrc_train_step(self.h, _ptr(q), _ptr(t), q.shape[0], ctypes.byref(loss)))
        return loss.value

    def time_steps(self, queries, targets, steps):
        """Milliseconds of `steps` optimiser steps on one staged batch."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        ms = ctypes.c_float()
        _check(load().rrhip_nrc_time_steps(self.h, _ptr(q), _ptr(t), q.shape[0], steps, ctypes.byref(ms)))
        return ms.value

    def gradients(self, queries, targets):
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        if t.shape[0] != q.shape[0]:
            raise ValueError("one target per query")
        loss = ctypes.c_float()
        gm = np.empty(NRC_MLP_PARAMS, np.float32)
        gg = np.empty(self.info()["grid_params"], np.float32)
        _check(load().rrhip_nrc_gradients(self.h, _ptr(q), _ptr(t), q.shape[0], ctypes.byref(loss), _ptr(gm),
                                          _ptr(gg)))
        return loss.value, gm, gg


# ---- batched jobs -------------------------------------------------------------

class Batch:
    """Many path-tracing jobs advanced through one launch sequence per
    pass (rrhip_batch_*). jobs: list of (PtScene, spp, RenderParams)."""

    def __init__(self, ctx, jobs, samples_per_pass=0):
        self.ctx = ctx
        self._scenes = [j[0] for j in jobs]  # the scenes must outlive the batch
        arr = (BatchJob * len(jobs))(*[BatchJob(sc.h, spp, 0, prm) for sc, spp, prm in jobs])
        self.npix = [sc.npix for sc, _, _ in jobs]
        h = ctypes.c_void_p()
        _check(load().rrhip_batch_create(ctx.h, len(jobs), ctypes.addressof(arr), samples_per_pass,
                                         ctypes.byref(h)))
        self.h = h
        self.n = len(jobs)

    def close(self):
        if getattr(self, "h", None):
            load().rrhip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def render(self):
        """Returns (per-job completion ms from the start, total ms)."""
        job_ms = np.zeros(self.n)
        total = ctypes.c_double()
        _check(load().rrhip_batch_render(self.h, _ptr(job_ms), ctypes.byref(total)))
        return job_ms, total.value

    def render_persistent(self):
        """Every job in one persistent launch: (per-job completion ms from
        the launch start, total ms)."""
        job_ms = np.zeros(self.n)
        total = ctypes.c_double()
        _check(load().rrhip_batch_render_persistent(self.h, _ptr(job_ms), ctypes.byref(total)))
        return job_ms, total.value

    def accum(self, job):
        out = np.empty((self.npix[job], ACCUM_STRIDE), np.float64)
        _check(load().rrhip_batch_read_accum(self.h, job, _ptr(out)))
        return out

    def stats(self):
        st = SessionStats()
        _check(load().rrhip_batch_get_stats(self.h, ctypes.byref(st)))
        return {"extension_rays": st.extension_rays, "shadow_rays": st.shadow_rays,
                "path_slots": st.path_slots, "samples_per_pass": st.samples_per_pass, "passes": st.reserved}
