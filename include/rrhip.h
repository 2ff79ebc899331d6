/*
 * rrhip — C ABI of the MI355X (gfx950) path-tracing backend for render.rs.
 *
 * render.rs selects a backend with `enum Backend { Cpu, Metal }`
 * (src/main.rs:11-17) and each dispatch arm calls the `raytracer::metal`
 * function set directly (src/raytracer/metal/mod.rs:13). This header is the
 * C-ABI counterpart of that set for a `Hip` arm: a Rust `extern "C"` block
 * binds these symbols. All entry points are blocking
 * unless they take a stream, return an `rrhip_status`, and leave a message
 * for `rrhip_last_error()` on failure. Handles are opaque; everything the
 * library allocates is freed by the matching *_destroy call.
 *
 * Threading: one context per host thread, like PtSession, which render.rs
 * documents as "create and use it on one thread"
 * (src/raytracer/metal/renderer.rs:256-259).
 */
#ifndef RRHIP_H
#define RRHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RRHIP_ABI_VERSION 10u
#define RRHIP_NO_HIT 0xFFFFFFFFu

/* Status codes. render.rs returns anyhow::Result with a message
 * (e.g. src/raytracer/metal/renderer.rs:72-91); here the class is the
 * return value and the message is rrhip_last_error(). */
typedef enum rrhip_status {
  RRHIP_OK = 0,
  RRHIP_ERR_INVALID_ARGUMENT = 1,
  RRHIP_ERR_NO_DEVICE = 2,
  RRHIP_ERR_DEVICE = 3, /* HIP runtime error or a kernel fault surfaced */
  RRHIP_ERR_OUT_OF_MEMORY = 4,
  RRHIP_ERR_UNSUPPORTED = 5,
  RRHIP_ERR_INTERNAL = 6
} rrhip_status;

typedef struct rrhip_ctx rrhip_ctx;
typedef struct rrhip_scene rrhip_scene;
typedef struct rrhip_session rrhip_session;
typedef struct rrhip_nrc rrhip_nrc;
typedef struct rrhip_batch rrhip_batch;

/* Query ray, 32 bytes. render.rs's GPU ray query takes [f32; 6]
 * (origin, direction) and tests t > 1e-6 (src/raytracer/metal/renderer.rs:
 * 813-877); its CPU mesh path accepts 1e-9 < t < t_max
 * (src/geometry/mesh.rs:141). Here both bounds travel with the ray:
 * a hit needs tmin < t < tmax, t measured along the direction as given
 * (not renormalised, as for render.rs's instance-local rays,
 * src/geometry/mesh.rs:316-323). */
typedef struct rrhip_ray {
  float ox, oy, oz, tmin;
  float dx, dy, dz, tmax;
} rrhip_ray;

/* Closest hit, 16 bytes: t, the ORIGINAL triangle index (render.rs's CPU
 * MeshHit.tri convention, src/accel/bvh.rs:168), barycentrics u, v of
 * v1 and v2. prim == RRHIP_NO_HIT and t == +inf on a miss. */
typedef struct rrhip_hit {
  float t;
  uint32_t prim;
  float u, v;
} rrhip_hit;

/* BVH build parameters. render.rs builds with 16 bins and leaves of up to
 * 4 (SAH may stop at 16) (src/accel/bvh.rs:91-92). Zero fields take the
 * defaults: 32 bins, leaves of at most 4 triangles. */
typedef struct rrhip_build_params {
  uint32_t bins;
  uint32_t max_leaf;
  uint32_t reserved0;
  uint32_t reserved1;
} rrhip_build_params;

/* One triangle mesh: f32 xyz positions, u32 vertex indices (3 per
 * triangle), optional per-vertex f32 shading normals (NULL: geometric
 * normals) and texture coordinates (st pairs, NULL: none) — render.rs's
 * Mesh storage (src/geometry/mesh.rs:20-34). */
typedef struct rrhip_mesh_desc {
  const float* positions;
  uint64_t num_vertices;
  const uint32_t* indices;
  uint64_t num_triangles;
  const float* normals;
  const float* st;
  /* Deformation motion (Mesh::positions1, mesh.rs:31): the vertices at
   * shutter close, lerped with `positions` at each ray's shutter time;
   * NULL: static. A deforming mesh shades with geometric normals at that
   * time (its vertex normals belong to the rest pose) and takes its
   * texture coordinates' density and gradients from the rest pose. */
  const float* positions1;
} rrhip_mesh_desc;

/* A set of curve segments (render.rs CurveSet, src/geometry/curves.rs:
 * 15-26): each segment is a rounded cone — a sphere swept from p0 to p1
 * with its radius lerped between them, spherical caps at both ends — given
 * as xyz + radius at each end (4 floats per end), and the curve parameter v
 * at both ends (0 root, 1 tip). A point (p0 == p1) is a sphere. Curves and
 * points are diced into segments by the caller (curves.rs:191-258,
 * scene_builder.rs:1050-1170). */
typedef struct rrhip_curves_desc {
  const float* p0;
  const float* p1;
  const float* v0;
  const float* v1;
  uint64_t num_segments;
} rrhip_curves_desc;

/* Instance kinds (render.rs GeomKind, src/geometry/mesh.rs:223-240). */
enum { RRHIP_INSTANCE_MESH = 0, RRHIP_INSTANCE_CURVES = 1 };

/* A placed copy of a mesh or curve set: row-major 3x4 object-to-world
 * affine transform and a material (render.rs Instance,
 * src/geometry/mesh.rs:223-305). */
typedef struct rrhip_instance_desc {
  uint32_t mesh;                /* index into meshes, or into curve_sets for RRHIP_INSTANCE_CURVES */
  uint32_t material;
  double object_to_world[12];
  double object_to_world1[12];  /* at shutter close (Instance::transform1) when `moving`; rays at
                                   time t see the element-wise lerp of the two (mesh.rs:294-305) */
  uint32_t moving;
  uint32_t kind;                /* RRHIP_INSTANCE_MESH or RRHIP_INSTANCE_CURVES */
} rrhip_instance_desc;

/* Loose world-space triangle (render.rs builds these from Polygon,
 * src/parser/scene_builder.rs:911-949; src/geometry/extra.rs:651-730). */
typedef struct rrhip_triangle_desc {
  double v0[3], v1[3], v2[3];
  uint32_t material;
  uint32_t reserved;
} rrhip_triangle_desc;

/* Loose sphere (render.rs Sphere with a translation, src/geometry/sphere.rs). */
typedef struct rrhip_sphere_desc {
  double center[3];
  double radius;
  uint32_t material;
  uint32_t reserved;
} rrhip_sphere_desc;

/* RenderMan quadrics as render.rs intersects them (src/geometry/sphere.rs,
 * cylinder.rs, cone.rs, extra.rs Disk / Paraboloid): object space is the
 * quadric's own (z the axis), placed by object_to_world (row-major 3x4);
 * thetamax in degrees. sphere: radius, zmin, zmax; cylinder: radius, zmin,
 * zmax; cone: height, radius (apex at the origin, radius at z = height);
 * disk: height (its z), radius; paraboloid: radius (= rmax), zmin, zmax;
 * hyperboloid: the line p1 - p2 swept about z; torus: radius (major),
 * height (minor), zmin / zmax (phimin / phimax, degrees). */
enum {
  RRHIP_QUADRIC_SPHERE = 0,
  RRHIP_QUADRIC_CYLINDER = 1,
  RRHIP_QUADRIC_CONE = 2,
  RRHIP_QUADRIC_DISK = 3,
  RRHIP_QUADRIC_PARABOLOID = 4,
  RRHIP_QUADRIC_HYPERBOLOID = 5,
  RRHIP_QUADRIC_TORUS = 6
};
typedef struct rrhip_quadric_desc {
  uint32_t kind;
  uint32_t material;
  double radius, zmin, zmax, height, thetamax;
  double p1[3], p2[3];
  double object_to_world[12];
} rrhip_quadric_desc;

/* Material: render.rs's PbrParams lobe stack (src/scene/pbr.rs:11-48) plus
 * Material::emission / area_light (src/scene/material.rs:14-40). */
typedef struct rrhip_material_desc {
  double diffuse_gain;
  double diffuse_color[3];
  double diffuse_roughness;       /* Oren-Nayar sigma; 0 = Lambert */
  double specular_face_color[3];  /* Schlick F0; black -> from specular_ior */
  double specular_edge_color[3];  /* Schlick F90 */
  double specular_roughness;
  double specular_ior;            /* 1 disables the specular lobe */
  double clearcoat_gain;
  double clearcoat_roughness;
  double fuzz_gain;
  double fuzz_color[3];
  double glass_gain;
  double glass_ior;
  double glass_roughness;
  double refraction_color[3];
  double glow[3];
  double emission[3];
  int32_t area_light;             /* light index for MIS on emitter hits, -1 = none */
  uint32_t interior;              /* 1 + index of the medium inside closed surfaces of this
                                     material (Material::interior); 0 = none. With no lobes and
                                     no emission the surface is an invisible volume hull */
  uint32_t id;                    /* object id for the id AOV (Material::id); 0 = none */
  double cutout;                  /* 1 - presence (PbrParams::presence): the probability that
                                     a ray passes through the surface; 0 = opaque */
  uint32_t hair;                  /* 1: a hair fiber (Material::hair, Bxdf "PxrMarschnerHair"):
                                     curve hits scatter with render.rs's R / TT / TRT fiber BSDF
                                     (src/raytracer/pt/hair.rs); other hits use the lobes above */
  uint32_t reserved_hair;
  double hair_sigma_a[3];         /* absorption per unit fiber diameter (HairParams::sigma_a) */
  double hair_beta_m;             /* longitudinal roughness, clamped to [0.01, 1] */
  double hair_beta_n;             /* azimuthal roughness, clamped to [0.01, 1] */
  double hair_eta;                /* fiber index of refraction */
} rrhip_material_desc;

/* Light kinds (render.rs LightType, src/scene/light.rs:7-33). */
enum {
  RRHIP_LIGHT_POINT = 0,
  RRHIP_LIGHT_DISTANT = 1,
  RRHIP_LIGHT_RECT = 2,
  RRHIP_LIGHT_SPHERE = 3,
  RRHIP_LIGHT_DISK = 4,
  RRHIP_LIGHT_DOME = 5 /* dome: constant radiance, or the scene desc's environment map */
};

/* Light: radiance = color * intensity. p = position (point), corner (rect),
 * centre (sphere/disk) or travel direction (distant); e1/e2 = rect edges or
 * disk axes; radius = sphere radius or distant angular radius. */
typedef struct rrhip_light_desc {
  uint32_t kind;
  uint32_t reserved;
  double intensity;
  double color[3];
  double p[3];
  double e1[3];
  double e2[3];
  double radius;
} rrhip_light_desc;

/* Camera in RenderMan's left-handed frame: forward = normalize(look_at -
 * eye), right = normalize(up x forward), up' = forward x right —
 * render.rs's Camera::new frame (src/scene/camera.rs: 96-119: at the RIB
 * default eye = 0, look +z, up +y it gives right = +x), placed at `eye`.
 * Pixel filter per src/scene/camera.rs:14-66 (0 box, 1 triangle, 2
 * gaussian); rays per Camera::generate_ray_lens (camera.rs:150-179):
 * perspective with an optional thin lens, or orthographic. */
typedef struct rrhip_camera_desc {
  uint32_t width, height;
  uint32_t filter;
  uint32_t projection;    /* 0 perspective, 1 orthographic */
  double fov_deg;         /* vertical (perspective) */
  double filter_width;
  double eye[3], look_at[3], up[3];
  double lens_radius;     /* aperture radius of the thin lens; 0 = pinhole */
  double focal_distance;  /* focal plane distance along forward (lens_radius > 0) */
  double ortho_half[2];   /* orthographic screen-window half extents along right, up */
  double motion_inv[12];  /* Camera::motion_inv (3x4 row-major): rays at shutter time t are mapped by
                             lerp(I, motion_inv, t) (camera.rs:134-146) when has_motion_inv */
  uint32_t has_motion_inv;
  uint32_t reserved;
} rrhip_camera_desc;

/* Scene description. `struct_size` must be sizeof(rrhip_scene_desc).
 * The caller's arrays are copied during rrhip_scene_upload.
 * Ray queries use exactly one mesh and no other geometry. Path-tracing
 * scenes place meshes through instances and add loose triangles and
 * spheres (render.rs Scene::objects, src/scene/mod.rs:27-56). */
/* A participating medium (render.rs Medium, src/scene/medium.rs:28-65):
 * absorption and scattering coefficients, Henyey-Greenstein anisotropy,
 * emission, extent along any ray, and an optional density field, either
 * - a voxel grid (nx x ny x nz samples, x fastest, values in [0, 1],
 *   trilinear between samples placed on the corners of [grid_lo, grid_hi];
 *   zero outside), or
 * - procedural = RRHIP_MEDIUM_FBM: render.rs's cloud field
 *   DensityField::Fbm (medium.rs:16-27, 55-63), evaluated per lookup at the
 *   world point p: rho = clamp((fbm(p) * 0.5 + 0.5 - (1 - coverage)) *
 *   sharpness, 0, 1), fbm with frequency, max(octaves, 1) (at most 64),
 *   gain, lacunarity and offset 0 (render.rs's RIB defaults: 0.3, 5, 0.55,
 *   2, coverage 0.5, sharpness 4); nx must be 0.
 * nx = 0 and procedural = 0: homogeneous. */
#define RRHIP_MEDIUM_FBM 1u
typedef struct rrhip_medium_desc {
  double sigma_a[3];
  double sigma_s[3];
  double g;
  double emission[3];
  double max_distance;  /* <= 0: unbounded */
  uint32_t nx, ny, nz, procedural;
  const float* density;
  double grid_lo[3];
  double grid_hi[3];
  uint32_t octaves, reserved;
  double frequency, gain, lacunarity, coverage, sharpness;
} rrhip_medium_desc;

/* A texture: linear RGB f32 texels of level 0, rows from the top
 * (render.rs's .tex contents, src/texture/tex.rs). The library builds the
 * mip chain with tex.rs's 2x2 box filter down to 1x1. */
typedef struct rrhip_texture_desc {
  uint32_t width, height;
  const float* texels;  /* width * height rgb triples */
} rrhip_texture_desc;

/* Pattern node kinds (render.rs PatternNode, src/texture/pattern.rs:42-68)
 * and texture wrap modes (cache.rs:20-35). */
enum {
  RRHIP_PATTERN_TEXTURE = 0,       /* PxrTexture: st-mapped texture (UDIM tiles optional) */
  RRHIP_PATTERN_CHECKER = 1,       /* PxrChecker, filtered towards the mean by the footprint */
  RRHIP_PATTERN_FRACTAL = 2,       /* PxrFractal: fBm over the world position, to [0, 1] */
  RRHIP_PATTERN_MIX = 3,           /* PxrMix: lerp(color1, color2, mix.r) */
  RRHIP_PATTERN_COLOR_CORRECT = 4, /* PxrColorCorrect: gain, offset, gamma, saturation */
  RRHIP_PATTERN_RAMP = 5,          /* PxrRamp: piecewise-linear over s or t */
  RRHIP_PATTERN_TRIPLANAR = 6      /* triplanar texture blended by |N| */
};
enum { RRHIP_WRAP_PERIODIC = 0, RRHIP_WRAP_CLAMP = 1, RRHIP_WRAP_BLACK = 2 };

/* One pattern node. Inputs name another node, or -1 for the constant
 * beside them; an input more than 16 levels down, or naming no node,
 * evaluates to magenta (pattern.rs:203-212). The GPU evaluates each
 * binding's input tree in post-order: at most 256 nodes per tree with at
 * most 8 inputs pending at once (RRHIP_ERR_UNSUPPORTED beyond). */
typedef struct rrhip_pattern_desc {
  uint32_t kind;
  uint32_t wrap;            /* texture, triplanar */
  int32_t texture;          /* texture, triplanar: texture index; -1 = missing (renders magenta) */
  uint32_t use_t;           /* ramp: over t instead of s */
  const int32_t* udim;      /* texture: NULL, or 100 texture indices of UDIM tiles 1001..1100 (-1: none) */
  int32_t input[3];         /* mix: color1, color2, mix; color correct: input */
  uint32_t octaves;         /* fractal layers (at most 64) */
  double constant[3][3];    /* the inputs' constants */
  double color_a[3], color_b[3];    /* checker */
  double cc_gain[3], cc_offset[3];  /* color correct */
  double scale[2];          /* texture scaleS / scaleT; checker sScale / tScale */
  double frequency;         /* fractal, triplanar */
  double gain, lacunarity;  /* fractal */
  double gamma, saturation; /* color correct */
  uint32_t num_stops;       /* ramp */
  uint32_t reserved;
  const double* positions;  /* ramp: num_stops positions */
  const double* colors;     /* ramp: num_stops rgb */
} rrhip_pattern_desc;

/* PbrParams fields a pattern drives (BoundField, pattern.rs:227-272):
 * colours take the pattern's rgb, scalars its red channel. */
enum {
  RRHIP_BIND_DIFFUSE_COLOR = 0,
  RRHIP_BIND_DIFFUSE_GAIN = 1,
  RRHIP_BIND_DIFFUSE_ROUGHNESS = 2,
  RRHIP_BIND_SPECULAR_FACE_COLOR = 3,
  RRHIP_BIND_SPECULAR_EDGE_COLOR = 4,
  RRHIP_BIND_SPECULAR_ROUGHNESS = 5, /* clamped to [0.005, 1] */
  RRHIP_BIND_CLEARCOAT_GAIN = 6,
  RRHIP_BIND_FUZZ_GAIN = 7,
  RRHIP_BIND_FUZZ_COLOR = 8,
  RRHIP_BIND_GLASS_GAIN = 9,
  RRHIP_BIND_GLASS_ROUGHNESS = 10,   /* clamped to [0.005, 1] */
  RRHIP_BIND_REFRACTION_COLOR = 11,
  RRHIP_BIND_GLOW = 12,
  RRHIP_BIND_PRESENCE = 13           /* clamped to [0, 1]; cutouts for camera and shadow rays */
};
typedef struct rrhip_binding_desc {  /* Material::pattern_bindings entry */
  uint32_t material;
  uint32_t field;    /* RRHIP_BIND_* */
  uint32_t pattern;  /* node index */
  uint32_t reserved;
} rrhip_binding_desc;

typedef struct rrhip_scene_desc {
  uint32_t struct_size;
  uint32_t num_meshes;
  const rrhip_mesh_desc* meshes;
  rrhip_build_params build;
  uint32_t num_instances;
  uint32_t num_triangles;
  const rrhip_instance_desc* instances;
  const rrhip_triangle_desc* triangles;
  uint32_t num_spheres;
  uint32_t num_materials;
  const rrhip_sphere_desc* spheres;
  const rrhip_material_desc* materials;
  uint32_t num_lights;
  uint32_t reserved0;
  const rrhip_light_desc* lights;
  rrhip_camera_desc camera;
  double background[3]; /* miss radiance when no dome light */
  /* Participating media; `atmosphere` (-1: none) is the medium every ray
   * travels in (render.rs Scene::atmosphere, src/scene/mod.rs). */
  uint32_t num_media;
  int32_t atmosphere;
  const rrhip_medium_desc* media;
  /* Lat-long environment map of the dome light (render.rs EnvMap,
   * src/scene/envmap.rs): env_width x env_height texels of linear rgb
   * radiance (3 floats each), row-major from the +y pole; a direction d
   * maps to u = atan2(d.x, -d.z) / 2 pi (wrapped to [0, 1)) and
   * v = acos(d.y) / pi; scaled by the dome's color * intensity. NULL: a
   * constant dome. Requires a dome light. */
  uint32_t env_width, env_height;
  const float* env_rgb;
  /* Loose quadrics (besides the spheres above), tested like them. */
  uint32_t num_quadrics;
  uint32_t reserved;
  const rrhip_quadric_desc* quadrics;
  /* Curve sets, placed by instances of kind RRHIP_INSTANCE_CURVES. */
  uint32_t num_curve_sets;
  uint32_t reserved_curves;
  const rrhip_curves_desc* curve_sets;
  /* Textures, pattern nodes and the materials' pattern bindings
   * (Scene::patterns, Material::pattern_bindings): a bound field of a
   * material follows its pattern at every shading point, filtered by the
   * ray cone's footprint (pt/mod.rs:389-440). Texture coordinates come
   * from meshes' st, spheres, quadrics (cones none) and curves ([0.5, v]);
   * loose triangles have none; a bound presence also decides shadow rays
   * at every blocker (pt/mod.rs:1294). At most 15 bindings per material;
   * radiance caches and batches refuse scenes with bindings
   * (RRHIP_ERR_UNSUPPORTED). */
  uint32_t num_textures;
  uint32_t num_patterns;
  const rrhip_texture_desc* textures;
  const rrhip_pattern_desc* patterns;
  uint32_t num_bindings;
  uint32_t reserved_patterns;
  const rrhip_binding_desc* bindings;
} rrhip_scene_desc;

/* Path-tracing parameters (render.rs constants, src/raytracer/pt/mod.rs:
 * 26-32). */
typedef struct rrhip_render_params {
  uint32_t struct_size;
  uint32_t max_depth;        /* bounces; render.rs CPU 10, Metal 8 */
  uint32_t rr_start;         /* Russian roulette from this depth (3) */
  uint32_t samples_per_pass; /* samples of every pixel in flight at once; 0 = auto */
  double firefly_clamp;      /* per-sample luminance clamp (60); <= 0 disables */
  uint64_t sample_base;      /* added to the sample index when seeding PCG32; 0 = render.rs */
  double ray_offset;         /* ray origins leaving a surface. 0 (default): error-bounded — the hit
                                point is rebuilt from the surface (barycentrics, sphere reprojection),
                                carries a bound on its f32 rounding error, and rays start outside
                                that bound along the geometric normal, one ulp further (Pharr, Jakob,
                                Humphreys, PBRT 3rd ed. sec. 3.9); shadow rays also stop short of the
                                light by the sampled light point's bound. > 0: render.rs's fixed
                                push of this length along the shading normal (1e-4,
                                src/raytracer/pt/mod.rs:30, 976-977, 1212-1224), which falls below
                                one ulp once coordinates pass ~1e3 */
} rrhip_render_params;

/* Per-pixel accumulator layout of rrhip_read_accum / orc_render:
 * 8 doubles per pixel, row-major: sum r, g, b; sum of squares r, g, b;
 * sample count; sum of squared luminance (0.2126 r + 0.7152 g + 0.0722 b).
 * All sums are of firefly-clamped samples. */
#define RRHIP_ACCUM_STRIDE 8

/* Per-pixel layout of rrhip_read_film / orc_render_film (render.rs's Film,
 * src/output/film.rs:41-55, as render_film fills it, pt/mod.rs:163-276):
 * 18 doubles per pixel, row-major: beauty r, g, b; diffuse r, g, b
 * (radiance whose first continuation was not specular-like); specular
 * r, g, b (max(beauty - diffuse, 0)); albedo r, g, b and shading normal
 * x, y, z at the first real hit (sample means, 0 where a sample hit
 * nothing); depth (mean distance travelled to the first real hit over the
 * samples that had one, 0 if none); dominant object id; that id's
 * fraction of the samples. */
#define RRHIP_FILM_CHANNELS 18

/* Scene statistics (what the upload built). */
typedef struct rrhip_scene_info {
  uint64_t num_triangles;
  uint64_t num_nodes;     /* 64-byte child-pair nodes */
  uint64_t num_leaves;
  uint64_t device_bytes;  /* nodes + triangle records */
  uint32_t depth;         /* node levels on the longest root-to-leaf path */
  uint32_t node_bytes;    /* 64 */
  uint32_t tri_bytes;     /* 48 */
  uint32_t depth4;        /* node levels of the four-wide tree (128-byte nodes); 0: not built */
  double build_ms;        /* host build + flatten time */
  float bounds_lo[3];
  float bounds_hi[3];
} rrhip_scene_info;

/* Launch configuration of the ray-query kernels (per scene). Each wave
 * owns `rays_per_wave` consecutive rays and refills finished lanes from
 * them; `stack` is the per-lane LDS traversal stack (>= the tree's depth).
 * No counterpart in render.rs (its Metal host fixes threadgroups of 64 /
 * 256, src/raytracer/metal/renderer.rs:390, 637). */
typedef struct rrhip_trace_tuning {
  uint32_t block;          /* 64, 128 or 256 threads */
  uint32_t stack;          /* 16, 24, 32, 48 or 64 entries */
  uint32_t rays_per_wave;  /* multiple of 64, <= 2^20; 0 = fill: the rays split evenly over the
                              wave slots the device holds at once (one round of resident waves) */
  uint32_t width;          /* tree width: 2 (64-byte child-pair nodes) or 4 (128-byte nodes); 0 = 2 */
} rrhip_trace_tuning;

/* ---- errors / version --------------------------------------------------- */
uint32_t rrhip_abi_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char* rrhip_last_error(void);

/* ---- context ------------------------------------------------------------ */
/* Opens HIP device `device` (gfx950 required) and creates the context's
 * stream. Replaces render.rs's PtSession device/queue setup
 * (src/raytracer/metal/renderer.rs:280-300). */
int rrhip_init(int device, rrhip_ctx** out);
int rrhip_destroy(rrhip_ctx* ctx);
/* Device name (e.g. "AMD Instinct MI355X") and arch (e.g. "gfx950:..."). */
int rrhip_device_name(rrhip_ctx* ctx, char* name, size_t name_len, char* arch, size_t arch_len);
/* The context's HIP stream (hipStream_t), for callers that order their
 * own work around the library's. */
void* rrhip_stream(rrhip_ctx* ctx);
int rrhip_synchronize(rrhip_ctx* ctx);

/* ---- device buffers (for device-resident batches) ------------------------ */
int rrhip_buffer_alloc(rrhip_ctx* ctx, uint64_t bytes, void** dptr);
int rrhip_buffer_free(rrhip_ctx* ctx, void* dptr);
int rrhip_copy_to_device(rrhip_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int rrhip_copy_to_host(rrhip_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Device-to-device copy by a plain kernel on `stream` (hipStream_t; no
 * wait). 16-byte aligned buffers, bytes a multiple of 16. No counterpart in
 * render.rs; its HBM counters calibrate traffic measurements. */
int rrhip_device_copy(void* stream, void* dst, const void* src, uint64_t bytes);

/* ---- scene -------------------------------------------------------------- */
/* Builds the binned-SAH BVH on the host and uploads the 64-byte nodes and
 * 48-byte leaf-ordered triangle records. Counterpart of render.rs's
 * Mesh::new -> Bvh::build (src/geometry/mesh.rs:48-84, src/accel/bvh.rs:
 * 96-108) followed by GpuPtScene::build's BLAS export
 * (src/raytracer/metal/gpu_scene.rs:473-531). */
int rrhip_scene_upload(rrhip_ctx* ctx, const rrhip_scene_desc* desc, rrhip_scene** out);
int rrhip_scene_destroy(rrhip_scene* scene);
int rrhip_scene_get_info(rrhip_scene* scene, rrhip_scene_info* info);
/* Runs the host build of rrhip_scene_upload without a device and reports
 * what it built: for a ray-query mesh the BVH (device_bytes = what an
 * upload would allocate); for a path-tracing scene the instance hierarchy
 * (num_nodes, BLAS + loose triangles, num_leaves = instances, depth = the
 * traversal stack entries a lane can need, bounds of all geometry). */
int rrhip_build_info(const rrhip_scene_desc* desc, rrhip_scene_info* info);

/* Current / explicit launch configuration (set validates against the
 * scene's depth and returns RRHIP_ERR_INVALID_ARGUMENT when unusable).
 * A scene keeps one configuration per query mode: set applies to both,
 * get reports the closest-hit one. */
int rrhip_scene_get_tuning(rrhip_scene* scene, rrhip_trace_tuning* tuning);
int rrhip_scene_set_tuning(rrhip_scene* scene, const rrhip_trace_tuning* tuning);
/* Autotuner: times every supported configuration on the caller's
 * device-resident rays (closest hit if any == 0, else any hit, writing
 * into `out`), keeps the fastest as the scene's configuration for that
 * mode and reports it with its per-launch milliseconds. */
int rrhip_scene_autotune(rrhip_scene* scene, const rrhip_ray* rays, uint64_t n, int any, void* out,
                         rrhip_trace_tuning* best, float* best_ms);

/* ---- ray queries -------------------------------------------------------- */
/* Closest hit for n host rays into n host hits. Counterpart of render.rs's
 * intersect_probe (src/raytracer/metal/renderer.rs:813-877) generalised to
 * a BVH: Bvh::traverse + Mesh::intersect_triangle (src/accel/bvh.rs:141-204,
 * src/geometry/mesh.rs:113-142). Hit ids and t equal the f64 CPU result. */
int rrhip_trace_closest(rrhip_scene* scene, const rrhip_ray* rays, uint64_t n, rrhip_hit* hits);
/* Any hit with tmin < t < tmax (occluded[i] = 0/1). Counterpart of
 * Bvh::any_hit + Instance::occludes (src/accel/bvh.rs:208-245,
 * src/geometry/mesh.rs:384-407). */
int rrhip_trace_any(rrhip_scene* scene, const rrhip_ray* rays, uint64_t n, uint8_t* occluded);

/* Device-resident variants: `rays`, `hits`, `occluded` are device
 * pointers; work is enqueued on `stream` (hipStream_t; NULL = the
 * context's stream) and the call returns without waiting. */
int rrhip_trace_closest_device(rrhip_scene* scene, const rrhip_ray* rays, uint64_t n,
                               rrhip_hit* hits, void* stream);
int rrhip_trace_any_device(rrhip_scene* scene, const rrhip_ray* rays, uint64_t n,
                           uint8_t* occluded, void* stream);

/* rrhip_trace_closest without any device allocation: no counterpart in
 * render.rs. For callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc). Rays and hits are staged through
 * a fixed per-device array of RRHIP_NOALLOC_MAX_RAYS x 48 bytes (48 MiB),
 * resolved by rrhip_init and held until the process ends; n above
 * RRHIP_NOALLOC_MAX_RAYS is rejected with RRHIP_ERR_INVALID_ARGUMENT.
 * Results equal rrhip_trace_closest's. Every call on a device shares that
 * device's array: calls that overlap on one device (from several threads or
 * contexts) race. */
#define RRHIP_NOALLOC_MAX_RAYS (1u << 20)
int rrhip_trace_closest_noalloc(rrhip_scene* scene, const rrhip_ray* rays, uint64_t n, rrhip_hit* hits);

/* Traversal cost of n host rays as the kernels execute it: totals of node
 * fetches (64 B each) and triangle fetches (48 B each). any != 0 counts the
 * any-hit traversal. out[0] = nodes, out[1] = triangles. No counterpart in
 * render.rs; it feeds the HBM traffic model of the ray-query kernel. */
int rrhip_trace_cost(rrhip_scene* scene, const rrhip_ray* rays, uint64_t n, int any,
                     uint64_t out[2]);

/* ---- path-tracing sessions ------------------------------------------------ */
typedef struct rrhip_session_stats {
  uint64_t extension_rays;   /* closest-hit rays traced since creation */
  uint64_t shadow_rays;      /* NEE shadow rays traced since creation */
  uint64_t samples_rendered; /* per pixel, summed over render calls */
  uint64_t path_slots;       /* samples_per_pass x pixels */
  uint32_t samples_per_pass;
  uint32_t reserved;
  double last_render_ms;     /* device time of the last rrhip_render_samples */
  uint64_t cache_queries;    /* radiance-cache queries (terminal vertices and training tails) */
  uint64_t cache_samples;    /* radiance-cache training samples */
} rrhip_session_stats;

/* Creates a wavefront path-tracing session on a scene with a camera.
 * Counterpart of PtSession::new / WfSession::new
 * (src/raytracer/metal/renderer.rs:260-300, 501-560). */
int rrhip_session_create(rrhip_scene* scene, const rrhip_render_params* params, rrhip_session** out);
int rrhip_session_destroy(rrhip_session* session);
/* Renders samples [start, start + count) of every pixel and adds them to
 * the f64 accumulator (blocking; with a radiance cache attached, the
 * optimiser steps of the last pass may still run on the cache's stream,
 * and every later use of the cache waits for them). Per-(pixel, sample)
 * PCG32 seeding makes sample ranges compose exactly. Counterpart of
 * PtSession::render_samples (renderer.rs:410-437). */
int rrhip_render_samples(rrhip_session* session, uint32_t start, uint32_t count);
/* Copies the accumulator (width x height x RRHIP_ACCUM_STRIDE doubles) to
 * the host. Counterpart of PtSession::sum_and_weight (renderer.rs:476-490);
 * the image is sum / count. */
int rrhip_read_accum(rrhip_session* session, double* out);
/* Replaces the accumulator (NULL clears it): checkpoint resume, as
 * render.rs's load_accum (renderer.rs:237-254). */
int rrhip_load_accum(rrhip_session* session, const double* in);
int rrhip_session_get_stats(rrhip_session* session, rrhip_session_stats* stats);
/* How passes are launched (bit flags, results identical in every mode):
 * 1 = each pass size captured once as a hipGraph and replayed (the first
 * sample index is fed from device memory; not with a radiance cache or
 * adaptive sampling, whose passes are launched directly); 2 = the live
 * queue sorted by shading class (miss / active-lobe set of the hit
 * material) before shade, so a wave runs one BSDF kind; 4 = a radiance
 * cache's optimiser steps stay on the rendering stream (by default they
 * run on the cache's own stream, overlapped with the next pass's tracing
 * also across rrhip_render_samples calls, and every later use of the
 * cache waits for them). 0 = plain kernel launches (default). No
 * counterpart in render.rs (its Metal host encodes one command buffer per
 * dispatch, renderer.rs:380-437). */
int rrhip_session_set_launch_mode(rrhip_session* session, uint32_t mode);
/* Autotuner: times the graph and sorted-queue launch modes (bits 0 and 1;
 * bit 2 stays as the caller set it) on `samples` samples per pixel,
 * keeps the fastest and reports it with its milliseconds. The trials are
 * discarded: the accumulator, the film accumulator, the statistics and
 * any attached cache are left as they were, also when a trial fails (the
 * launch mode then stays the caller's). */
int rrhip_session_autotune(rrhip_session* session, uint32_t samples, uint32_t* best_mode, float* best_ms);
/* Adaptive sampling: with tolerance > 0, rrhip_render_samples checks every
 * pixel before samples 32, 48, 64, ... (index relative to sample_base) and
 * stops sampling those whose 95 % confidence half-width of the mean
 * luminance, 1.96 sqrt(var / n), is below tolerance * max(mean, 0.05);
 * stopped pixels stay stopped until this call or rrhip_load_accum. 0 turns
 * it off. Counterpart of WfSession::set_adaptive (renderer.rs:580-584) with
 * the schedule of pt::render_adaptive (src/raytracer/pt/mod.rs:285-336):
 * rendering samples [0, max_spp) from a cleared accumulator is that
 * function. Passes then run as plain launches (no graph replay). */
int rrhip_session_set_adaptive(rrhip_session* session, double tolerance);
/* Mean sample count over the pixels (the accumulator's count column).
 * Counterpart of WfSession::average_spp (renderer.rs:587-591). */
int rrhip_session_average_spp(rrhip_session* session, double* out);
/* Film mode (render.rs's render_pt_film / pt::render_film): with enable != 0
 * every later sample also accumulates the AOVs of RRHIP_FILM_CHANNELS
 * (cleared now, and with the accumulator by rrhip_load_accum); 0 turns it
 * off and frees its buffers. The id vote keeps the first four distinct ids
 * a pixel sees (exact when a pixel sees at most four; ties go to the id
 * seen first). RRHIP_ERR_UNSUPPORTED with a radiance cache attached. */
int rrhip_session_set_film(rrhip_session* session, int enable);
/* Copies the film (width x height x RRHIP_FILM_CHANNELS doubles) to `out`;
 * RRHIP_ERR_INVALID_ARGUMENT when film mode is off. */
int rrhip_read_film(rrhip_session* session, double* out);
/* Profiling: with enable != 0 every later pass launches counting variants
 * of the trace kernels (same results) and times each bounce's kernels and
 * the radiance cache's work with device events, waiting for each pass
 * (graph replay is off meanwhile); the profile starts empty now. 0 turns
 * it off. No counterpart in render.rs; it feeds the traversal bytes model
 * of the path-tracing benchmarks. */
typedef struct rrhip_pt_profile {
  uint64_t extend_rays, extend_nodes, extend_triangles, extend_instances;
  uint64_t shadow_rays, shadow_nodes, shadow_triangles, shadow_instances;
  /* nodes: 64-byte TLAS / BLAS node fetches; triangles: 48-byte triangle
   * records tested; instances: instances entered (64-byte record + 16-byte
   * mesh record); loose objects are tested by every ray and not counted */
  uint64_t cache_queries, cache_samples;
  uint64_t passes, bounce_rounds;
  double extend_ms, shade_ms, shadow_ms; /* summed device time (shade includes a queue sort) */
  double cache_infer_ms;                 /* radiance-cache inference */
  double cache_steps_ms;                 /* optimiser steps, first begin to last end (cache stream) */
} rrhip_pt_profile;
int rrhip_session_set_profiling(rrhip_session* session, int enable);
/* The profile since profiling was enabled (RRHIP_ERR_INVALID_ARGUMENT
 * when it is off). */
int rrhip_session_get_profile(rrhip_session* session, rrhip_pt_profile* out);

/* relMSE of `session`'s mean image against `reference`'s (same size and
 * context): mean over pixels and channels of (x - r)^2 / (r^2 + eps),
 * reduced in FP64 on the GPU. No counterpart in render.rs (its tests use
 * RMSE, tests/pt_tests.rs:77-88); it is the error measure of the
 * time-to-quality benchmark (eps = 0.01). */
int rrhip_relmse(rrhip_session* session, rrhip_session* reference, double eps, double* out);

/* Diagnostic: the first `count` (1..4096) u32 draws of the PCG32 stream of
 * each (pixels[i], samples[i]), generated on the device by the seeding and
 * step the camera-path kernel uses, into out[i * count + k]. Counterpart of
 * Pcg32::for_pixel_sample + next_u32 (src/raytracer/pt/sampler.rs:24-42);
 * lets a caller check the device stream against the CPU path's. */
int rrhip_pcg32_draws(rrhip_ctx* ctx, const uint64_t* pixels, const uint64_t* samples, uint32_t n,
                      uint32_t count, uint32_t* out);

/* Diagnostic: render.rs's hair fiber BSDF (src/raytracer/pt/hair.rs:
 * f, pdf, sample) as the path tracer's shade kernel evaluates it on the
 * device, for n queries in the fiber frame (x along the fiber): f (rgb) and
 * pdf towards wi[3i..] from wo[3i..] at azimuthal offset h[i], and a sample
 * drawn from the uniforms u[4i..4i+3] in hair.rs's order (lobe, two
 * longitudinal, azimuthal): its direction ws, value fs and pdf pdfs (0 where
 * sampling fails). `material` must have hair = 1. Host arrays; lets a
 * caller check the device fiber BSDF against the CPU path's. */
int rrhip_hair_bsdf(rrhip_ctx* ctx, const rrhip_material_desc* material, uint32_t n, const float* wo,
                    const float* wi, const float* h, const float* u, float* f, float* pdf, float* ws, float* fs,
                    float* pdfs);

/* ---- neural radiance cache ------------------------------------------------ */
/* render.rs has no radiance cache: this section is the learned part of the
 * HIP backend (a multiresolution hash grid + 48-64x5-3 ReLU MLP with bf16
 * weights on the matrix cores, trained online with Adam on the relative
 * L2 loss). */
typedef struct rrhip_nrc_config {
  uint32_t struct_size;     /* sizeof(rrhip_nrc_config) */
  uint32_t levels;          /* hash-grid levels, 1..16 */
  uint32_t log2_table;      /* entries per level = 2^log2_table, 8..24 */
  uint32_t seed;            /* parameter initialisation */
  float base_res, max_res;  /* cells per axis of the coarsest / finest level */
  float lo[3], hi[3];       /* position bounds; hi <= lo on an axis: the scene's bounds */
  float learning_rate, beta1, beta2, epsilon; /* Adam */
  uint32_t batch;           /* samples per optimiser step, 64..65536 */
  uint32_t steps_per_pass;  /* optimiser steps per rendering pass */
  uint32_t train_paths;     /* training paths traced per rendering pass */
  float spread;             /* area-spread termination constant c */
  float ema;                /* inference uses the debiased moving average of the
                               parameters with this decay, [0, 1); 0 = the
                               trained parameters themselves */
} rrhip_nrc_config;

/* A cache query / training sample (64 bytes): the network input at a
 * path vertex. dir points from the vertex towards the viewer; normal is
 * the shading normal on dir's side. */
typedef struct rrhip_nrc_query {
  float pos[3], roughness;
  float dir[3], pad0;
  float normal[3], pad1;
  float albedo[3], pad2;
} rrhip_nrc_query;

typedef struct rrhip_nrc_info {
  uint32_t levels, log2_table;
  uint32_t res[16];         /* cells per axis of each level */
  uint64_t mlp_params;      /* f32 count of rrhip_nrc_get_params' mlp array */
  uint64_t grid_params;     /* levels * 2^log2_table * 2 */
  uint64_t steps;           /* optimiser steps taken */
  float last_loss;          /* loss of the last training batch */
  float lo[3], hi[3];       /* bounds in use */
  uint32_t reserved;
} rrhip_nrc_info;

/* Creates a cache with freshly initialised parameters (grid U(-1e-4,
 * 1e-4), He-uniform weights, both from PCG32(seed)). config = NULL takes
 * the defaults: 16 levels, 2^19 entries, resolutions 16..2048, Adam
 * (1e-2, 0.9, 0.99, 1e-15), 16 steps of 4096 samples (65536 samples)
 * and 65536 training paths per pass, c = 1e-4, no moving average. */
int rrhip_nrc_create(rrhip_ctx* ctx, const rrhip_nrc_config* config, rrhip_nrc** out);
int rrhip_nrc_destroy(rrhip_nrc* nrc);
int rrhip_nrc_get_info(rrhip_nrc* nrc, rrhip_nrc_info* info);
/* Parameters in f32: mlp = W0 [64][48], W1..W4 [64][64], W5 [3][64]
 * (row = output unit), grid = [level][entry][feature]. Either may be NULL.
 * set_params also clears the optimiser state. */
int rrhip_nrc_get_params(rrhip_nrc* nrc, float* mlp, float* grid);
int rrhip_nrc_set_params(rrhip_nrc* nrc, const float* mlp, const float* grid);
/* Cache radiance (rgb, n x 3 floats) of n host queries. */
int rrhip_nrc_infer(rrhip_nrc* nrc, const rrhip_nrc_query* queries, uint64_t n, float* rgb);
/* One optimiser step on n <= 65536 samples with targets (n x 3); the
 * batch's loss (before the step) into *loss when non-NULL. */
int rrhip_nrc_train_step(rrhip_nrc* nrc, const rrhip_nrc_query* queries, const float* targets, uint64_t n,
                         float* loss);
/* Timing: stages the n samples once, runs `steps` optimiser steps on them
 * back to back between HIP events on the context stream and returns the
 * milliseconds of the whole sequence (the parameters advance as with
 * rrhip_nrc_train_step). */
int rrhip_nrc_time_steps(rrhip_nrc* nrc, const rrhip_nrc_query* queries, const float* targets, uint64_t n,
                         uint32_t steps, float* elapsed_ms);
/* Loss and gradients of a batch without a step (mlp_grad / grid_grad in
 * the get_params layouts; either may be NULL). */
int rrhip_nrc_gradients(rrhip_nrc* nrc, const rrhip_nrc_query* queries, const float* targets, uint64_t n,
                        float* loss, float* mlp_grad, float* grid_grad);
/* Attaches a cache to a session (NULL detaches): render passes then
 * terminate paths into the cache by the area-spread criterion and train
 * it online on the suffixes of extra training paths. Cache and session
 * must share a context. RRHIP_ERR_UNSUPPORTED in film mode. */
int rrhip_session_set_nrc(rrhip_session* session, rrhip_nrc* nrc);

/* ---- batched jobs ---------------------------------------------------------- */
/* Many independent renders (different scenes, resolutions, sample counts,
 * media) advanced together: every pass runs one raygen, one extend /
 * shade / shadow per bounce and one resolve over the path slots of all
 * unfinished jobs, instead of one launch sequence per job. No counterpart
 * in render.rs, which renders one scene per process; each job's
 * accumulator equals that of a session with the batch's
 * samples_per_pass. */
typedef struct rrhip_batch_job {
  rrhip_scene* scene;          /* a path-tracing scene (camera set) */
  uint32_t spp;                /* samples of every pixel */
  uint32_t reserved;
  rrhip_render_params params;  /* max_depth, rr_start, firefly_clamp, sample_base (samples_per_pass ignored) */
} rrhip_batch_job;

/* samples_per_pass > 0: every unfinished job renders that many samples per
 * pass (its accumulator then equals a session's with the same
 * samples_per_pass). 0: passes of up to 2^24 path slots filled shortest
 * job first (fewest samples x pixels), so small jobs complete early. */
int rrhip_batch_create(rrhip_ctx* ctx, uint32_t num_jobs, const rrhip_batch_job* jobs, uint32_t samples_per_pass,
                       rrhip_batch** out);
int rrhip_batch_destroy(rrhip_batch* batch);
/* Renders every job to completion (blocking, accumulators cleared
 * first). job_ms (num_jobs entries, may be NULL) receives each job's
 * completion time from the start, measured with device events at pass
 * granularity; total_ms the whole batch. */
int rrhip_batch_render(rrhip_batch* batch, double* job_ms, double* total_ms);
/* Every job in ONE persistent launch: each resident lane pulls (job,
 * pixel) work items from a device counter — pixels of the shortest jobs
 * (samples x pixels) first — and runs all of the pixel's samples through
 * the same camera / closest-hit / shading / shadow device functions as the
 * passes, adding them in sample order; the last pixel of a job stamps the
 * device clock. job_ms (may be NULL) receives each job's completion time
 * from the launch's first block start (device clock), total_ms the
 * launch (events). Accumulators are cleared first and hold what a session
 * with samples_per_pass = the job's spp renders. Error-bounded ray origins
 * only (RRHIP_ERR_UNSUPPORTED if a job sets ray_offset). No counterpart in
 * render.rs. */
int rrhip_batch_render_persistent(rrhip_batch* batch, double* job_ms, double* total_ms);
int rrhip_batch_read_accum(rrhip_batch* batch, uint32_t job, double* out);
/* Rays traced by the last render (extension, shadow), passes, samples_per_pass. */
int rrhip_batch_get_stats(rrhip_batch* batch, rrhip_session_stats* stats);

/* ---- timing ------------------------------------------------------------- */
/* Enqueues `reps` back-to-back closest-hit (any == 0) or any-hit launches
 * over device-resident rays on the context stream, bracketed by HIP
 * events, waits, and returns the event-measured milliseconds of the
 * whole sequence. No counterpart in render.rs (which times nothing). */
int rrhip_time_trace_device(rrhip_scene* scene, const rrhip_ray* rays, uint64_t n, int any,
                            void* out, uint32_t reps, float* elapsed_ms);

/* ---- workload generators (deterministic, host) --------------------------- */
/* Heightfield over [0,1]^2 with nx x ny quads (two triangles each),
 * z = amplitude * fbm(frequency*x + o0, frequency*y + o1, o2) using
 * render.rs's hash-gradient Perlin and fBm (src/geometry/displace.rs:36-97)
 * with `octaves`, gain 0.5, lacunarity 2; the offset o = 1000*u, u from
 * PCG32(seed, stream 0x5eed) (render.rs's Perlin has no seed; the offset is
 * its only handle). positions: (nx+1)(ny+1)*3 floats, row-major in y then
 * x; indices: nx*ny*6. */
int rrhip_gen_heightfield(uint32_t nx, uint32_t ny, uint64_t seed, float amplitude,
                          float frequency, uint32_t octaves, float* positions, uint32_t* indices);
/* n query rays: origins uniform in the box centre +/- (inflate/2 + 1/2) *
 * extent per axis, directions uniform on the sphere (z = 1-2u,
 * phi = 2*pi*v), tmin = 1e-9, tmax = +inf. Ray i draws from render.rs's
 * per-(pixel, sample) PCG32 seeding with pixel = i, sample = seed
 * (src/raytracer/pt/sampler.rs:24-32), in the order ox, oy, oz, u, v. */
int rrhip_gen_query_rays(const float lo[3], const float hi[3], float inflate, uint64_t seed,
                         uint64_t n, rrhip_ray* rays);

/* Density grid of render.rs's fBm cloud field (Medium::density_at,
 * src/scene/medium.rs:55-63): rho = clamp((fbm(p) * 0.5 + 0.5 - (1 -
 * coverage)) * sharpness, 0, 1) with fbm(frequency, octaves, gain 0.5,
 * lacunarity 2, offset from PCG32(seed, 0x5eed) as for the heightfield)
 * sampled on n x n x n points spanning [lo, hi] (x fastest). */
int rrhip_gen_density_grid(uint32_t n, uint64_t seed, float frequency, uint32_t octaves, float coverage,
                           float sharpness, const float lo[3], const float hi[3], float* out);

/* Unit icosphere subdivided `levels` times: 10*4^levels + 2 vertices and
 * 20*4^levels triangles. Vertices move along their smooth normals by
 * amplitude * fbm(p) (frequency, octaves, gain 0.5, lacunarity 2, offset
 * from PCG32(seed, 0x5eed) as for the heightfield), then normals are
 * rebuilt area-weighted from the displaced faces — render.rs's
 * displace_mesh and smooth_normals (src/geometry/displace.rs:99-123,
 * src/geometry/subdiv.rs:329-358). */
int rrhip_gen_icosphere(uint32_t levels, uint64_t seed, float amplitude, float frequency,
                        uint32_t octaves, float* positions, uint32_t* indices, float* normals);

/* render.rs's fBm (src/geometry/displace.rs:60-97: sum over `octaves` of
 * amp * perlin(p * freq + offset), amp *= gain, freq *= lacunarity, from
 * frequency and amplitude 1) at n points (xyz, f64), on the host: the
 * value RIB `Displace "noise"` moves tessellated vertices by (times its
 * amplitude, along the smooth normals; displace.rs:101-123). Replaces
 * displace_mesh's noise evaluation for callers that dice on the host. */
int rrhip_gen_fbm(uint64_t n, const double* points, double frequency, uint32_t octaves, double gain,
                  double lacunarity, const double offset[3], double* out);

#ifdef __cplusplus
}
#endif

#endif /* RRHIP_H */
