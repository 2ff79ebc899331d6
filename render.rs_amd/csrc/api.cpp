// C-ABI layer: contexts, device buffers, scene upload and ray queries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "api_internal.h"
#include "bvh_build.h"
#include "pt_kernels.h"
#include "pt_scene.h"
#include "scene_internal.h"
#include "trace.h"

namespace rrhip {

static thread_local std::string g_last_error;

int fail(int status, const std::string& msg) {
  g_last_error = msg;
  return status;
}

int fail_hip(hipError_t e, const char* where) {
  std::string m = std::string(where) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
  int status = RRHIP_ERR_DEVICE;
  if (e == hipErrorOutOfMemory) status = RRHIP_ERR_OUT_OF_MEMORY;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) status = RRHIP_ERR_NO_DEVICE;
  return fail(status, m);
}

}  // namespace rrhip

using namespace rrhip;

namespace {

int set_device(rrhip_ctx* ctx) {
  RRHIP_HIP_TRY(hipSetDevice(ctx->device));
  return RRHIP_OK;
}

int ensure(void** p, uint64_t* cap, uint64_t bytes) {
  if (*cap >= bytes) return RRHIP_OK;
  if (*p) RRHIP_HIP_TRY(hipFree(*p));
  *p = nullptr;
  *cap = 0;
  RRHIP_HIP_TRY(hipMalloc(p, bytes));
  *cap = bytes;
  return RRHIP_OK;
}

hipStream_t pick_stream(rrhip_scene* s, void* stream) {
  return stream ? (hipStream_t)stream : s->ctx->stream;
}

int check_scene(rrhip_scene* s, const void* rays, uint64_t n, const void* out) {
  if (!s) return fail(RRHIP_ERR_INVALID_ARGUMENT, "null scene");
  if (!s->has_query) return fail(RRHIP_ERR_UNSUPPORTED, "scene has no ray-query mesh");
  if (n && (!rays || !out)) return fail(RRHIP_ERR_INVALID_ARGUMENT, "null ray or output buffer");
  return RRHIP_OK;
}

// Launch configuration of a query mode (closest and any hit are tuned
// apart: their kernels differ in registers and so in resident waves).
const TraceTuning& tuning_for(const rrhip_scene* s, TraceMode mode) {
  return mode == TraceMode::kAny ? s->tuning_any : s->tuning;
}

// Host-buffer query: copy in, trace, copy out, wait.
int trace_host(rrhip_scene* s, const rrhip_ray* rays, uint64_t n, TraceMode mode, void* out,
               size_t out_elem) {
  if (int st = check_scene(s, rays, n, out)) return st;
  if (n == 0) return RRHIP_OK;
  if (int st = set_device(s->ctx)) return st;
  if (int st = ensure(&s->d_in, &s->in_cap, n * sizeof(rrhip_ray))) return st;
  if (int st = ensure(&s->d_out, &s->out_cap, n * out_elem)) return st;
  hipStream_t st = s->ctx->stream;
  RRHIP_HIP_TRY(hipMemcpyAsync(s->d_in, rays, n * sizeof(rrhip_ray), hipMemcpyHostToDevice, st));
  if (s->bvh.num_nodes == 0) {
    // Empty scene: every ray misses.
    if (mode == TraceMode::kAny) {
      std::memset(out, 0, n);
    } else {
      auto* h = (rrhip_hit*)out;
      for (uint64_t i = 0; i < n; i++) h[i] = rrhip_hit{INFINITY, RRHIP_NO_HIT, 0.f, 0.f};
    }
    RRHIP_HIP_TRY(hipStreamSynchronize(st));
    return RRHIP_OK;
  }
  RRHIP_HIP_TRY(launch_trace(s->bvh, (const RayRec*)s->d_in, n, mode,
                             mode == TraceMode::kClosest ? (HitRec*)s->d_out : nullptr,
                             mode == TraceMode::kAny ? (uint8_t*)s->d_out : nullptr, nullptr, tuning_for(s, mode), st));
  RRHIP_HIP_TRY(hipMemcpyAsync(out, s->d_out, n * out_elem, hipMemcpyDeviceToHost, st));
  RRHIP_HIP_TRY(hipStreamSynchronize(st));
  return RRHIP_OK;
}

// Device-buffer query: enqueue only.
int trace_dev(rrhip_scene* s, const rrhip_ray* rays, uint64_t n, TraceMode mode, void* out,
              void* stream) {
  if (int st = check_scene(s, rays, n, out)) return st;
  if (n == 0) return RRHIP_OK;
  if (int st = set_device(s->ctx)) return st;
  hipStream_t st = pick_stream(s, stream);
  if (s->bvh.num_nodes == 0) {
    if (mode == TraceMode::kAny) {
      RRHIP_HIP_TRY(hipMemsetAsync(out, 0, n, st));
      return RRHIP_OK;
    }
    std::vector<rrhip_hit> miss(n, rrhip_hit{INFINITY, RRHIP_NO_HIT, 0.f, 0.f});
    RRHIP_HIP_TRY(hipMemcpyAsync(out, miss.data(), n * sizeof(rrhip_hit), hipMemcpyHostToDevice, st));
    RRHIP_HIP_TRY(hipStreamSynchronize(st));
    return RRHIP_OK;
  }
  RRHIP_HIP_TRY(launch_trace(s->bvh, (const RayRec*)rays, n, mode,
                             mode == TraceMode::kClosest ? (HitRec*)out : nullptr,
                             mode == TraceMode::kAny ? (uint8_t*)out : nullptr, nullptr, tuning_for(s, mode), st));
  return RRHIP_OK;
}

}  // namespace

extern "C" {

uint32_t rrhip_abi_version(void) { return RRHIP_ABI_VERSION; }

const char* rrhip_last_error(void) { return g_last_error.c_str(); }

int rrhip_init(int device, rrhip_ctx** out) {
  if (!out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_init: null output");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return fail(RRHIP_ERR_NO_DEVICE, "rrhip_init: no HIP device visible");
  if (device < 0 || device >= count)
    return fail(RRHIP_ERR_NO_DEVICE, "rrhip_init: device index out of range");
  hipDeviceProp_t prop;
  RRHIP_HIP_TRY(hipGetDeviceProperties(&prop, device));
  std::string arch = prop.gcnArchName;
  if (arch.rfind("gfx950", 0) != 0)
    return fail(RRHIP_ERR_UNSUPPORTED, "rrhip_init: device is " + arch + ", gfx950 (MI355X) required");
  auto* ctx = new (std::nothrow) rrhip_ctx();
  if (!ctx) return fail(RRHIP_ERR_OUT_OF_MEMORY, "rrhip_init: host allocation failed");
  ctx->device = device;
  ctx->name = prop.name;
  ctx->arch = arch;
  e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
  if (e != hipSuccess) {
    rrhip_destroy(ctx);
    return fail_hip(e, "rrhip_init");
  }
  if (int st = noalloc_setup(device)) {
    rrhip_destroy(ctx);
    return st;
  }
  *out = ctx;
  return RRHIP_OK;
}

int rrhip_destroy(rrhip_ctx* ctx) {
  if (!ctx) return RRHIP_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RRHIP_OK;
}

int rrhip_device_name(rrhip_ctx* ctx, char* name, size_t name_len, char* arch, size_t arch_len) {
  if (!ctx) return fail(RRHIP_ERR_INVALID_ARGUMENT, "null context");
  if (name && name_len) {
    std::strncpy(name, ctx->name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (arch && arch_len) {
    std::strncpy(arch, ctx->arch.c_str(), arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  return RRHIP_OK;
}

void* rrhip_stream(rrhip_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int rrhip_synchronize(rrhip_ctx* ctx) {
  if (!ctx) return fail(RRHIP_ERR_INVALID_ARGUMENT, "null context");
  if (int st = set_device(ctx)) return st;
  RRHIP_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return RRHIP_OK;
}

int rrhip_buffer_alloc(rrhip_ctx* ctx, uint64_t bytes, void** dptr) {
  if (!ctx || !dptr) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_buffer_alloc: null argument");
  *dptr = nullptr;
  if (int st = set_device(ctx)) return st;
  RRHIP_HIP_TRY(hipMalloc(dptr, bytes ? bytes : 1));
  return RRHIP_OK;
}

int rrhip_buffer_free(rrhip_ctx* ctx, void* dptr) {
  if (!ctx) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_buffer_free: null context");
  if (!dptr) return RRHIP_OK;
  if (int st = set_device(ctx)) return st;
  RRHIP_HIP_TRY(hipStreamSynchronize(ctx->stream));
  RRHIP_HIP_TRY(hipFree(dptr));
  return RRHIP_OK;
}

int rrhip_copy_to_device(rrhip_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_copy_to_device: null argument");
  if (int st = set_device(ctx)) return st;
  RRHIP_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  RRHIP_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return RRHIP_OK;
}

int rrhip_copy_to_host(rrhip_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_copy_to_host: null argument");
  if (int st = set_device(ctx)) return st;
  RRHIP_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  RRHIP_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return RRHIP_OK;
}

namespace {

// A scene is a ray-query mesh (exactly one mesh, no instances, no loose
// objects), a path-tracing scene (camera width > 0), or both.
bool is_query(const rrhip_scene_desc* d) {
  return d->num_meshes == 1 && d->num_instances == 0 && d->meshes;
}
bool is_pt(const rrhip_scene_desc* d) { return d->camera.width > 0; }

int validate_desc(const rrhip_scene_desc* desc) {
  if (desc->struct_size != sizeof(rrhip_scene_desc))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "scene desc: struct_size mismatch (ABI version)");
  if (!is_query(desc) && !is_pt(desc))
    return fail(RRHIP_ERR_INVALID_ARGUMENT,
                "scene desc: need exactly one mesh (ray queries) or a camera (path tracing)");
  if (is_query(desc)) {
    const rrhip_mesh_desc& m = desc->meshes[0];
    if (m.num_triangles && (!m.positions || !m.indices))
      return fail(RRHIP_ERR_INVALID_ARGUMENT, "scene desc: null mesh arrays");
  }
  return RRHIP_OK;
}

int build_from_desc(const rrhip_scene_desc* desc, BuiltBvh* built, rrhip_scene_info* info) {
  const rrhip_mesh_desc& m = desc->meshes[0];
  BuildParams bp;
  bp.bins = desc->build.bins ? (int)desc->build.bins : 32;
  bp.max_leaf = desc->build.max_leaf ? (int)desc->build.max_leaf : 4;
  std::string err;
  auto t0 = std::chrono::steady_clock::now();
  if (!build_triangle_bvh(m.positions, m.num_vertices, m.indices, m.num_triangles, bp, built, &err))
    return fail(RRHIP_ERR_INVALID_ARGUMENT, "BVH build: " + err);
  auto t1 = std::chrono::steady_clock::now();
  *info = rrhip_scene_info{};
  info->num_triangles = m.num_triangles;
  info->num_nodes = built->nodes.size();
  info->num_leaves = built->leaves;
  info->device_bytes = built->nodes.size() * sizeof(Node2) + built->tris.size() * sizeof(TriRec);
  info->depth = (uint32_t)(built->nodes.empty() ? 0 : built->depth + 1);
  info->node_bytes = sizeof(Node2);
  info->tri_bytes = sizeof(TriRec);
  info->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  for (int a = 0; a < 3; a++) {
    info->bounds_lo[a] = built->root_lo[a];
    info->bounds_hi[a] = built->root_hi[a];
  }
  return RRHIP_OK;
}

}  // namespace

int rrhip_build_info(const rrhip_scene_desc* desc, rrhip_scene_info* info) {
  if (!desc || !info) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_build_info: null argument");
  if (int st = validate_desc(desc)) return st;
  if (is_query(desc)) {
    BuiltBvh built;
    return build_from_desc(desc, &built, info);
  }
  if (!is_pt(desc)) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_build_info: no ray-query mesh or camera");
  // Path-tracing scene: instance hierarchy statistics; depth = the
  // traversal stack entries a lane can need.
  PtSceneHost h;
  std::string err;
  if (!build_pt_scene(*desc, &h, &err)) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_build_info: " + err);
  *info = rrhip_scene_info{};
  info->num_triangles = h.btris.size() + h.tris.size();
  info->num_nodes = h.nodes.size();
  info->num_leaves = h.insts.size();
  info->depth = h.stack_need;
  info->node_bytes = sizeof(Node2);
  info->tri_bytes = sizeof(TriRec);
  for (int a = 0; a < 3; a++) {
    info->bounds_lo[a] = h.lo[a];
    info->bounds_hi[a] = h.hi[a];
  }
  return RRHIP_OK;
}

}  // extern "C"

namespace {

template <class T>
hipError_t upload_vec(const std::vector<T>& v, void** slot, const T** view) {
  *view = nullptr;
  if (v.empty()) return hipSuccess;
  hipError_t e = hipMalloc(slot, v.size() * sizeof(T));
  if (e == hipSuccess) e = hipMemcpy(*slot, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  *view = (const T*)*slot;
  return e;
}

hipError_t upload_pt(rrhip_scene* s, const PtSceneHost& h) {
  PtSceneView& v = s->pt;
  hipError_t e = upload_vec(h.tris, &s->pt_mem[0], &v.tris);
  if (e == hipSuccess) e = upload_vec(h.spheres, &s->pt_mem[1], &v.spheres);
  if (e == hipSuccess) e = upload_vec(h.materials, &s->pt_mem[2], &v.materials);
  if (e == hipSuccess) e = upload_vec(h.lights, &s->pt_mem[3], &v.lights);
  if (e == hipSuccess) e = upload_vec(h.lnodes, &s->pt_mem[4], &v.lnodes);
  if (e == hipSuccess) e = upload_vec(h.light_leaf, &s->pt_mem[5], &v.light_leaf);
  if (e == hipSuccess) e = upload_vec(h.inf_lights, &s->pt_mem[6], &v.inf_lights);
  if (e == hipSuccess) e = upload_vec(h.inf_cdf, &s->pt_mem[7], &v.inf_cdf);
  if (e == hipSuccess) e = upload_vec(h.nodes, &s->pt_mem[8], &v.nodes);
  if (e == hipSuccess) e = upload_vec(h.btris, &s->pt_mem[9], &v.btris);
  if (e == hipSuccess) e = upload_vec(h.insts, &s->pt_mem[10], &v.insts);
  if (e == hipSuccess) e = upload_vec(h.meshes, &s->pt_mem[11], &v.meshes);
  if (e == hipSuccess) e = upload_vec(h.mesh_idx, &s->pt_mem[12], &v.mesh_idx);
  if (e == hipSuccess) e = upload_vec(h.mesh_nrm, &s->pt_mem[13], &v.mesh_nrm);
  // media: each density grid in its own allocation, then the medium table
  std::vector<PtMedium> media = h.media;
  for (size_t i = 0; i < media.size() && e == hipSuccess; i++) {
    if (h.media_grids[i].empty()) continue;
    void* gp = nullptr;
    const float* view = nullptr;
    e = upload_vec(h.media_grids[i], &gp, &view);
    if (gp) s->media_mem.push_back(gp);
    media[i].grid = view;
  }
  if (e == hipSuccess) e = upload_vec(media, &s->pt_mem[14], &v.media);
  v.atmosphere = h.atmosphere;
  v.stack_need = h.stack_need;
  v.transparent = h.transparent;
  v.motion = h.motion;
  if (e == hipSuccess) e = upload_vec(h.inst_motion, &s->pt_mem[15], &v.inst_motion);
  if (e == hipSuccess) e = upload_vec(h.env_rgb, &s->pt_mem[16], &v.env_rgb);
  if (e == hipSuccess) e = upload_vec(h.env_marg, &s->pt_mem[17], &v.env_marg);
  if (e == hipSuccess) e = upload_vec(h.env_cond, &s->pt_mem[18], &v.env_cond);
  if (e == hipSuccess) e = upload_vec(h.quadrics, &s->pt_mem[19], &v.quadrics);
  if (e == hipSuccess) e = upload_vec(h.inst_fwd, &s->pt_mem[20], &v.inst_fwd);
  if (e == hipSuccess) e = upload_vec(h.csegs, &s->pt_mem[21], &v.csegs);
  if (e == hipSuccess) e = upload_vec(h.btris1, &s->pt_mem[35], &v.btris1);
  if (h.patterns) {  // pattern-bound materials: their data and the hits' texture coordinates
    if (e == hipSuccess) e = upload_vec(h.raw_pbr, &s->pt_mem[22], &v.raw_pbr);
    if (e == hipSuccess) e = upload_vec(h.binds, &s->pt_mem[23], &v.binds);
    if (e == hipSuccess) e = upload_vec(h.prog, &s->pt_mem[24], &v.prog);
    if (e == hipSuccess) e = upload_vec(h.pats, &s->pt_mem[25], &v.pats);
    if (e == hipSuccess) e = upload_vec(h.ramp, &s->pt_mem[26], &v.ramp);
    if (e == hipSuccess) e = upload_vec(h.udim, &s->pt_mem[27], &v.udim);
    if (e == hipSuccess) e = upload_vec(h.texs, &s->pt_mem[28], &v.texs);
    if (e == hipSuccess) e = upload_vec(h.tlev, &s->pt_mem[29], &v.tlev);
    if (e == hipSuccess) e = upload_vec(h.texels, &s->pt_mem[30], &v.texels);
    if (e == hipSuccess) e = upload_vec(h.mesh_st, &s->pt_mem[31], &v.mesh_st);
    if (e == hipSuccess) e = upload_vec(h.st, &s->pt_mem[32], &v.st);
    if (e == hipSuccess) e = upload_vec(h.inst_scale, &s->pt_mem[33], &v.inst_scale);
    if (e == hipSuccess) e = upload_vec(h.quad_st, &s->pt_mem[34], &v.quad_st);
  }
  v.patterns = h.patterns;
  v.pixel_spread = h.pixel_spread;
  v.num_quadrics = (uint32_t)h.quadrics.size();
  v.curves = h.curves;
  v.hair = h.hair;
  v.env_w = h.env_w;
  v.env_h = h.env_h;
  v.env_total = h.env_total;
  v.env_pdf_k = h.env_rgb.empty() ? 0.0
                                  : (double)h.env_w * (double)h.env_h /
                                        (h.env_total * 2.0 * 3.14159265358979323846 * 3.14159265358979323846);
  v.num_insts = (uint32_t)h.insts.size();
  v.num_tris = (uint32_t)h.tris.size();
  v.num_spheres = (uint32_t)h.spheres.size();
  v.num_lights = (uint32_t)h.lights.size();
  v.num_lnodes = (uint32_t)h.lnodes.size();
  v.num_inf = (uint32_t)h.inf_lights.size();
  v.dome = h.dome;
  v.p_infinite = h.p_infinite;
  v.inf_total = h.inf_total;
  for (int c = 0; c < 3; c++) v.background[c] = h.background[c];
  v.cam = h.cam;
  for (int c = 0; c < 3; c++) {
    s->pt_lo[c] = h.lo[c];
    s->pt_hi[c] = h.hi[c];
  }
  s->pt_eq = false;
  if (h.atmosphere >= 0)
    for (const PtLight& l : h.lights) s->pt_eq |= l.kind == kLightPoint;
  return e;
}

}  // namespace

extern "C" {

int rrhip_scene_upload(rrhip_ctx* ctx, const rrhip_scene_desc* desc, rrhip_scene** out) {
  if (!ctx || !desc || !out) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_scene_upload: null argument");
  *out = nullptr;
  if (int st = validate_desc(desc)) return st;
  BuiltBvh built;
  rrhip_scene_info info{};
  if (is_query(desc)) {
    if (int st = build_from_desc(desc, &built, &info)) return st;
  }
  PtSceneHost pth;
  if (is_pt(desc)) {
    std::string err;
    if (!build_pt_scene(*desc, &pth, &err)) {
      const bool unsupported = err.rfind("UNSUPPORTED", 0) == 0;
      return fail(unsupported ? RRHIP_ERR_UNSUPPORTED : RRHIP_ERR_INVALID_ARGUMENT, "rrhip_scene_upload: " + err);
    }
    if (pth.stack_need > kPtTraceStack)
      return fail(RRHIP_ERR_UNSUPPORTED, "rrhip_scene_upload: instance hierarchy too deep for the " +
                                             std::to_string(kPtTraceStack) + "-entry traversal stack");
  }
  if (int st = set_device(ctx)) return st;
  auto* s = new (std::nothrow) rrhip_scene();
  if (!s) return fail(RRHIP_ERR_OUT_OF_MEMORY, "rrhip_scene_upload: host allocation failed");
  s->ctx = ctx;
  s->has_query = is_query(desc);
  hipError_t e = hipSuccess;
  std::vector<Node4> nodes4;
  int depth4 = 0;
  collapse_bvh4(built.nodes, &nodes4, &depth4);
  if (depth4 + 1 > kMaxStack || nodes4.size() > kNode4Mask) nodes4.clear();  // beyond the four-wide stack entries
  if (!nodes4.empty()) {
    e = hipMalloc((void**)&s->d_nodes4, nodes4.size() * sizeof(Node4));
    if (e == hipSuccess)
      e = hipMemcpy(s->d_nodes4, nodes4.data(), nodes4.size() * sizeof(Node4), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && !built.nodes.empty()) {
    e = hipMalloc((void**)&s->d_nodes, built.nodes.size() * sizeof(Node2));
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_tris, built.tris.size() * sizeof(TriRec));
    if (e == hipSuccess)
      e = hipMemcpy(s->d_nodes, built.nodes.data(), built.nodes.size() * sizeof(Node2), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(s->d_tris, built.tris.data(), built.tris.size() * sizeof(TriRec), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && is_pt(desc)) {
    s->has_pt = true;
    e = upload_pt(s, pth);
  }
  if (e != hipSuccess) {
    rrhip_scene_destroy(s);
    return fail_hip(e, "rrhip_scene_upload");
  }
  s->bvh.nodes = s->d_nodes;
  s->bvh.tris = s->d_tris;
  s->bvh.num_nodes = built.nodes.size();
  s->bvh.num_tris = built.tris.size();
  s->bvh.depth = (int)info.depth;
  s->bvh.nodes4 = s->d_nodes4;
  s->bvh.num_nodes4 = nodes4.size();
  s->bvh.depth4 = depth4;
  s->tuning.stack = (uint32_t)std::max(16, stack_for_depth((int)info.depth));
  s->tuning_any = s->tuning;
  info.depth4 = nodes4.empty() ? 0u : (uint32_t)(depth4 + 1);
  info.device_bytes += nodes4.size() * sizeof(Node4);
  s->info = info;
  *out = s;
  return RRHIP_OK;
}

int rrhip_scene_destroy(rrhip_scene* s) {
  if (!s) return RRHIP_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  if (s->d_nodes) (void)hipFree(s->d_nodes);
  if (s->d_tris) (void)hipFree(s->d_tris);
  if (s->d_nodes4) (void)hipFree(s->d_nodes4);
  if (s->d_in) (void)hipFree(s->d_in);
  if (s->d_out) (void)hipFree(s->d_out);
  for (void* p : s->media_mem) (void)hipFree(p);
  for (void* p : s->pt_mem)
    if (p) (void)hipFree(p);
  delete s;
  return RRHIP_OK;
}

int rrhip_scene_get_info(rrhip_scene* s, rrhip_scene_info* info) {
  if (!s || !info) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_scene_get_info: null argument");
  *info = s->info;
  return RRHIP_OK;
}

int rrhip_scene_get_tuning(rrhip_scene* s, rrhip_trace_tuning* t) {
  if (!s || !t) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_scene_get_tuning: null argument");
  *t = rrhip_trace_tuning{s->tuning.block, s->tuning.stack, s->tuning.chunk, s->tuning.width};
  return RRHIP_OK;
}

int rrhip_scene_set_tuning(rrhip_scene* s, const rrhip_trace_tuning* t) {
  if (!s || !t) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_scene_set_tuning: null argument");
  TraceTuning tu;
  tu.block = t->block;
  tu.stack = t->stack;
  tu.chunk = t->rays_per_wave;
  tu.width = t->width ? t->width : 2;
  if (!tuning_valid(tu, s->bvh))
    return fail(RRHIP_ERR_INVALID_ARGUMENT,
                "rrhip_scene_set_tuning: unsupported block/stack/rays_per_wave/width for this scene (stack need " +
                    std::to_string(stack_need(s->bvh, tu.width)) + ")");
  s->tuning = tu;
  s->tuning_any = tu;
  return RRHIP_OK;
}

int rrhip_scene_autotune(rrhip_scene* s, const rrhip_ray* rays, uint64_t n, int any, void* out,
                         rrhip_trace_tuning* best, float* best_ms) {
  if (int st = check_scene(s, rays, n, out)) return st;
  if (n == 0 || s->bvh.num_nodes == 0) {
    const TraceTuning& cur = any ? s->tuning_any : s->tuning;
    if (best) *best = rrhip_trace_tuning{cur.block, cur.stack, cur.chunk, cur.width};
    if (best_ms) *best_ms = 0.f;
    return RRHIP_OK;
  }
  if (int st = set_device(s->ctx)) return st;
  hipStream_t st = s->ctx->stream;
  const TraceMode mode = any ? TraceMode::kAny : TraceMode::kClosest;
  HitRec* h = any ? nullptr : (HitRec*)out;
  uint8_t* o = any ? (uint8_t*)out : nullptr;
  TraceTuning chosen = any ? s->tuning_any : s->tuning;
  float chosen_ms = 0.f;
  bool have = false;
  for (uint32_t width : {2u, 4u}) {
   const int s0 = stack_for_depth(stack_need(s->bvh, width));
   if (s0 == 0) continue;
   for (uint32_t block : {64u, 128u, 256u}) {
    for (int stack : {16, 24, 32, 48, 64}) {
      if (stack < s0 || stack > std::max(s0, 32)) continue;  // the smallest fitting sizes only
      TraceTuning base;
      base.block = block;
      base.stack = (uint32_t)stack;
      base.width = width;
      // The fill chunk (one round of resident waves) and a few larger
      // chunks (fewer waves, in case fewer fit than the occupancy query
      // reports), plus fixed sizes.
      const uint32_t fill = fill_chunk(n, trace_resident_waves(base, mode));
      const uint32_t cands[] = {0u, (fill * 9 / 8 + 63) / 64 * 64, (fill * 5 / 4 + 63) / 64 * 64,
                                (fill * 3 / 2 + 63) / 64 * 64, 256u, 1024u};
      for (uint32_t chunk : cands) {
        TraceTuning tu = base;
        tu.chunk = chunk;
        if (!tuning_valid(tu, s->bvh)) continue;
        RRHIP_HIP_TRY(launch_trace(s->bvh, (const RayRec*)rays, n, mode, h, o, nullptr, tu, st));
        RRHIP_HIP_TRY(hipEventRecord(s->ctx->ev0, st));
        const int reps = 3;
        for (int r = 0; r < reps; r++)
          RRHIP_HIP_TRY(launch_trace(s->bvh, (const RayRec*)rays, n, mode, h, o, nullptr, tu, st));
        RRHIP_HIP_TRY(hipEventRecord(s->ctx->ev1, st));
        RRHIP_HIP_TRY(hipEventSynchronize(s->ctx->ev1));
        float ms = 0.f;
        RRHIP_HIP_TRY(hipEventElapsedTime(&ms, s->ctx->ev0, s->ctx->ev1));
        ms /= reps;
        // a later candidate must win by more than 1 % (timing noise), so
        // near-ties keep the earlier, smaller configuration
        if (!have || ms < chosen_ms * 0.99f) {
          chosen = tu;
          chosen_ms = ms;
          have = true;
        }
      }
    }
   }
  }
  (any ? s->tuning_any : s->tuning) = chosen;
  if (best) *best = rrhip_trace_tuning{chosen.block, chosen.stack, chosen.chunk, chosen.width};
  if (best_ms) *best_ms = chosen_ms;
  return RRHIP_OK;
}

int rrhip_trace_closest(rrhip_scene* s, const rrhip_ray* rays, uint64_t n, rrhip_hit* hits) {
  return trace_host(s, rays, n, TraceMode::kClosest, hits, sizeof(rrhip_hit));
}

int rrhip_trace_any(rrhip_scene* s, const rrhip_ray* rays, uint64_t n, uint8_t* occluded) {
  return trace_host(s, rays, n, TraceMode::kAny, occluded, 1);
}

int rrhip_trace_closest_device(rrhip_scene* s, const rrhip_ray* rays, uint64_t n, rrhip_hit* hits,
                               void* stream) {
  return trace_dev(s, rays, n, TraceMode::kClosest, hits, stream);
}

int rrhip_trace_any_device(rrhip_scene* s, const rrhip_ray* rays, uint64_t n, uint8_t* occluded,
                           void* stream) {
  return trace_dev(s, rays, n, TraceMode::kAny, occluded, stream);
}

int rrhip_trace_cost(rrhip_scene* s, const rrhip_ray* rays, uint64_t n, int any, uint64_t out[2]) {
  if (int st = check_scene(s, rays, n, out)) return st;
  out[0] = out[1] = 0;
  if (n == 0 || s->bvh.num_nodes == 0) return RRHIP_OK;
  if (int st = set_device(s->ctx)) return st;
  hipStream_t st = s->ctx->stream;
  const TraceMode mode = any ? TraceMode::kAny : TraceMode::kClosest;
  const uint64_t out_elem = any ? 1 : sizeof(rrhip_hit);
  if (int r = ensure(&s->d_in, &s->in_cap, n * sizeof(rrhip_ray))) return r;
  if (int r = ensure(&s->d_out, &s->out_cap, n * out_elem)) return r;
  uint32_t* d_counts = nullptr;
  RRHIP_HIP_TRY(hipMalloc((void**)&d_counts, n * 2 * sizeof(uint32_t)));
  std::vector<uint32_t> counts(2 * n);
  hipError_t e = hipMemcpyAsync(s->d_in, rays, n * sizeof(rrhip_ray), hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = launch_trace(s->bvh, (const RayRec*)s->d_in, n, mode,
                     any ? nullptr : (HitRec*)s->d_out, any ? (uint8_t*)s->d_out : nullptr, d_counts,
                     tuning_for(s, mode), st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(counts.data(), d_counts, n * 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d_counts);
  if (e != hipSuccess) return fail_hip(e, "rrhip_trace_cost");
  for (uint64_t i = 0; i < n; i++) {
    out[0] += counts[2 * i];
    out[1] += counts[2 * i + 1];
  }
  return RRHIP_OK;
}

int rrhip_time_trace_device(rrhip_scene* s, const rrhip_ray* rays, uint64_t n, int any, void* out,
                            uint32_t reps, float* elapsed_ms) {
  if (int st = check_scene(s, rays, n, out)) return st;
  if (!elapsed_ms || reps == 0) return fail(RRHIP_ERR_INVALID_ARGUMENT, "rrhip_time_trace_device: bad reps/output");
  if (int st = set_device(s->ctx)) return st;
  hipStream_t st = s->ctx->stream;
  const TraceMode mode = any ? TraceMode::kAny : TraceMode::kClosest;
  RRHIP_HIP_TRY(hipEventRecord(s->ctx->ev0, st));
  for (uint32_t r = 0; r < reps; r++) {
    if (n && s->bvh.num_nodes)
      RRHIP_HIP_TRY(launch_trace(s->bvh, (const RayRec*)rays, n, mode,
                                 any ? nullptr : (HitRec*)out, any ? (uint8_t*)out : nullptr, nullptr,
                                 tuning_for(s, mode), st));
  }
  RRHIP_HIP_TRY(hipEventRecord(s->ctx->ev1, st));
  RRHIP_HIP_TRY(hipEventSynchronize(s->ctx->ev1));
  RRHIP_HIP_TRY(hipEventElapsedTime(elapsed_ms, s->ctx->ev0, s->ctx->ev1));
  return RRHIP_OK;
}

}  // extern "C"
