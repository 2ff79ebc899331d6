// Texture coordinates at hits, texture lookups and the pattern graph on
// the device (render.rs src/texture/pattern.rs, cache.rs; st of
// src/geometry/mesh.rs:145-197, 339-370, sphere.rs and extra.rs's
// quadrics; pt/mod.rs:389-440 shade_params with its ray cone).
//
// Textures are whole mip chains in HBM (rgb + pad per texel, 16-byte
// loads), not render.rs's tile cache: a 4096^2 texture with its chain is
// 358 MB, and 288 GB holds every texture a scene names. Lookups run in f32
// over the f32 texels; the fractal's fBm runs in f64 as render.rs's does.
#pragma once
#include "../../include/rrhip.h"
#include "pt_material.h"
#include "pt_trace.h"
#include "sampling.h"

namespace rrhip {
namespace pt {

// ---- texture coordinates at a hit ------------------------------------------------

struct HitSt {
  float s, t, density;
  V3 gs, gt;  // world gradients of s and t (zero where render.rs has none)
};

// extra.rs:35-41: the azimuth as a fraction of the sweep
__device__ __forceinline__ float st_phi_u(float x, float y, float thetamax) {
  float phi = atan2f(y, x) * (180.f / kPi);
  if (phi < 0.f) phi += 360.f;
  return clampf(phi / fmaxf(thetamax, 1e-9f), 0.f, 1.f);
}

// Quadric st and density from the object-space hit point (sphere.rs,
// cylinder.rs, disk.rs, paraboloid.rs, hyperboloid.rs, torus.rs): 1 /
// sqrt(|dP/du| |dP/dv|) over the transform's scale; cones carry none.
__device__ __forceinline__ void quadric_st(const PtQuadric& q, float2 qs, V3 p, HitSt& o) {
  const float rad_theta = q.thetamax * (kPi / 180.f);
  const float scale = qs.x;
  float u = 0.f, v = 0.f, d = 0.f;
  switch (q.kind) {
    case RRHIP_QUADRIC_SPHERE: {
      u = st_phi_u(p.x, p.y, q.thetamax);
      const float zr = fmaxf(q.zmax - q.zmin, 1e-12f);
      v = clampf((p.z - q.zmin) / zr, 0.f, 1.f);
      const float ring = fmaxf(sqrtf(p.x * p.x + p.y * p.y), 1e-6f * fmaxf(q.radius, 1e-12f));
      d = 1.f / sqrtf(fmaxf(ring * rad_theta * (zr * q.radius / ring), 1e-24f)) / scale;
      break;
    }
    case RRHIP_QUADRIC_CYLINDER: {
      u = st_phi_u(p.x, p.y, q.thetamax);
      const float zr = fmaxf(q.zmax - q.zmin, 1e-12f);
      v = clampf((p.z - q.zmin) / zr, 0.f, 1.f);
      d = 1.f / sqrtf(fmaxf(fmaxf(q.radius, 1e-12f) * rad_theta * zr, 1e-24f)) / scale;
      break;
    }
    case RRHIP_QUADRIC_DISK: {
      const float rh = sqrtf(p.x * p.x + p.y * p.y);
      u = st_phi_u(p.x, p.y, q.thetamax);
      v = clampf((q.radius - rh) / fmaxf(q.radius, 1e-12f), 0.f, 1.f);
      const float dpdu = fmaxf(rh, 1e-6f * fmaxf(q.radius, 1e-12f)) * rad_theta;
      d = 1.f / sqrtf(fmaxf(dpdu * q.radius, 1e-24f)) / scale;
      break;
    }
    case RRHIP_QUADRIC_PARABOLOID: {
      u = st_phi_u(p.x, p.y, q.thetamax);
      const float zr = fmaxf(q.zmax - q.zmin, 1e-12f);
      v = clampf((p.z - q.zmin) / zr, 0.f, 1.f);
      const float ring = fmaxf(sqrtf(p.x * p.x + p.y * p.y), 1e-6f * fmaxf(q.radius, 1e-12f));
      d = 1.f / sqrtf(fmaxf(ring * rad_theta * zr, 1e-24f)) / scale;
      break;
    }
    case RRHIP_QUADRIC_HYPERBOLOID: {
      if (fabsf(q.z2 - q.z1) < kObjEps) break;  // the annulus case carries none
      u = st_phi_u(p.x, p.y, q.thetamax);
      v = clampf((p.z - q.z1) / (q.z2 - q.z1), 0.f, 1.f);
      const float ring = fmaxf(sqrtf(p.x * p.x + p.y * p.y), 1e-12f);
      d = 1.f / sqrtf(fmaxf(ring * rad_theta * qs.y, 1e-24f)) / scale;
      break;
    }
    case RRHIP_QUADRIC_TORUS: {
      const float R = q.radius, r = q.height;
      const float rxy = fmaxf(sqrtf(p.x * p.x + p.y * p.y), 1e-12f);
      u = st_phi_u(p.x, p.y, q.thetamax);
      float vphi = atan2f(p.z, rxy - R) * (180.f / kPi);
      while (vphi < q.zmin) vphi += 360.f;
      while (vphi >= q.zmin + 360.f) vphi -= 360.f;
      const float vr = fmaxf(q.zmax - q.zmin, 1e-9f);
      v = clampf((vphi - q.zmin) / vr, 0.f, 1.f);
      d = 1.f / sqrtf(fmaxf(rxy * rad_theta * (fmaxf(r, 1e-12f) * vr * (kPi / 180.f)), 1e-24f)) / scale;
      break;
    }
    default:
      break;
  }
  o.s = u;
  o.t = v;
  o.density = d;
}

// st of the closest hit h (world hit point h.p): meshes with per-vertex st
// interpolate them, with the density sqrt(st area / object area) over the
// instance's scale and world gradients through the instance's inverse
// (mesh.rs:145-197, 357-370); curves give [0.5, v] (mesh.rs:339-342);
// spheres and quadrics their parametric st; loose triangles none. A
// moving curve instance takes v at the ray's shutter time.
template <bool MOTION>
__device__ __forceinline__ HitSt hit_st(const PtSceneView& S, const Hit& h, float time) {
  HitSt o;
  o.s = o.t = o.density = 0.f;
  o.gs = o.gt = v3(0.f, 0.f, 0.f);
  if (h.inst != 0xFFFFFFFFu) {
    const PtInstance& I = S.insts[h.inst];
    if (S.meshes[I.mesh].curves) {
      const CurveRec c = S.csegs[h.slot];
      V3 lp;
      if (!MOTION || I.motion == 0xFFFFFFFFu) {
        lp = xform_point(I.inv, h.p);
      } else {
        float inv[12];
        instance_inverse(S, I, time, inv);
        lp = xform_point(inv, h.p);
      }
      const V3 p0 = v3(c.p0[0], c.p0[1], c.p0[2]), ax = v3(c.p1[0], c.p1[1], c.p1[2]) - p0;
      const float l2 = dot(ax, ax);
      const float along = l2 > 1e-24f ? clampf(dot(lp - p0, ax) / l2, 0.f, 1.f) : 0.5f;
      o.s = 0.5f;
      o.t = c.v0 + (c.v1 - c.v0) * along;
      return o;
    }
    const uint32_t base = S.mesh_st[I.mesh];
    if (base == 0xFFFFFFFFu) return o;
    const TriRec tr = S.btris[h.slot];
    const uint32_t* ix = S.mesh_idx + S.meshes[I.mesh].idx_base + 3 * (size_t)tr.prim;
    const float* st0 = S.st + 2 * ((size_t)base + ix[0]);
    const float* st1 = S.st + 2 * ((size_t)base + ix[1]);
    const float* st2 = S.st + 2 * ((size_t)base + ix[2]);
    const float w = 1.f - h.bu - h.bv;
    o.s = st0[0] * w + st1[0] * h.bu + st2[0] * h.bv;
    o.t = st0[1] * w + st1[1] * h.bu + st2[1] * h.bv;
    const float ds1 = st1[0] - st0[0], dt1 = st1[1] - st0[1];
    const float ds2 = st2[0] - st0[0], dt2 = st2[1] - st0[1];
    const float st_area = 0.5f * fabsf(ds1 * dt2 - ds2 * dt1);
    const V3 v0 = v3(tr.v0x, tr.v0y, tr.v0z);
    const V3 e1 = v3(tr.v1x, tr.v1y, tr.v1z) - v0, e2 = v3(tr.v2x, tr.v2y, tr.v2z) - v0;
    const V3 n = cross(e1, e2);
    const float n2 = dot(n, n);
    const float geo_area = 0.5f * sqrtf(n2);
    o.density = (geo_area > 1e-18f ? sqrtf(st_area / geo_area) : 0.f) / S.inst_scale[h.inst];
    if (n2 >= 1e-24f) {
      // the dual basis of (e1, e2) in their plane: grad = dual1 ds1 + dual2 ds2
      const V3 dual1 = cross(e2, n) / n2, dual2 = cross(e1, n) / -n2;
      o.gs = xform_normal(I.inv, dual1 * ds1 + dual2 * ds2);
      o.gt = xform_normal(I.inv, dual1 * dt1 + dual2 * dt2);
    }
    return o;
  }
  const uint32_t kind = h.obj >> 30, idx = h.obj & kObjIndex;
  if (kind == kObjSphere) {  // a full sphere: thetamax 360, z in [-r, r]
    const PtSphere sp = S.spheres[idx];
    const V3 lp = h.p - v3(sp.c);
    const float zr = fmaxf(2.f * sp.r, 1e-12f);
    const float ring = fmaxf(sqrtf(lp.x * lp.x + lp.y * lp.y), 1e-6f * fmaxf(sp.r, 1e-12f));
    o.s = st_phi_u(lp.x, lp.y, 360.f);
    o.t = clampf((lp.z + sp.r) / zr, 0.f, 1.f);
    o.density = 1.f / sqrtf(fmaxf(ring * 2.f * kPi * (zr * sp.r / ring), 1e-24f));
  } else if (kind == kObjQuadric) {
    const PtQuadric& q = S.quadrics[idx];
    quadric_st(q, S.quad_st[idx], xform_point(q.inv, h.p), o);
  }
  return o;
}

// ---- texture lookups (cache.rs) ------------------------------------------------------

__device__ __forceinline__ V3 tex_texel(const PtSceneView& S, const PtTexLevel& L, int x, int y, uint32_t wrap) {
  const int w = (int)L.w, h = (int)L.h;
  if (wrap == RRHIP_WRAP_PERIODIC) {
    x %= w;
    if (x < 0) x += w;
    y %= h;
    if (y < 0) y += h;
  } else if (wrap == RRHIP_WRAP_CLAMP) {
    x = min(max(x, 0), w - 1);
    y = min(max(y, 0), h - 1);
  } else if (x < 0 || x >= w || y < 0 || y >= h) {
    return v3(0.f, 0.f, 0.f);
  }
  const float4 t = S.texels[L.offset + (size_t)y * L.w + x];
  return v3(t.x, t.y, t.z);
}

// floor of a texel coordinate, kept inside int range
__device__ __forceinline__ int tex_floor(float f) { return (int)floorf(clampf(f, -1e9f, 1e9f)); }

// cache.rs:205-228: four taps around (s w - 0.5, t h - 0.5)
__device__ __forceinline__ V3 tex_bilinear(const PtSceneView& S, const PtTexLevel& L, float s, float t,
                                           uint32_t wrap) {
  const float fx = s * (float)L.w - 0.5f, fy = t * (float)L.h - 0.5f;
  const float x0 = floorf(fx), y0 = floorf(fy);
  const float ax = fx - x0, ay = fy - y0;
  const int ix = tex_floor(fx), iy = tex_floor(fy);
  V3 c = tex_texel(S, L, ix, iy, wrap) * ((1.f - ax) * (1.f - ay));
  c = c + tex_texel(S, L, ix + 1, iy, wrap) * (ax * (1.f - ay));
  c = c + tex_texel(S, L, ix, iy + 1, wrap) * ((1.f - ax) * ay);
  c = c + tex_texel(S, L, ix + 1, iy + 1, wrap) * (ax * ay);
  return c;
}

// cache.rs:230-250: trilinear between the two levels around log2 of the
// footprint in texels
__device__ __forceinline__ V3 tex_sample(const PtSceneView& S, uint32_t tex, float s, float t, float footprint,
                                         uint32_t wrap) {
  const PtTexture T = S.texs[tex];
  const float base = (float)max(T.w, T.h);
  const float texels = fmaxf(footprint * base, 1e-9f);
  const float mip_f = clampf(log2f(texels), 0.f, (float)(T.nlevels - 1));
  const uint32_t m0 = (uint32_t)floorf(mip_f);
  const uint32_t m1 = min(m0 + 1, T.nlevels - 1);
  const float a = mip_f - (float)m0;
  const V3 lo = tex_bilinear(S, S.tlev[T.first_level + m0], s, t, wrap);
  if (a < 1e-6f || m1 == m0) return lo;
  const V3 hi = tex_bilinear(S, S.tlev[T.first_level + m1], s, t, wrap);
  return lo * (1.f - a) + hi * a;
}

// cache.rs:287-378: Heckbert's elliptical Gaussian over the level of the
// minor axis (anisotropy capped at 16, at most 4096 texels, else trilinear)
__device__ __forceinline__ V3 tex_sample_ewa(const PtSceneView& S, uint32_t tex, float s, float t, float dux_,
                                             float duy_, float dvx_, float dvy_, uint32_t wrap) {
  float major = sqrtf(dux_ * dux_ + duy_ * duy_), minor = sqrtf(dvx_ * dvx_ + dvy_ * dvy_);
  float ma0 = dux_, ma1 = duy_, mi0 = dvx_, mi1 = dvy_;
  if (minor > major) {
    const float tm = major;
    major = minor;
    minor = tm;
    ma0 = dvx_;
    ma1 = dvy_;
    mi0 = dux_;
    mi1 = duy_;
  }
  if (major <= 1e-9f) return tex_sample(S, tex, s, t, 0.f, wrap);
  if (minor < major / 16.f) minor = major / 16.f;
  const PtTexture T = S.texs[tex];
  const float base = (float)max(T.w, T.h);
  const float level = clampf(log2f(fmaxf(minor * base, 1e-9f)), 0.f, (float)(T.nlevels - 1));
  const PtTexLevel L = S.tlev[T.first_level + (uint32_t)floorf(level)];
  const float mw = (float)L.w, mh = (float)L.h;
  const float px = s * mw - 0.5f, py = t * mh - 0.5f;
  const float dux = ma0 * mw, duy = ma1 * mh, dvx = mi0 * mw, dvy = mi1 * mh;
  float A = duy * duy + dvy * dvy + 1.f;
  float B = -2.f * (dux * duy + dvx * dvy);
  float C = dux * dux + dvx * dvx + 1.f;
  const float F = A * C - 0.25f * B * B;
  if (F <= 1e-12f) return tex_sample(S, tex, s, t, minor, wrap);
  A /= F;
  B /= F;
  C /= F;
  const float det = -B * B + 4.f * A * C;
  if (det <= 0.f) return tex_sample(S, tex, s, t, minor, wrap);
  const float urad = sqrtf(4.f * C * det) / det, vrad = sqrtf(4.f * A * det) / det;
  const int x0 = tex_floor(px - urad), x1 = (int)ceilf(clampf(px + urad, -1e9f, 1e9f));
  const int y0 = tex_floor(py - vrad), y1 = (int)ceilf(clampf(py + vrad, -1e9f, 1e9f));
  if ((long long)(x1 - x0) * (long long)(y1 - y0) > 4096) return tex_sample(S, tex, s, t, minor, wrap);
  // Row by row over the box, in cache.rs's order; each row visits only
  // the span where A dx^2 + B dx dy + C dy^2 < 1 can hold (the roots of
  // the row's quadratic, widened by a texel each side; A > 0), with the
  // texel column wrapped incrementally. The r2 < 1 test still decides, so
  // the sum takes the same texels in the same order as the whole box.
  V3 sum = v3(0.f, 0.f, 0.f);
  float wsum = 0.f;
  const int w = (int)L.w, h = (int)L.h;
  const float inv2a = 0.5f / A;
  for (int y = y0; y <= y1; y++) {
    const float dy = (float)y - py;
    const float b = B * dy, c = C * dy * dy - 1.f;
    const float disc = b * b - 4.f * A * c;
    if (disc < -1e-6f) continue;
    const float sq = sqrtf(fmaxf(disc, 0.f));
    const int xs = max(x0, tex_floor(px + (-b - sq) * inv2a) - 1);
    const int xe = min(x1, tex_floor(px + (-b + sq) * inv2a) + 2);
    if (xs > xe) continue;
    int yy = y;
    if (wrap == RRHIP_WRAP_PERIODIC) {
      yy %= h;
      if (yy < 0) yy += h;
    } else if (wrap == RRHIP_WRAP_CLAMP) {
      yy = min(max(yy, 0), h - 1);
    } else if (yy < 0 || yy >= h) {
      // black rows: their texels weigh in as zero
      for (int x = xs; x <= xe; x++) {
        const float dx = (float)x - px;
        const float r2 = A * dx * dx + B * dx * dy + C * dy * dy;
        if (r2 < 1.f) wsum += __expf(-2.f * r2);
      }
      continue;
    }
    const float4* row = S.texels + L.offset + (size_t)yy * L.w;
    int xw = xs;
    if (wrap == RRHIP_WRAP_PERIODIC) {
      xw %= w;
      if (xw < 0) xw += w;
    }
    for (int x = xs; x <= xe; x++) {
      const float dx = (float)x - px;
      const float r2 = A * dx * dx + B * dx * dy + C * dy * dy;
      if (r2 < 1.f) {
        const float wt = __expf(-2.f * r2);
        V3 tc = v3(0.f, 0.f, 0.f);
        if (wrap == RRHIP_WRAP_PERIODIC) {
          const float4 t4 = row[xw];
          tc = v3(t4.x, t4.y, t4.z);
        } else if (wrap == RRHIP_WRAP_CLAMP) {
          const float4 t4 = row[min(max(x, 0), w - 1)];
          tc = v3(t4.x, t4.y, t4.z);
        } else if (x >= 0 && x < w) {
          const float4 t4 = row[x];
          tc = v3(t4.x, t4.y, t4.z);
        }
        sum = sum + tc * wt;
        wsum += wt;
      }
      if (++xw == w) xw = 0;
    }
  }
  if (wsum <= 1e-9f) return tex_sample(S, tex, s, t, minor, wrap);
  return sum / wsum;
}

// ---- pattern graph (pattern.rs) --------------------------------------------------------

struct PatCtx {
  float s, t;
  V3 p, n;
  float footprint;
  float maj[2], min[2];  // EWA axes in st
};

__device__ __forceinline__ V3 pat_missing() { return v3(1.f, 0.f, 1.f); }  // pattern.rs:85

// TextureRef::resolve (pattern.rs:24-38): texture and local st, false when
// the node names none (or its UDIM tile holds none)
__device__ __forceinline__ bool pat_resolve(const PtSceneView& S, const PtPattern& n, float s, float t, uint32_t& tex,
                                            float& ls, float& lt) {
  if (n.udim < 0) {
    if (n.texture < 0) return false;
    tex = (uint32_t)n.texture;
    ls = s;
    lt = t;
    return true;
  }
  const float sc = clampf(floorf(s), 0.f, 9.f);
  const float tc = fmaxf(floorf(t), 0.f);
  if (tc > 9.f) return false;  // past tile 1100
  const int id = S.udim[n.udim + (int)sc + 10 * (int)tc];
  if (id < 0) return false;
  tex = (uint32_t)id;
  ls = s - sc;
  lt = t - tc;
  return true;
}

// st has t up, texture rows run down
__device__ __forceinline__ V3 pat_texture(const PtSceneView& S, const PtPattern& n, float s, float t,
                                          float footprint) {
  uint32_t tex;
  float ls, lt;
  if (!pat_resolve(S, n, s, t, tex, ls, lt)) return pat_missing();
  return tex_sample(S, tex, ls, 1.f - lt, footprint, n.wrap);
}

// The leaf kinds: texture, checker, fractal, ramp, triplanar.
__device__ __forceinline__ V3 pat_leaf(const PtSceneView& S, const PtPattern& n, const PatCtx& c) {
  switch (n.kind) {
    case RRHIP_PATTERN_TEXTURE: {
      const float s = c.s * n.scale[0], t = c.t * n.scale[1];
      const float aniso = fabsf(c.maj[0]) + fabsf(c.maj[1]) + fabsf(c.min[0]) + fabsf(c.min[1]);
      if (aniso > 1e-12f) {
        uint32_t tex;
        float ls, lt;
        if (!pat_resolve(S, n, s, t, tex, ls, lt)) return pat_missing();
        return tex_sample_ewa(S, tex, ls, 1.f - lt, c.maj[0] * n.scale[0], -(c.maj[1] * n.scale[1]),
                              c.min[0] * n.scale[0], -(c.min[1] * n.scale[1]), n.wrap);
      }
      return pat_texture(S, n, s, t, c.footprint * fmaxf(fabsf(n.scale[0]), fabsf(n.scale[1])));
    }
    case RRHIP_PATTERN_CHECKER: {  // filtered toward the mean as the footprint reaches a cell
      const float fs = c.s * n.scale[0], ft = c.t * n.scale[1];
      const long long sum = (long long)(floorf(fs) + floorf(ft));
      const V3 a = v3(n.color_a), b = v3(n.color_b);
      const V3 sharp = (((sum % 2) + 2) % 2) == 0 ? a : b;
      const float w = fminf(c.footprint * fmaxf(fabsf(n.scale[0]), fabsf(n.scale[1])), 1.f);
      return sharp * (1.f - w) + (a + b) * (0.5f * w);
    }
    case RRHIP_PATTERN_FRACTAL: {
      const double off[3] = {0.0, 0.0, 0.0};
      const double f = fbm3((double)c.p.x, (double)c.p.y, (double)c.p.z, (double)n.frequency, n.octaves,
                            (double)n.gain, (double)n.lacunarity, off);
      const float v = (float)fmin(fmax(f * 0.5 + 0.5, 0.0), 1.0);
      return v3(v, v, v);
    }
    case RRHIP_PATTERN_RAMP: {
      const float x = clampf(n.use_t ? c.t : c.s, 0.f, 1.f);
      if (n.nstops == 0) return pat_missing();
      float4 prev = S.ramp[n.ramp];
      if (x <= prev.x) return v3(prev.y, prev.z, prev.w);
      for (uint32_t i = 1; i < n.nstops; i++) {
        const float4 cur = S.ramp[n.ramp + i];
        if (x <= cur.x) {
          const float a = (x - prev.x) / fmaxf(cur.x - prev.x, 1e-9f);
          return v3(prev.y, prev.z, prev.w) * (1.f - a) + v3(cur.y, cur.z, cur.w) * a;
        }
        prev = cur;
      }
      return v3(prev.y, prev.z, prev.w);
    }
    case RRHIP_PATTERN_TRIPLANAR: {
      const V3 an = vabs(c.n);
      const float total = fmaxf(an.x + an.y + an.z, 1e-9f);
      const float f = n.frequency;
      const V3 cx = pat_texture(S, n, c.p.y * f, c.p.z * f, c.footprint);
      const V3 cy = pat_texture(S, n, c.p.z * f, c.p.x * f, c.footprint);
      const V3 cz = pat_texture(S, n, c.p.x * f, c.p.y * f, c.footprint);
      return cx * (an.x / total) + cy * (an.y / total) + cz * (an.z / total);
    }
    default:
      return pat_missing();
  }
}

// One binding's program: its node tree in post-order over a value stack.
__device__ __forceinline__ V3 pat_run(const PtSceneView& S, const PtBinding& b, const PatCtx& c) {
  V3 stk[kPatStack];
  int sp = 0;
  for (uint32_t i = 0; i < b.len; i++) {
    const PtPatOp op = S.prog[b.first + i];
    if (op.node < 0) {
      stk[sp++] = v3(op.c[0], op.c[1], op.c[2]);
      continue;
    }
    const PtPattern& n = S.pats[op.node];
    if (n.kind == RRHIP_PATTERN_MIX) {  // lerp(color1, color2, clamp(mix.r))
      const float m = clampf(stk[sp - 1].x, 0.f, 1.f);
      const V3 r = stk[sp - 3] * (1.f - m) + stk[sp - 2] * m;
      sp -= 2;
      stk[sp - 1] = r;
    } else if (n.kind == RRHIP_PATTERN_COLOR_CORRECT) {  // gain, offset, gamma, then saturation
      const V3 in = stk[sp - 1];
      const float g = 1.f / fmaxf(n.gamma, 1e-6f);
      V3 o = v3(powf(fmaxf(in.x * n.cc_gain[0] + n.cc_offset[0], 0.f), g),
                powf(fmaxf(in.y * n.cc_gain[1] + n.cc_offset[1], 0.f), g),
                powf(fmaxf(in.z * n.cc_gain[2] + n.cc_offset[2], 0.f), g));
      if (fabsf(n.saturation - 1.f) > 1e-9f) {
        const float l = lum(o);
        o = v3(l, l, l) + (o - v3(l, l, l)) * n.saturation;
      }
      stk[sp - 1] = o;
    } else {
      stk[sp++] = pat_leaf(S, n, c);
    }
  }
  return stk[0];
}

// Material `mat`'s pattern-bound fields evaluated at a hit, written into
// its PbrParams (BoundField::apply, pattern.rs:274-290), and the lobe
// constants derived again from them.
__device__ __forceinline__ void pat_resolve_material(const PtSceneView& S, uint32_t mat, uint32_t bind,
                                                     const PatCtx& c, PtMaterial& m) {
  RawPbr<float> r = S.raw_pbr[mat];
  float presence = m.presence;
  const uint32_t first = bind >> 4, count = bind & 15u;
  for (uint32_t k = 0; k < count; k++) {
    const PtBinding b = S.binds[first + k];
    const V3 v = pat_run(S, b, c);
    switch (b.field) {
      case RRHIP_BIND_DIFFUSE_COLOR: r.diffuse_color[0] = v.x; r.diffuse_color[1] = v.y; r.diffuse_color[2] = v.z; break;
      case RRHIP_BIND_DIFFUSE_GAIN: r.diffuse_gain = v.x; break;
      case RRHIP_BIND_DIFFUSE_ROUGHNESS: r.diffuse_roughness = v.x; break;
      case RRHIP_BIND_SPECULAR_FACE_COLOR:
        r.specular_face_color[0] = v.x; r.specular_face_color[1] = v.y; r.specular_face_color[2] = v.z; break;
      case RRHIP_BIND_SPECULAR_EDGE_COLOR:
        r.specular_edge_color[0] = v.x; r.specular_edge_color[1] = v.y; r.specular_edge_color[2] = v.z; break;
      case RRHIP_BIND_SPECULAR_ROUGHNESS: r.specular_roughness = clampf(v.x, 0.005f, 1.f); break;
      case RRHIP_BIND_CLEARCOAT_GAIN: r.clearcoat_gain = v.x; break;
      case RRHIP_BIND_FUZZ_GAIN: r.fuzz_gain = v.x; break;
      case RRHIP_BIND_FUZZ_COLOR: r.fuzz_color[0] = v.x; r.fuzz_color[1] = v.y; r.fuzz_color[2] = v.z; break;
      case RRHIP_BIND_GLASS_GAIN: r.glass_gain = v.x; break;
      case RRHIP_BIND_GLASS_ROUGHNESS: r.glass_roughness = clampf(v.x, 0.005f, 1.f); break;
      case RRHIP_BIND_REFRACTION_COLOR:
        r.refraction_color[0] = v.x; r.refraction_color[1] = v.y; r.refraction_color[2] = v.z; break;
      case RRHIP_BIND_GLOW: r.glow[0] = v.x; r.glow[1] = v.y; r.glow[2] = v.z; break;
      case RRHIP_BIND_PRESENCE: presence = clampf(v.x, 0.f, 1.f); break;
      default: break;
    }
  }
  lobe_constants(r, m);
  m.presence = presence;
}

// A blocker's presence for shadow rays: the material's, or its bound
// pattern's at the blocker with no footprint (pt/mod.rs:1294).
__device__ __forceinline__ float pat_presence(const PtSceneView& S, uint32_t bind, const PatCtx& c, float presence) {
  const uint32_t first = bind >> 4, count = bind & 15u;
  for (uint32_t k = 0; k < count; k++) {
    const PtBinding b = S.binds[first + k];
    if (b.field == RRHIP_BIND_PRESENCE) presence = clampf(pat_run(S, b, c).x, 0.f, 1.f);
  }
  return presence;
}

// shade_params (pt/mod.rs:389-440): the shading context of a hit from its
// st record, the ray direction and the ray cone's width — the cone's disk
// projected on the surface as an ellipse (minor axis the width, major axis
// stretched by 1 / |cos| along the view direction), mapped to st by the
// st gradients for EWA; the isotropic footprint is width times density.
__device__ __forceinline__ PatCtx pat_ctx(float4 st4, float4 gs4, float4 gt4, V3 p, V3 n, V3 rd, float width) {
  PatCtx c;
  c.s = st4.x;
  c.t = st4.y;
  c.p = p;
  c.n = n;
  c.footprint = width * st4.z;
  c.maj[0] = c.maj[1] = c.min[0] = c.min[1] = 0.f;
  const V3 gs = v3(gs4.x, gs4.y, gs4.z), gt = v3(gt4.x, gt4.y, gt4.z);
  if (dot(gs, gs) + dot(gt, gt) > 1e-18f && width > 0.f) {
    const V3 d = normalize(rd);
    const float cs = fmaxf(fabsf(dot(d, n)), 1.f / 16.f);
    V3 t1 = d - n * dot(d, n);
    if (dot(t1, t1) < 1e-12f) t1 = fabsf(n.x) < 0.9f ? cross(v3(1.f, 0.f, 0.f), n) : cross(v3(0.f, 1.f, 0.f), n);
    t1 = normalize(t1) * (width / cs);
    const V3 t2 = normalize(cross(n, t1)) * width;
    c.maj[0] = dot(gs, t1);
    c.maj[1] = dot(gt, t1);
    c.min[0] = dot(gs, t2);
    c.min[1] = dot(gt, t2);
  }
  return c;
}

}  // namespace pt
}  // namespace rrhip
