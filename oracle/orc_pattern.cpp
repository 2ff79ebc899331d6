// ORACLE — test infrastructure only (see orc_math.hpp header).
//
// Pattern graph and texture lookups of render.rs (src/texture/pattern.rs,
// cache.rs, tex.rs) in f64 over f32 texels; declarations in orc_pattern.hpp.
#include "orc_pattern.hpp"

#include <algorithm>
#include <cmath>

#include "orc_sampling.hpp"

namespace orc {

void Texture::build_chain(uint32_t w, uint32_t h, const float* rgb) {
  levels.clear();
  Level base;
  base.w = w;
  base.h = h;
  base.rgb.assign(rgb, rgb + 3 * (size_t)w * h);
  levels.push_back(std::move(base));
  while (levels.back().w > 1 || levels.back().h > 1) {
    const Level& a = levels.back();
    Level b;
    b.w = std::max(a.w / 2, 1u);
    b.h = std::max(a.h / 2, 1u);
    b.rgb.assign(3 * (size_t)b.w * b.h, 0.f);
    for (uint32_t y = 0; y < b.h; y++) {
      const uint32_t y0 = std::min(2 * y, a.h - 1), y1 = std::min(2 * y + 1, a.h - 1);
      for (uint32_t x = 0; x < b.w; x++) {
        const uint32_t x0 = std::min(2 * x, a.w - 1), x1 = std::min(2 * x + 1, a.w - 1);
        for (int c = 0; c < 3; c++) {
          // f32 sum in tex.rs's order, then * 0.25
          const float s = a.rgb[3 * ((size_t)y0 * a.w + x0) + c] + a.rgb[3 * ((size_t)y0 * a.w + x1) + c] +
                          a.rgb[3 * ((size_t)y1 * a.w + x0) + c] + a.rgb[3 * ((size_t)y1 * a.w + x1) + c];
          b.rgb[3 * ((size_t)y * b.w + x) + c] = s * 0.25f;
        }
      }
    }
    levels.push_back(std::move(b));
  }
}

void Texture::texel(size_t mip, int64_t x, int64_t y, Wrap wrap, float out[3]) const {
  const Level& L = levels[mip];
  const int64_t w = L.w, h = L.h;
  switch (wrap) {
    case Wrap::Periodic:
      x = ((x % w) + w) % w;
      y = ((y % h) + h) % h;
      break;
    case Wrap::Clamp:
      x = std::clamp<int64_t>(x, 0, w - 1);
      y = std::clamp<int64_t>(y, 0, h - 1);
      break;
    case Wrap::Black:
      if (x < 0 || x >= w || y < 0 || y >= h) {
        out[0] = out[1] = out[2] = 0.f;
        return;
      }
      break;
  }
  const float* p = L.rgb.data() + 3 * ((size_t)y * L.w + x);
  out[0] = p[0];
  out[1] = p[1];
  out[2] = p[2];
}

void Texture::bilinear(size_t mip, double s, double t, Wrap wrap, float out[3]) const {
  const Level& L = levels[mip];
  const double fx = s * L.w - 0.5, fy = t * L.h - 0.5;
  const double x0 = std::floor(fx), y0 = std::floor(fy);
  const double ax = fx - x0, ay = fy - y0;
  const int64_t ix = (int64_t)x0, iy = (int64_t)y0;
  const struct {
    int dx, dy;
    double w;
  } taps[4] = {{0, 0, (1.0 - ax) * (1.0 - ay)}, {1, 0, ax * (1.0 - ay)}, {0, 1, (1.0 - ax) * ay}, {1, 1, ax * ay}};
  out[0] = out[1] = out[2] = 0.f;
  for (const auto& tp : taps) {
    float px[3];
    texel(mip, ix + tp.dx, iy + tp.dy, wrap, px);
    for (int c = 0; c < 3; c++) out[c] += px[c] * (float)tp.w;
  }
}

void Texture::sample(double s, double t, double footprint, Wrap wrap, float out[3]) const {
  const double base = (double)std::max(levels[0].w, levels[0].h);
  const double texels = std::fmax(footprint * base, 1e-9);
  const double mip_f = std::clamp(std::log2(texels), 0.0, (double)(levels.size() - 1));
  const size_t m0 = (size_t)std::floor(mip_f);
  const size_t m1 = std::min(m0 + 1, levels.size() - 1);
  const double a = mip_f - (double)m0;
  bilinear(m0, s, t, wrap, out);
  if (a < 1e-6 || m1 == m0) return;
  float hi[3];
  bilinear(m1, s, t, wrap, hi);
  for (int c = 0; c < 3; c++) out[c] = out[c] * (float)(1.0 - a) + hi[c] * (float)a;
}

void Texture::sample_ewa(double s, double t, const double du[2], const double dv[2], Wrap wrap,
                         float out[3]) const {
  double major = std::sqrt(du[0] * du[0] + du[1] * du[1]);
  double minor = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1]);
  double amaj[2] = {du[0], du[1]}, amin[2] = {dv[0], dv[1]};
  if (minor > major) {
    std::swap(major, minor);
    std::swap(amaj[0], amin[0]);
    std::swap(amaj[1], amin[1]);
  }
  if (major <= 1e-9) return sample(s, t, 0.0, wrap, out);
  if (minor < major / 16.0) minor = major / 16.0;  // anisotropy cap
  const double base = (double)std::max(levels[0].w, levels[0].h);
  const double level = std::clamp(std::log2(std::fmax(minor * base, 1e-9)), 0.0, (double)(levels.size() - 1));
  const size_t mip = (size_t)std::floor(level);
  const Level& L = levels[mip];
  const double mw = L.w, mh = L.h;
  const double px = s * mw - 0.5, py = t * mh - 0.5;
  const double dux = amaj[0] * mw, duy = amaj[1] * mh, dvx = amin[0] * mw, dvy = amin[1] * mh;
  // Heckbert's implicit ellipse A x^2 + B xy + C y^2 = F, normalised to F = 1
  double A = duy * duy + dvy * dvy + 1.0;
  double B = -2.0 * (dux * duy + dvx * dvy);
  double C = dux * dux + dvx * dvx + 1.0;
  const double F = A * C - 0.25 * B * B;
  if (F <= 1e-12) return sample(s, t, minor, wrap, out);
  A /= F;
  B /= F;
  C /= F;
  const double det = -B * B + 4.0 * A * C;
  if (det <= 0.0) return sample(s, t, minor, wrap, out);
  const double urad = std::sqrt(4.0 * C * det) / det, vrad = std::sqrt(4.0 * A * det) / det;
  const int64_t x0 = (int64_t)std::floor(px - urad), x1 = (int64_t)std::ceil(px + urad);
  const int64_t y0 = (int64_t)std::floor(py - vrad), y1 = (int64_t)std::ceil(py + vrad);
  if ((x1 - x0) * (y1 - y0) > 4096) return sample(s, t, minor, wrap, out);
  float sum[3] = {0.f, 0.f, 0.f}, wsum = 0.f;
  for (int64_t y = y0; y <= y1; y++) {
    const double dy = (double)y - py;
    for (int64_t x = x0; x <= x1; x++) {
      const double dx = (double)x - px;
      const double r2 = A * dx * dx + B * dx * dy + C * dy * dy;
      if (r2 < 1.0) {
        const float w = (float)std::exp(-2.0 * r2);
        float c3[3];
        texel(mip, x, y, wrap, c3);
        for (int k = 0; k < 3; k++) sum[k] += c3[k] * w;
        wsum += w;
      }
    }
  }
  if (wsum <= 1e-9f) return sample(s, t, minor, wrap, out);
  for (int k = 0; k < 3; k++) out[k] = sum[k] / wsum;
}

namespace {

const Vec3 kMissing{1.0, 0.0, 1.0};  // pattern.rs:85

// TextureRef::resolve (pattern.rs:24-38): the texture and local st.
bool resolve(const PatternNode& n, double s, double t, int* tex, double* ls, double* lt) {
  if (n.udim.empty()) {
    if (n.texture < 0) return false;
    *tex = n.texture;
    *ls = s;
    *lt = t;
    return true;
  }
  const double sc = std::clamp(std::floor(s), 0.0, 9.0);
  const double tc = std::fmax(std::floor(t), 0.0);
  const int64_t tile = 1001 + (int64_t)sc + 10 * (int64_t)tc;
  if (tile < 1001 || tile > 1100 || n.udim[(size_t)(tile - 1001)] < 0) return false;
  *tex = n.udim[(size_t)(tile - 1001)];
  *ls = s - sc;
  *lt = t - tc;
  return true;
}

Vec3 sample_tex(const PatternNode& n, const std::vector<Texture>& textures, double s, double t, double footprint) {
  int id;
  double ls, lt;
  if (!resolve(n, s, t, &id, &ls, &lt)) return kMissing;
  float c[3];
  textures[(size_t)id].sample(ls, 1.0 - lt, footprint, n.wrap, c);  // st has t up, rows run down
  return {(double)c[0], (double)c[1], (double)c[2]};
}

Vec3 sample_tex_ewa(const PatternNode& n, const std::vector<Texture>& textures, double s, double t, const double du[2],
                    const double dv[2]) {
  int id;
  double ls, lt;
  if (!resolve(n, s, t, &id, &ls, &lt)) return kMissing;
  const double u2[2] = {du[0], -du[1]}, v2[2] = {dv[0], -dv[1]};
  float c[3];
  textures[(size_t)id].sample_ewa(ls, 1.0 - lt, u2, v2, n.wrap, c);
  return {(double)c[0], (double)c[1], (double)c[2]};
}

Vec3 eval_at(const std::vector<PatternNode>& nodes, const std::vector<Texture>& textures, int index,
             const ShadeCtx& ctx, int depth);

Vec3 eval_input(const std::vector<PatternNode>& nodes, const std::vector<Texture>& textures, const PatternNode& n,
                int k, const ShadeCtx& ctx, int depth) {
  return n.input[k] >= 0 ? eval_at(nodes, textures, n.input[k], ctx, depth) : n.constant[k];
}

Vec3 eval_at(const std::vector<PatternNode>& nodes, const std::vector<Texture>& textures, int index,
             const ShadeCtx& ctx, int depth) {
  if (depth > 16 || index < 0 || (size_t)index >= nodes.size()) return kMissing;
  const PatternNode& n = nodes[(size_t)index];
  switch (n.kind) {
    case PAT_TEXTURE: {
      const double s = ctx.st[0] * n.scale[0], t = ctx.st[1] * n.scale[1];
      const double aniso = std::fabs(ctx.dst_major[0]) + std::fabs(ctx.dst_major[1]) + std::fabs(ctx.dst_minor[0]) +
                           std::fabs(ctx.dst_minor[1]);
      if (aniso > 1e-12) {
        const double du[2] = {ctx.dst_major[0] * n.scale[0], ctx.dst_major[1] * n.scale[1]};
        const double dv[2] = {ctx.dst_minor[0] * n.scale[0], ctx.dst_minor[1] * n.scale[1]};
        return sample_tex_ewa(n, textures, s, t, du, dv);
      }
      return sample_tex(n, textures, s, t, ctx.footprint * std::fmax(std::fabs(n.scale[0]), std::fabs(n.scale[1])));
    }
    case PAT_CHECKER: {  // analytically filtered toward the mean as the footprint reaches a cell
      const double fs = ctx.st[0] * n.scale[0], ft = ctx.st[1] * n.scale[1];
      const int64_t sum = (int64_t)(std::floor(fs) + std::floor(ft));
      const Vec3 sharp = ((sum % 2) + 2) % 2 == 0 ? n.color_a : n.color_b;
      const Vec3 mean = (n.color_a + n.color_b) * 0.5;
      const double w = std::fmin(ctx.footprint * std::fmax(std::fabs(n.scale[0]), std::fabs(n.scale[1])), 1.0);
      return sharp * (1.0 - w) + mean * w;
    }
    case PAT_FRACTAL: {
      FbmParams fp;
      fp.frequency = n.frequency;
      fp.octaves = n.octaves;
      fp.gain = n.gain;
      fp.lacunarity = n.lacunarity;
      const double v = std::clamp(fbm(ctx.p.x, ctx.p.y, ctx.p.z, fp) * 0.5 + 0.5, 0.0, 1.0);
      return {v, v, v};
    }
    case PAT_MIX: {
      const Vec3 a = eval_input(nodes, textures, n, 0, ctx, depth + 1);
      const Vec3 b = eval_input(nodes, textures, n, 1, ctx, depth + 1);
      const double m = std::clamp(eval_input(nodes, textures, n, 2, ctx, depth + 1).x, 0.0, 1.0);
      return a * (1.0 - m) + b * m;
    }
    case PAT_COLOR_CORRECT: {
      const Vec3 c = eval_input(nodes, textures, n, 0, ctx, depth + 1);
      const double g = 1.0 / std::fmax(n.gamma, 1e-6);
      Vec3 out{std::pow(std::fmax(c.x * n.cc_gain.x + n.cc_offset.x, 0.0), g),
               std::pow(std::fmax(c.y * n.cc_gain.y + n.cc_offset.y, 0.0), g),
               std::pow(std::fmax(c.z * n.cc_gain.z + n.cc_offset.z, 0.0), g)};
      if (std::fabs(n.saturation - 1.0) > 1e-9) {
        const double l = luminance(out);
        const Vec3 grey{l, l, l};
        out = grey + (out - grey) * n.saturation;
      }
      return out;
    }
    case PAT_RAMP: {
      const double x = std::clamp(n.use_t ? ctx.st[1] : ctx.st[0], 0.0, 1.0);
      if (n.colors.empty()) return kMissing;
      if (x <= n.positions[0]) return n.colors[0];
      for (size_t i = 1; i < n.positions.size() && i < n.colors.size(); i++) {
        if (x <= n.positions[i]) {
          const double span = std::fmax(n.positions[i] - n.positions[i - 1], 1e-9);
          const double a = (x - n.positions[i - 1]) / span;
          return n.colors[i - 1] * (1.0 - a) + n.colors[i] * a;
        }
      }
      return n.colors.back();
    }
    case PAT_TRIPLANAR: {
      const Vec3 an{std::fabs(ctx.n.x), std::fabs(ctx.n.y), std::fabs(ctx.n.z)};
      const double total = std::fmax(an.x + an.y + an.z, 1e-9);
      const Vec3 w{an.x / total, an.y / total, an.z / total};
      const double f = n.frequency, fp = ctx.footprint;
      const Vec3 cx = sample_tex(n, textures, ctx.p.y * f, ctx.p.z * f, fp);
      const Vec3 cy = sample_tex(n, textures, ctx.p.z * f, ctx.p.x * f, fp);
      const Vec3 cz = sample_tex(n, textures, ctx.p.x * f, ctx.p.y * f, fp);
      return cx * w.x + cy * w.y + cz * w.z;
    }
  }
  return kMissing;
}

}  // namespace

Vec3 pattern_eval(const std::vector<PatternNode>& nodes, const std::vector<Texture>& textures, int index,
                  const ShadeCtx& ctx) {
  return eval_at(nodes, textures, index, ctx, 0);
}

}  // namespace orc
