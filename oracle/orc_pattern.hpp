// ORACLE — test infrastructure only (see orc_math.hpp header).
//
// f64 restatement of render.rs's pattern graph and texture lookups:
// - src/texture/pattern.rs: the node kinds (texture, checker, fractal, mix,
//   color correct, ramp, triplanar), their evaluation at a shading point,
//   UDIM tile picks and the PbrParams fields a pattern drives;
// - src/texture/cache.rs: texel wrap modes, bilinear, trilinear (mip from
//   the footprint) and EWA (Heckbert's elliptical Gaussian, anisotropy
//   capped at 16) lookups; texels are f32 as the .tex files hold them;
// - src/texture/tex.rs: the 2x2 box downsample of the mip chain.
#pragma once
#include <cstdint>
#include <vector>

#include "orc_math.hpp"

namespace orc {

enum class Wrap { Periodic = 0, Clamp = 1, Black = 2 };  // cache.rs:20-35

// One texture's mip chain, linear RGB f32, rows from the top (tex.rs).
struct Texture {
  struct Level {
    uint32_t w = 0, h = 0;
    std::vector<float> rgb;
  };
  std::vector<Level> levels;
  // tex.rs:83-108: every level from the previous by the 2x2 box filter
  // (odd sizes clamp the second tap) down to 1x1.
  void build_chain(uint32_t w, uint32_t h, const float* rgb);
  void texel(size_t mip, int64_t x, int64_t y, Wrap wrap, float out[3]) const;        // cache.rs:186-203
  void bilinear(size_t mip, double s, double t, Wrap wrap, float out[3]) const;      // cache.rs:205-228
  void sample(double s, double t, double footprint, Wrap wrap, float out[3]) const;  // cache.rs:230-250
  void sample_ewa(double s, double t, const double du[2], const double dv[2], Wrap wrap,
                  float out[3]) const;  // cache.rs:287-378
};

enum PatternKind {
  PAT_TEXTURE = 0,
  PAT_CHECKER = 1,
  PAT_FRACTAL = 2,
  PAT_MIX = 3,
  PAT_COLOR_CORRECT = 4,
  PAT_RAMP = 5,
  PAT_TRIPLANAR = 6
};

// pattern.rs:42-68. Inputs: node index or -1 for the constant beside it.
struct PatternNode {
  int kind = PAT_CHECKER;
  int texture = -1;               // texture / triplanar: index, -1 = missing (magenta)
  std::vector<int> udim;          // 100 entries (tiles 1001..1100), empty = a single texture
  Wrap wrap = Wrap::Periodic;
  double scale[2] = {1, 1};       // texture scaleS / scaleT, checker sScale / tScale
  Vec3 color_a, color_b;          // checker
  double frequency = 1.0;         // fractal, triplanar
  uint32_t octaves = 4;
  double gain = 0.5, lacunarity = 2.0;
  int input[3] = {-1, -1, -1};    // mix: color1, color2, mix; color correct: input
  Vec3 constant[3];               // the inputs' constants
  Vec3 cc_gain{1, 1, 1}, cc_offset;
  double gamma = 1.0, saturation = 1.0;
  std::vector<double> positions;  // ramp
  std::vector<Vec3> colors;
  bool use_t = false;
};

// pattern.rs:70-83
struct ShadeCtx {
  double st[2] = {0, 0};
  Vec3 p, n;
  double footprint = 0.0;
  double dst_major[2] = {0, 0}, dst_minor[2] = {0, 0};
};

// pattern.rs:227-257 in rrhip's numbering (RRHIP_BIND_*)
enum BoundField {
  BIND_DIFFUSE_COLOR = 0,
  BIND_DIFFUSE_GAIN,
  BIND_DIFFUSE_ROUGHNESS,
  BIND_SPECULAR_FACE_COLOR,
  BIND_SPECULAR_EDGE_COLOR,
  BIND_SPECULAR_ROUGHNESS,
  BIND_CLEARCOAT_GAIN,
  BIND_FUZZ_GAIN,
  BIND_FUZZ_COLOR,
  BIND_GLASS_GAIN,
  BIND_GLASS_ROUGHNESS,
  BIND_REFRACTION_COLOR,
  BIND_GLOW,
  BIND_PRESENCE,
  BIND_COUNT
};

// pattern.rs:203-212 (cycle guard at depth 16 as well)
Vec3 pattern_eval(const std::vector<PatternNode>& nodes, const std::vector<Texture>& textures, int index,
                  const ShadeCtx& ctx);

}  // namespace orc
