// ORACLE — test infrastructure only (see orc_math.hpp).
// f64 neural radiance cache: the specification of the HIP cache kernels.
#include "orc_nrc.hpp"

#include <cmath>
#include <cstring>

namespace orc {
namespace nrc {

int level_res(const Config& c, int level) {
  if (c.levels <= 1) return (int)std::floor(c.base_res + 1e-6);
  const double b = std::exp((std::log(c.max_res) - std::log(c.base_res)) / (double)(c.levels - 1));
  return (int)std::floor(c.base_res * std::pow(b, (double)level) + 1e-6);
}

uint64_t table_size(const Config& c) { return 1ull << c.log2_table; }

// Dense indexing while the level's (N + 1)^3 corners fit in the table,
// else the spatial hash of Teschner et al. 2003 as used by instant-ngp
// (Mueller et al. 2022, eq. 4): (x * 1) ^ (y * 2654435761) ^ (z * 805459861)
// in 32-bit arithmetic, modulo the table size.
uint32_t corner_index(const Config& c, int level, uint32_t x, uint32_t y, uint32_t z) {
  const uint64_t n1 = (uint64_t)level_res(c, level) + 1;
  const uint64_t T = table_size(c);
  if (n1 * n1 * n1 <= T) return (uint32_t)(x + n1 * (y + n1 * z));
  const uint32_t h = (x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u);
  return h & (uint32_t)(T - 1);
}

double bf16_round(double x) {
  const float f = (float)x;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return (double)r;
}

void make_input(const Config& c, const float* table, const Query& q, double* x, Corners* corners) {
  for (int i = 0; i < kIn; i++) x[i] = 0.0;
  const uint64_t T = table_size(c);
  float pn[3];
  for (int a = 0; a < 3; a++) {
    // f32 normalisation and scaling, as the kernels do it
    const float inv = 1.f / (c.hi[a] - c.lo[a]);
    float v = (q.pos[a] - c.lo[a]) * inv;
    pn[a] = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
  }
  for (int l = 0; l < c.levels; l++) {
    const int N = level_res(c, l);
    uint32_t i0[3];
    double fr[3];
    for (int a = 0; a < 3; a++) {
      const float s = pn[a] * (float)N;
      float fl = std::floor(s);
      if (fl > (float)(N - 1)) fl = (float)(N - 1);
      i0[a] = (uint32_t)fl;
      fr[a] = (double)(s - fl);
    }
    double f0 = 0.0, f1 = 0.0;
    for (int k = 0; k < 8; k++) {
      const uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
      const double w = (dx ? fr[0] : 1.0 - fr[0]) * (dy ? fr[1] : 1.0 - fr[1]) * (dz ? fr[2] : 1.0 - fr[2]);
      const uint32_t id = corner_index(c, l, i0[0] + dx, i0[1] + dy, i0[2] + dz);
      const float* e = table + 2 * ((uint64_t)l * T + id);
      f0 += w * (double)e[0];
      f1 += w * (double)e[1];
      if (corners) {
        corners->idx[l][k] = id;
        corners->w[l][k] = w;
      }
    }
    x[2 * l] = f0;
    x[2 * l + 1] = f1;
  }
  x[32] = q.dir[0];
  x[33] = q.dir[1];
  x[34] = q.dir[2];
  x[35] = q.roughness;
  x[36] = q.normal[0];
  x[37] = q.normal[1];
  x[38] = q.normal[2];
  x[39] = 1.0;
  x[40] = q.albedo[0];
  x[41] = q.albedo[1];
  x[42] = q.albedo[2];
}

namespace {

inline double rd(double v, bool mimic) { return mimic ? bf16_round(v) : v; }

// out[r] = sum_k W[r][k] in[k] for a row-major rows x cols matrix.
void matvec(const double* W, int rows, int cols, const double* in, bool mimic, double* out) {
  for (int r = 0; r < rows; r++) {
    double s = 0.0;
    for (int k = 0; k < cols; k++) s += rd(W[r * cols + k], mimic) * in[k];
    out[r] = s;
  }
}

// out[k] = sum_r W[r][k] in[r]
void matvec_t(const double* W, int rows, int cols, const double* in, bool mimic, double* out) {
  for (int k = 0; k < cols; k++) out[k] = 0.0;
  for (int r = 0; r < rows; r++)
    for (int k = 0; k < cols; k++) out[k] += rd(W[r * cols + k], mimic) * in[r];
}

}  // namespace

void forward(const double* w, const double* x, bool mimic, Activations* act) {
  for (int i = 0; i < kIn; i++) act->x[i] = rd(x[i], mimic);
  double z[kWidth];
  matvec(w + kOffW[0], kWidth, kIn, act->x, mimic, z);
  for (int i = 0; i < kWidth; i++) act->a[0][i] = rd(z[i] > 0.0 ? z[i] : 0.0, mimic);
  for (int l = 1; l < kHidden; l++) {
    matvec(w + kOffW[l], kWidth, kWidth, act->a[l - 1], mimic, z);
    for (int i = 0; i < kWidth; i++) act->a[l][i] = rd(z[i] > 0.0 ? z[i] : 0.0, mimic);
  }
  matvec(w + kOffW[5], kOut, kWidth, act->a[kHidden - 1], mimic, act->y);
}

void backward(const double* w, const Activations& act, const double* dy, bool mimic, double* grad_w,
              double* dx) {
  // Back-propagation through W5 .. W0; the deltas are what the kernels
  // round to bf16 before each product.
  double delta[kWidth], dprev[kWidth];
  const double* a_last = act.a[kHidden - 1];
  if (grad_w)
    for (int r = 0; r < kOut; r++)
      for (int k = 0; k < kWidth; k++) grad_w[kOffW[5] + r * kWidth + k] += dy[r] * a_last[k];
  matvec_t(w + kOffW[5], kOut, kWidth, dy, mimic, dprev);
  for (int i = 0; i < kWidth; i++) delta[i] = rd(a_last[i] > 0.0 ? dprev[i] : 0.0, mimic);
  for (int l = kHidden - 1; l >= 1; l--) {
    const double* ain = act.a[l - 1];
    if (grad_w)
      for (int r = 0; r < kWidth; r++)
        for (int k = 0; k < kWidth; k++) grad_w[kOffW[l] + r * kWidth + k] += delta[r] * ain[k];
    matvec_t(w + kOffW[l], kWidth, kWidth, delta, mimic, dprev);
    for (int i = 0; i < kWidth; i++) delta[i] = rd(ain[i] > 0.0 ? dprev[i] : 0.0, mimic);
  }
  if (grad_w)
    for (int r = 0; r < kWidth; r++)
      for (int k = 0; k < kIn; k++) grad_w[kOffW[0] + r * kIn + k] += delta[r] * act.x[k];
  if (dx) matvec_t(w + kOffW[0], kWidth, kIn, delta, mimic, dx);
}

double loss_and_gradients(const Config& c, const float* table, const double* w, const Query* q,
                          const float* targets, uint64_t n, bool mimic, double* grad_w,
                          double* grad_table) {
  const uint64_t T = table_size(c);
  const double inv = 1.0 / (3.0 * (double)n);
  double loss = 0.0;
  std::vector<double> x(kIn);
  Corners cs;
  Activations act;
  for (uint64_t s = 0; s < n; s++) {
    make_input(c, table, q[s], x.data(), &cs);
    forward(w, x.data(), mimic, &act);
    double dy[kOut];
    for (int k = 0; k < kOut; k++) {
      const double y = act.y[k], t = targets[3 * s + k];
      const double den = y * y + kLossEps;  // stop-gradient denominator
      loss += (y - t) * (y - t) / den * inv;
      dy[k] = rd(2.0 * (y - t) / den * inv, mimic);
    }
    if (!grad_w && !grad_table) continue;
    double dx[kIn];
    backward(w, act, dy, mimic, grad_w, grad_table ? dx : nullptr);
    if (grad_table) {
      for (int l = 0; l < c.levels; l++)
        for (int k = 0; k < 8; k++) {
          double* g = grad_table + 2 * ((uint64_t)l * T + cs.idx[l][k]);
          g[0] += dx[2 * l] * cs.w[l][k];
          g[1] += dx[2 * l + 1] * cs.w[l][k];
        }
    }
  }
  return loss;
}

void adam(double* p, const double* g, double* m, double* v, uint64_t n, uint64_t t, double lr, double b1,
          double b2, double eps) {
  const double c1 = 1.0 - std::pow(b1, (double)t), c2 = 1.0 - std::pow(b2, (double)t);
  for (uint64_t i = 0; i < n; i++) {
    m[i] = b1 * m[i] + (1.0 - b1) * g[i];
    v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
    const double mh = m[i] / c1, vh = v[i] / c2;
    p[i] -= lr * mh / (std::sqrt(vh) + eps);
  }
}

}  // namespace nrc
}  // namespace orc

// ---------------------------------------------------------------------------
// C API for the Python test harness.
using namespace orc::nrc;

namespace {
Config make_cfg(const double* cfg) {
  // cfg: levels, log2_table, base_res, max_res, lo[3], hi[3]
  Config c;
  c.levels = (int)cfg[0];
  c.log2_table = (int)cfg[1];
  c.base_res = cfg[2];
  c.max_res = cfg[3];
  for (int a = 0; a < 3; a++) {
    c.lo[a] = (float)cfg[4 + a];
    c.hi[a] = (float)cfg[7 + a];
  }
  return c;
}
}  // namespace

extern "C" {

void orc_nrc_level_res(const double* cfg, int* out) {
  const Config c = make_cfg(cfg);
  for (int l = 0; l < c.levels; l++) out[l] = level_res(c, l);
}

uint32_t orc_nrc_corner_index(const double* cfg, int level, uint32_t x, uint32_t y, uint32_t z) {
  return corner_index(make_cfg(cfg), level, x, y, z);
}

// Network inputs (n x 48, natural order) of n queries.
void orc_nrc_inputs(const double* cfg, const float* table, const void* queries, uint64_t n, double* out) {
  const Config c = make_cfg(cfg);
  const Query* q = (const Query*)queries;
  for (uint64_t i = 0; i < n; i++) make_input(c, table, q[i], out + kIn * i, nullptr);
}

// MLP outputs (n x 3) for inputs (n x 48); w is the natural layout.
void orc_nrc_forward(const double* w, const double* x, uint64_t n, int mimic, double* out) {
  Activations act;
  for (uint64_t i = 0; i < n; i++) {
    forward(w, x + kIn * i, mimic != 0, &act);
    for (int k = 0; k < kOut; k++) out[kOut * i + k] = act.y[k];
  }
}

// Cache outputs (n x 3) of n queries.
void orc_nrc_infer(const double* cfg, const float* table, const double* w, const void* queries, uint64_t n,
                   int mimic, double* out) {
  const Config c = make_cfg(cfg);
  const Query* q = (const Query*)queries;
  double x[kIn];
  Activations act;
  for (uint64_t i = 0; i < n; i++) {
    make_input(c, table, q[i], x, nullptr);
    forward(w, x, mimic != 0, &act);
    for (int k = 0; k < kOut; k++) out[kOut * i + k] = act.y[k];
  }
}

// Loss over the batch; gradients accumulated into grad_w (kMlpParams) and
// grad_table (levels * 2^log2_table * 2), either may be null.
double orc_nrc_gradients(const double* cfg, const float* table, const double* w, const void* queries,
                         const float* targets, uint64_t n, int mimic, double* grad_w, double* grad_table) {
  return loss_and_gradients(make_cfg(cfg), table, w, (const Query*)queries, targets, n, mimic != 0, grad_w,
                            grad_table);
}

// Loss of the MLP alone on raw inputs x (n x 48); gradients accumulated
// into grad_w and written to grad_x (n x 48), either may be null.
double orc_nrc_mlp_gradients(const double* w, const double* x, const float* targets, uint64_t n, int mimic,
                             double* grad_w, double* grad_x) {
  const double inv = 1.0 / (3.0 * (double)n);
  double loss = 0.0;
  Activations act;
  for (uint64_t s = 0; s < n; s++) {
    forward(w, x + kIn * s, mimic != 0, &act);
    double dy[kOut];
    for (int k = 0; k < kOut; k++) {
      const double y = act.y[k], t = targets[3 * s + k];
      const double den = y * y + kLossEps;
      loss += (y - t) * (y - t) / den * inv;
      dy[k] = 2.0 * (y - t) / den * inv;
      if (mimic) dy[k] = bf16_round(dy[k]);
    }
    backward(w, act, dy, mimic != 0, grad_w, grad_x ? grad_x + kIn * s : nullptr);
  }
  return loss;
}

void orc_nrc_adam(double* p, const double* g, double* m, double* v, uint64_t n, uint64_t t, double lr, double b1,
                  double b2, double eps) {
  adam(p, g, m, v, n, t, lr, b1, b2, eps);
}

double orc_bf16_round(double x) { return bf16_round(x); }

}  // extern "C"
