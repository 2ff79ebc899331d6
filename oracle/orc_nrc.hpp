// ORACLE — test infrastructure only (see orc_math.hpp).
//
// f64 restatement of the neural radiance cache: multiresolution hash-grid
// encoding, the 48 -> 64 x 5 (ReLU) -> 3 MLP, the relative-L2 loss, its
// gradients with respect to the weights and the grid, and Adam. render.rs
// has no radiance cache; this is the specification the HIP kernels are
// checked against, pinned by finite differences, by closed forms (a
// trilinear grid reproduces linear functions), by torch autograd golden
// vectors and by the textbook Adam update (Kingma & Ba 2015, Algorithm 1).
//
// `mimic` mode rounds to bfloat16 (round to nearest even, through f32)
// exactly where the GPU kernels feed the matrix cores: network inputs,
// weights, hidden activations and the back-propagated deltas. Exact mode
// rounds nowhere (finite-difference checks).
#pragma once
#include <cstdint>
#include <vector>

namespace orc {
namespace nrc {

constexpr int kIn = 48;      // network input width (43 used, 5 zero)
constexpr int kWidth = 64;   // hidden width
constexpr int kHidden = 5;   // hidden layers
constexpr int kOut = 3;
constexpr int kMaxLevels = 16;
// Natural parameter layout: W0 [64][48], W1..W4 [64][64], W5 [3][64].
constexpr int kOffW[6] = {0, 3072, 7168, 11264, 15360, 19456};
constexpr int kMlpParams = 19456 + 3 * 64;
constexpr double kLossEps = 0.01;  // relative L2: (y - t)^2 / (y^2 + 0.01)

struct Config {
  int levels = 16;
  int log2_table = 19;
  double base_res = 16.0;
  double max_res = 2048.0;
  float lo[3] = {0.f, 0.f, 0.f};
  float hi[3] = {1.f, 1.f, 1.f};
};

// One cache query: world position, outgoing direction, shading normal,
// material roughness and albedo (the 64-byte rrhip_nrc_query record).
struct Query {
  float pos[3], roughness;
  float dir[3], pad0;
  float normal[3], pad1;
  float albedo[3], pad2;
};
static_assert(sizeof(Query) == 64, "query record is 64 bytes");

int level_res(const Config& c, int level);            // cells per axis
uint64_t table_size(const Config& c);                 // entries per level
uint32_t corner_index(const Config& c, int level, uint32_t x, uint32_t y, uint32_t z);

double bf16_round(double x);  // f64 -> f32 -> bf16 (RNE) -> f64

// Corner entries and trilinear weights of one query at every level.
struct Corners {
  uint32_t idx[kMaxLevels][8];
  double w[kMaxLevels][8];
};

// Network input (natural order, kIn values) of a query; fills `corners`
// when given.
void make_input(const Config& c, const float* table, const Query& q, double* x, Corners* corners);

struct Activations {
  double x[kIn];
  double a[kHidden][kWidth];  // post-ReLU hidden activations
  double y[kOut];
};

// y = W5 relu(W4 ... relu(W0 x)).
void forward(const double* w, const double* x, bool mimic, Activations* act);

// Adds the batch-summed weight gradients of one sample to grad_w (when
// given) and writes the input gradient to dx (when given), from the
// output gradient dy.
void backward(const double* w, const Activations& act, const double* dy, bool mimic, double* grad_w,
              double* dx);

// Mean relative-L2 loss over a batch and, when grad_w / grad_table are
// given, its gradients (accumulated into them).
double loss_and_gradients(const Config& c, const float* table, const double* w, const Query* q,
                          const float* targets, uint64_t n, bool mimic, double* grad_w,
                          double* grad_table);

// One Adam step on every parameter (t is the step number after the
// increment, >= 1).
void adam(double* p, const double* g, double* m, double* v, uint64_t n, uint64_t t, double lr, double b1,
          double b2, double eps);

}  // namespace nrc
}  // namespace orc
