// ORACLE — test infrastructure only. CPU f64 restatement of render.rs's
// math types. Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline may load anything under oracle/; the product library never does.
//
// Follows src/math/vector.rs, src/math/point.rs and src/math/matrix.rs of
// render.rs (f64 throughout, EPSILON = 1e-6 from src/math/mod.rs:9).
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

constexpr double EPSILON = 1e-6;  // math/mod.rs:9
constexpr double PI = 3.14159265358979323846;

struct Vec3 {
  double x = 0, y = 0, z = 0;
  Vec3() = default;
  Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
  static Vec3 zero() { return {0, 0, 0}; }
  static Vec3 one() { return {1, 1, 1}; }
  double length_squared() const { return x * x + y * y + z * z; }
  double length() const { return std::sqrt(length_squared()); }
  // vector.rs:31-38 — divides only when the length exceeds EPSILON.
  Vec3 normalize() const {
    double len = length();
    if (len > EPSILON) return {x / len, y / len, z / len};
    return *this;
  }
  double dot(const Vec3& o) const { return x * o.x + y * o.y + z * o.z; }
  // vector.rs:44-50
  Vec3 cross(const Vec3& o) const {
    return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x};
  }
  // vector.rs:52-54
  Vec3 reflect(const Vec3& n) const { return *this - n * 2.0 * dot(n); }
  Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  Vec3 operator-() const { return {-x, -y, -z}; }
  Vec3 operator*(double s) const { return {x * s, y * s, z * s}; }
  Vec3 operator/(double s) const { return {x / s, y / s, z / s}; }
  // Component-wise product (vector.rs Mul<Vec3>).
  Vec3 operator*(const Vec3& o) const { return {x * o.x, y * o.y, z * o.z}; }
  double operator[](int a) const { return a == 0 ? x : (a == 1 ? y : z); }
};

inline double luminance(const Vec3& c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }
inline double max_component(const Vec3& c) { return std::fmax(std::fmax(c.x, c.y), c.z); }

// Ray (raytracer/ray.rs:3-29). `make` normalises like Ray::new; `raw`
// keeps the direction as given (mesh.rs:316-323 builds local rays that way).
struct Ray {
  Vec3 origin, direction;
  double time = 0;
  static Ray make(const Vec3& o, const Vec3& d) { return Ray{o, d.normalize(), 0.0}; }
  static Ray raw(const Vec3& o, const Vec3& d) { return Ray{o, d, 0.0}; }
  Vec3 at(double t) const { return origin + direction * t; }
};

// Row-major 4x4 (matrix.rs). Only what the oracle needs.
struct Mat4 {
  double m[4][4];
  static Mat4 identity() {
    Mat4 r{};
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) r.m[i][j] = (i == j) ? 1.0 : 0.0;
    return r;
  }
  // matrix.rs:199-210
  Vec3 transform_point(const Vec3& p) const {
    double X = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    double Y = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    double Z = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
    double W = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
    if (std::fabs(W - 1.0) < EPSILON) return {X, Y, Z};
    return {X / W, Y / W, Z / W};
  }
  // matrix.rs:212-217
  Vec3 transform_vec(const Vec3& v) const {
    return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z,
            m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
            m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
  }
  // matrix.rs:222-227 (transpose of the upper 3x3)
  Vec3 transform_normal(const Vec3& n) const {
    return {m[0][0] * n.x + m[1][0] * n.y + m[2][0] * n.z,
            m[0][1] * n.x + m[1][1] * n.y + m[2][1] * n.z,
            m[0][2] * n.x + m[1][2] * n.y + m[2][2] * n.z};
  }
  Mat4 operator*(const Mat4& o) const {
    Mat4 r{};
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        double s = 0;
        for (int k = 0; k < 4; k++) s += m[i][k] * o.m[k][j];
        r.m[i][j] = s;
      }
    return r;
  }
  bool inverse(Mat4& out) const;  // cofactor inverse (matrix.rs:86-197)
  // matrix.rs:246-251: the cube root of the axes' stretch product
  double approx_scale() const {
    const double lx = transform_vec({1, 0, 0}).length(), ly = transform_vec({0, 1, 0}).length(),
                 lz = transform_vec({0, 0, 1}).length();
    return std::fmax(std::cbrt(lx * ly * lz), 1e-12);
  }
};

}  // namespace orc
