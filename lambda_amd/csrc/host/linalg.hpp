// Host-side small linear algebra with the reference's evaluation order.
//
// Every sum is accumulated left to right from 0.f and every product is a
// separate IEEE-754 single rounding (this file is compiled with
// -ffp-contract=off), so the matrices that reach the device have the same
// bits as the reference's nl::bra operators produce:
//   mat*mat, mat*vec       lib/numli/bra.hpp:310-336
//   cross, dot, normalize  lib/numli/bra.hpp:202-216, 386-393
#pragma once
#include <cmath>
#include <cstdint>

namespace lamh {

struct V3 {
  float x = 0.f, y = 0.f, z = 0.f;
  V3() = default;
  V3(float a, float b, float c) : x(a), y(b), z(c) {}
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {a.x * -1.f, a.y * -1.f, a.z * -1.f}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
// cross product, bra.hpp:388-389
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// dot product, bra.hpp:392-393
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// len2 / l2 / normalized, bra.hpp:202-216
inline float len2(V3 a) {
  float s = 0.f;
  s += a.x * a.x;
  s += a.y * a.y;
  s += a.z * a.z;
  return s;
}
inline float l2(V3 a) { return std::sqrt(len2(a)); }
inline V3 normalized(V3 a) {
  float l = l2(a);
  return {a.x / l, a.y / l, a.z / l};
}

// 4x4 row-major; row 3 kept explicitly so the products match the reference
struct M4 {
  float m[16];
  void identity() {
    for (int i = 0; i < 16; i++) m[i] = (i % 5 == 0) ? 1.f : 0.f;
  }
  float operator()(int i, int j) const { return m[i * 4 + j]; }
  float &operator()(int i, int j) { return m[i * 4 + j]; }
};
// bra.hpp:310-321
inline M4 matmul(const M4 &A, const M4 &B) {
  M4 C;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      float s = 0.f;
      for (int k = 0; k < 4; k++) s += A.m[i * 4 + k] * B.m[k * 4 + j];
      C.m[i * 4 + j] = s;
    }
  return C;
}
// bra.hpp:325-336 with b = (v, w)
inline V3 matvec(const M4 &A, V3 v, float w) {
  float b[4] = {v.x, v.y, v.z, w};
  float r[3];
  for (int i = 0; i < 3; i++) {
    float s = 0.f;
    for (int k = 0; k < 4; k++) s += A.m[i * 4 + k] * b[k];
    r[i] = s;
  }
  return {r[0], r[1], r[2]};
}

// transform (cg.hpp:422-496)
struct Xform {
  M4 M, Minv;
  Xform() { M.identity(); Minv.identity(); }
  V3 pos() const { return {M(0, 3), M(1, 3), M(2, 3)}; }
  V3 column(int j) const { return {M(0, j), M(1, j), M(2, j)}; }
  V3 pntToWorld(V3 p) const { return matvec(M, p, 1.f); }
  V3 vecToWorld(V3 v) const { return matvec(M, v, 0.f); }
  // subMatT<3,3>(M_inv) * n   (cg.hpp:442-443, bra.hpp:419-427)
  V3 normalToWorld(V3 n) const {
    float r[3];
    for (int i = 0; i < 3; i++) {
      float s = 0.f;
      for (int k = 0; k < 3; k++) s += Minv(k, i) * n[k];
      r[i] = s;
    }
    return {r[0], r[1], r[2]};
  }
};

// transform::scale (cg.hpp:446-451)
inline Xform xf_scale(V3 x) {
  Xform T;
  for (int i = 0; i < 3; i++) {
    T.M(i, i) = x[i];
    T.Minv(i, i) = 1.f / x[i];
  }
  return T;
}
// transform::rotate (cg.hpp:452-480).  The reference writes the transpose
// of the textbook Rodrigues matrix, i.e. a rotation by -degrees; kept.
inline Xform xf_rotate(V3 axis, float degrees) {
  const V3 u = normalized(axis);
  Xform T;
  const float pi = 3.14159265358979323846f;
  const float th = degrees * pi / 180.f;
  const float c = std::cos(th);
  const float s = std::sin(th);
  T.M(0, 0) = c + (1 - c) * u[0] * u[0];
  T.M(0, 1) = (1 - c) * u[0] * u[1] + s * u[2];
  T.M(0, 2) = (1 - c) * u[0] * u[2] - s * u[1];
  T.M(1, 0) = (1 - c) * u[1] * u[0] - s * u[2];
  T.M(1, 1) = c + (1 - c) * u[1] * u[1];
  T.M(1, 2) = (1 - c) * u[1] * u[2] + s * u[0];
  T.M(2, 0) = (1 - c) * u[2] * u[0] + s * u[1];
  T.M(2, 1) = (1 - c) * u[2] * u[1] - s * u[0];
  T.M(2, 2) = c + (1 - c) * u[2] * u[2];
  T.Minv(0, 0) = c + (1 - c) * u[0] * u[0];
  T.Minv(0, 1) = (1 - c) * u[0] * u[1] - s * u[2];
  T.Minv(0, 2) = (1 - c) * u[0] * u[2] + s * u[1];
  T.Minv(1, 0) = (1 - c) * u[1] * u[0] + s * u[2];
  T.Minv(1, 1) = c + (1 - c) * u[1] * u[1];
  T.Minv(1, 2) = (1 - c) * u[1] * u[2] - s * u[0];
  T.Minv(2, 0) = (1 - c) * u[2] * u[0] - s * u[1];
  T.Minv(2, 1) = (1 - c) * u[2] * u[1] + s * u[0];
  T.Minv(2, 2) = c + (1 - c) * u[2] * u[2];
  return T;
}
// transform::translate (cg.hpp:481-486)
inline Xform xf_translate(V3 x) {
  Xform T;
  for (int i = 0; i < 3; i++) {
    T.M(i, 3) = x[i];
    T.Minv(i, 3) = -x[i];
  }
  return T;
}
// operator<< : T2 applied after T1 (cg.hpp:490-496)
inline Xform xf_then(const Xform &T1, const Xform &T2) {
  Xform T;
  T.M = matmul(T2.M, T1.M);
  T.Minv = matmul(T1.Minv, T2.Minv);
  return T;
}

}  // namespace lamh
