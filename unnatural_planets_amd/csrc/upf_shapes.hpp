// upf_shapes.hpp -- the extended base shapes (UPF_SHAPE_CYLINDER .. UPF_SHAPE_HEMISPHERES)
// as one template over the scalar type, compiled for three types:
//   float     the field kernels;
//   Grad      value and d/dx, d/dy, d/dz (normals and the adjoint's position gradient);
//   a host operation counter (the flop table, upf_common.cpp).
// Formulas are the reference's compositions (sources/sdf.cpp, cited per shape) over the
// textbook primitives; the primitives' axis conventions are those documented at
// upf_shape_id.  Every operation is a separate +, -, *, / or sqrt (no fused forms), so the
// operation count is the flop count.
#pragma once

#include <math.h>

#include "../../include/upf.h"

#if defined(__HIPCC__)
#define UPF_HD __host__ __device__ __forceinline__
#else
#define UPF_HD inline
#endif

namespace upf {

// value + gradient (forward mode, 3 directions)
struct Grad {
  float v, dx, dy, dz;
};
UPF_HD Grad gvar(float v, int axis) { return {v, axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f}; }
UPF_HD Grad gc(float c) { return {c, 0.0f, 0.0f, 0.0f}; }
UPF_HD float value_of(float a) { return a; }
// a parameter or constant as the scalar type (no operation)
UPF_HD float lift(float, float v) { return v; }
UPF_HD Grad lift(const Grad&, float v) { return gc(v); }
UPF_HD float value_of(const Grad& a) { return a.v; }
UPF_HD Grad operator+(Grad a, Grad b) { return {a.v + b.v, a.dx + b.dx, a.dy + b.dy, a.dz + b.dz}; }
UPF_HD Grad operator-(Grad a, Grad b) { return {a.v - b.v, a.dx - b.dx, a.dy - b.dy, a.dz - b.dz}; }
UPF_HD Grad operator-(Grad a) { return {-a.v, -a.dx, -a.dy, -a.dz}; }
UPF_HD Grad operator*(Grad a, Grad b) {
  return {a.v * b.v, a.dx * b.v + a.v * b.dx, a.dy * b.v + a.v * b.dy, a.dz * b.v + a.v * b.dz};
}
UPF_HD Grad operator/(Grad a, Grad b) {
  const float q = a.v / b.v;
  return {q, (a.dx - q * b.dx) / b.v, (a.dy - q * b.dy) / b.v, (a.dz - q * b.dz) / b.v};
}
UPF_HD Grad operator+(Grad a, float b) { return {a.v + b, a.dx, a.dy, a.dz}; }
UPF_HD Grad operator-(Grad a, float b) { return {a.v - b, a.dx, a.dy, a.dz}; }
UPF_HD Grad operator*(Grad a, float b) { return {a.v * b, a.dx * b, a.dy * b, a.dz * b}; }
UPF_HD Grad operator/(Grad a, float b) { return {a.v / b, a.dx / b, a.dy / b, a.dz / b}; }
UPF_HD Grad operator+(float a, Grad b) { return {a + b.v, b.dx, b.dy, b.dz}; }
UPF_HD Grad operator-(float a, Grad b) { return {a - b.v, -b.dx, -b.dy, -b.dz}; }
UPF_HD Grad operator*(float a, Grad b) { return {a * b.v, a * b.dx, a * b.dy, a * b.dz}; }
UPF_HD float sqrt_f(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_sqrtf(a);
#else
  return sqrtf(a);
#endif
}
UPF_HD Grad sq_root(Grad a) {
  const float r = sqrt_f(a.v);
  const float k = r > 0.0f ? 0.5f / r : 0.0f;
  return {r, a.dx * k, a.dy * k, a.dz * k};
}
UPF_HD float sq_root(float a) { return sqrt_f(a); }
template <class T> UPF_HD T abs_v(T a) { return value_of(a) < 0.0f ? -a : a; }
template <class T> UPF_HD T min_v(T a, T b) { return value_of(b) < value_of(a) ? b : a; }
template <class T> UPF_HD T max_v(T a, T b) { return value_of(a) < value_of(b) ? b : a; }
template <class T> UPF_HD T clamp01(T a, T zero, T one) { return min_v(max_v(a, zero), one); }

// ------------------------------------------------------------------ primitives
template <class T> UPF_HD T len2(T a, T b) { return sq_root(a * a + b * b); }
template <class T> UPF_HD T len3(T a, T b, T c) { return sq_root(a * a + b * b + c * c); }

// ring of radius R in the xz-plane (axis y), tube radius r
template <class T> UPF_HD T sd_torus(T x, T y, T z, T R, T r) { return len2(len2(x, z) - R, y) - r; }

// axis z, |z| <= h / 2, radius r
template <class T> UPF_HD T sd_cylinder(T x, T y, T z, T h, T r, T zero) {
  const T dr = len2(x, y) - r;
  const T dz = abs_v(z) - h * 0.5f;
  return min_v(max_v(dr, dz), zero) + len2(max_v(dr, zero), max_v(dz, zero));
}

// box of half extents (hx, hy, hz)
template <class T> UPF_HD T sd_box(T x, T y, T z, T hx, T hy, T hz, T zero) {
  const T qx = abs_v(x) - hx, qy = abs_v(y) - hy, qz = abs_v(z) - hz;
  return len3(max_v(qx, zero), max_v(qy, zero), max_v(qz, zero)) + min_v(max_v(qx, max_v(qy, qz)), zero);
}

// exact octahedron |x| + |y| + |z| = s (Quilez, "distance functions")
template <class T> UPF_HD T sd_octahedron(T x, T y, T z, T s, T zero) {
  const T ax = abs_v(x), ay = abs_v(y), az = abs_v(z);
  const T m = ax + ay + az - s;
  T qx, qy, qz;
  if (value_of(3.0f * ax) < value_of(m)) { qx = ax; qy = ay; qz = az; }
  else if (value_of(3.0f * ay) < value_of(m)) { qx = ay; qy = az; qz = ax; }
  else if (value_of(3.0f * az) < value_of(m)) { qx = az; qy = ax; qz = ay; }
  else return m * 0.57735027f;
  const T k = min_v(max_v((qz - qy + s) * 0.5f, zero), s);
  return len3(qx, qy - s + k, qz - k);
}

// smoothMin of sources/math.cpp:28-33
template <class T> UPF_HD T smooth_min_v(T a, T b, T k, T zero) {
  const T h = clamp01((b - a) / k * 0.5f + 0.5f, zero, lift(zero, 1.0f));
  return b + (a - b) * h - k * h * (1.0f - h);
}

// ------------------------------------------------------------------ shapes
// Value of extended shape `shape` at (x, y, z); sp = shape parameters.
template <class T> UPF_HD T shape_ext(int shape, const float* spf, T x, T y, T z, T zero) {
  const T sp[4] = {lift(zero, spf[0]), lift(zero, spf[1]), lift(zero, spf[2]), lift(zero, spf[3])};
  const T c150 = lift(zero, 150.0f), c100 = lift(zero, 100.0f), c1000 = lift(zero, 1000.0f);
  const T c1300 = lift(zero, 1300.0f), c10000 = lift(zero, 10000.0f), c200 = lift(zero, 200.0f);
  const T c800 = lift(zero, 800.0f), c400 = lift(zero, 400.0f);
  switch (shape) {
    case UPF_SHAPE_CYLINDER:  // sdfTube / sdfDisk over sdfCylinder, sources/sdf.cpp:30-38
      return sd_cylinder(x, y, z, sp[0], sp[1], zero);
    case UPF_SHAPE_PLANE:  // sdfHexagon / sdfSquare over sdfPlane, sources/sdf.cpp:10-18
      return x * sp[0] + y * sp[1] + z * sp[2] - sp[3];
    case UPF_SHAPE_OCTAHEDRON:  // sources/sdf.cpp:55-58
      return sd_octahedron(x, y, z, sp[0], zero) - sp[1];
    case UPF_SHAPE_DOUBLE_TORUS:  // sources/sdf.cpp:292-299
      return smooth_min_v(sd_torus(y, z, x, sp[0], sp[1]), sd_torus(y, x, z, sp[0], sp[1]), sp[2], zero);
    case UPF_SHAPE_TORUS_CROSS:  // sources/sdf.cpp:301-308
      return smooth_min_v(sd_torus(y - sp[0], z, x, sp[0], sp[1]), sd_torus(y + sp[0], x, z, sp[0], sp[1]), sp[2],
                          zero);
    case UPF_SHAPE_H2O: {  // sources/sdf.cpp:141-149
      const T px = x * 0.75f, py = y * 0.75f, pz = z * 0.75f;
      const T h1 = len3(px + 550.0f, py - 300.0f, pz) - 450.0f;
      const T h2 = len3(px - 550.0f, py - 300.0f, pz) - 450.0f;
      const T o = len3(px, py + 100.0f, pz) - 650.0f;
      return smooth_min_v(o, min_v(h1, h2), c150, zero) / 0.75f;
    }
    case UPF_SHAPE_BOWL:  // sources/sdf.cpp:310-313
      return 2000.0f - len3(x, y, z + 3000.0f);
    case UPF_SHAPE_INSIDE_CUBE:  // sources/sdf.cpp:315-318
      return -(sd_box(x, y, z, c1300, c1300, c1300, zero) - 200.0f);
    case UPF_SHAPE_WORMHOLE: {  // sources/sdf.cpp:399-404
      const T a = min_v(len2(x, z) - 1400.0f, zero);
      return 700.0f - len2(a, y);
    }
    case UPF_SHAPE_PIPE: {  // sources/sdf.cpp:355-366
      T px = x, py = y, pz = z;
      if (value_of(px) < 0.0f) {
        px = -px;
        const T t = py;
        py = pz;
        pz = t;
      }
      py = abs_v(py) - 800.0f;
      return sd_torus(px, pz, py, c800, c400);
    }
    case UPF_SHAPE_HEMISPHERES: {  // sources/sdf.cpp:383-397
      const float s = 0.70710678f;
      const T ax = abs_v(x), ay = abs_v(y);
      // q = rotate45(p + (-2000, 0, 0)), r2 = rotate45(p + (0, -2000, 0))
      const T qx = ((ax - 2000.0f) - ay) * s, qy = ((ax - 2000.0f) + ay) * s;
      const T rx = (ax - (ay - 2000.0f)) * s, ry = (ax + (ay - 2000.0f)) * s;
      const T c = sd_cylinder(qy, z, qx, c10000, c200, zero);
      const T s1 = len3(qx, qy, z) - c1000;
      const T s2 = len3(rx, ry, z) - c1000;
      return smooth_min_v(c, min_v(s1, s2), c100, zero);
    }
    default:
      return zero;
  }
}

}  // namespace upf
