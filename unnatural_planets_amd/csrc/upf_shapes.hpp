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

#include "upf_math.hpp"

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
// Correctly rounded square root (the compositions below rely on sqrt(x^2 + y^2 + z^2)
// >= |z|, e.g. acos(z / r) in the Mandelbulb).
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
// transcendental functions of the scalar type (the C library's for float, except pow:
// the library's own, upf_math.hpp; forward-mode derivatives for Grad)
UPF_HD float pow_v(float a, float e) { return upf_math::powf_(a, e); }
UPF_HD Grad pow_v(const Grad& a, float e) {
  const float v = upf_math::powf_(a.v, e);
  const float d = a.v != 0.0f ? e * upf_math::powf_(a.v, e - 1.0f) : 0.0f;  // d/da a^e, 0 at a = 0
  return {v, d * a.dx, d * a.dy, d * a.dz};
}
UPF_HD float floor_v(float a) { return floorf(a); }
UPF_HD Grad floor_v(const Grad& a) { return gc(floorf(a.v)); }
UPF_HD float sin_v(float a) { return sinf(a); }
// acc + w h with a constant weight w: one fused multiply-add per component
UPF_HD float fma_s(float w, float h, float acc) { return fmaf(w, h, acc); }
UPF_HD Grad sin_v(const Grad& a) {
  const float c = cosf(a.v);
  return {sinf(a.v), c * a.dx, c * a.dy, c * a.dz};
}
UPF_HD Grad fma_s(float w, const Grad& h, const Grad& acc) {
  return {fmaf(w, h.v, acc.v), fmaf(w, h.dx, acc.dx), fmaf(w, h.dy, acc.dy), fmaf(w, h.dz, acc.dz)};
}
UPF_HD float cos_v(float a) { return cosf(a); }
UPF_HD Grad cos_v(const Grad& a) {
  const float s = -sinf(a.v);
  return {cosf(a.v), s * a.dx, s * a.dy, s * a.dz};
}
UPF_HD float atan2_v(float y, float x) { return atan2f(y, x); }
UPF_HD Grad atan2_v(const Grad& y, const Grad& x) {
  const float r2 = x.v * x.v + y.v * y.v;
  const float k = r2 > 0.0f ? 1.0f / r2 : 0.0f;  // d atan2 = (x dy - y dx) / (x^2 + y^2)
  return {atan2f(y.v, x.v), (x.v * y.dx - y.v * x.dx) * k, (x.v * y.dy - y.v * x.dy) * k,
          (x.v * y.dz - y.v * x.dz) * k};
}
UPF_HD float acos_v(float a) { return acosf(a); }
UPF_HD Grad acos_v(const Grad& a) {
  const float t = 1.0f - a.v * a.v;
  const float k = t > 0.0f ? -1.0f / sqrt_f(t) : 0.0f;
  return {acosf(a.v), k * a.dx, k * a.dy, k * a.dz};
}
UPF_HD float log_v(float a) { return logf(a); }
UPF_HD Grad log_v(const Grad& a) {
  const float k = 1.0f / a.v;
  return {logf(a.v), k * a.dx, k * a.dy, k * a.dz};
}
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

// Closest point on triangle (a, b, c) to p by Voronoi regions of its vertices, edges and
// face (Ericson, Real-Time Collision Detection, 5.1.5); returns |p - closest|.  The
// triangle corners are constants (no derivative through them).
template <class T>
UPF_HD T dist_triangle(T px, T py, T pz, const float a[3], const float b[3], const float c[3], T zero) {
  const float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const T apx = px - a[0], apy = py - a[1], apz = pz - a[2];
  const T d1 = apx * ab[0] + apy * ab[1] + apz * ab[2], d2 = apx * ac[0] + apy * ac[1] + apz * ac[2];
  const T bpx = px - b[0], bpy = py - b[1], bpz = pz - b[2];
  const T d3 = bpx * ab[0] + bpy * ab[1] + bpz * ab[2], d4 = bpx * ac[0] + bpy * ac[1] + bpz * ac[2];
  const T cpx = px - c[0], cpy = py - c[1], cpz = pz - c[2];
  const T d5 = cpx * ab[0] + cpy * ab[1] + cpz * ab[2], d6 = cpx * ac[0] + cpy * ac[1] + cpz * ac[2];
  const T va = d3 * d6 - d5 * d4, vb = d5 * d2 - d1 * d6, vc = d1 * d4 - d3 * d2;
  T qx, qy, qz;  // closest point
  if (value_of(d1) <= 0.0f && value_of(d2) <= 0.0f) { qx = lift(zero, a[0]); qy = lift(zero, a[1]); qz = lift(zero, a[2]); }
  else if (value_of(d3) >= 0.0f && value_of(d4) <= value_of(d3)) {
    qx = lift(zero, b[0]); qy = lift(zero, b[1]); qz = lift(zero, b[2]);
  } else if (value_of(vc) <= 0.0f && value_of(d1) >= 0.0f && value_of(d3) <= 0.0f) {
    const T v = d1 / (d1 - d3);
    qx = a[0] + v * ab[0]; qy = a[1] + v * ab[1]; qz = a[2] + v * ab[2];
  } else if (value_of(d6) >= 0.0f && value_of(d5) <= value_of(d6)) {
    qx = lift(zero, c[0]); qy = lift(zero, c[1]); qz = lift(zero, c[2]);
  } else if (value_of(vb) <= 0.0f && value_of(d2) >= 0.0f && value_of(d6) <= 0.0f) {
    const T w = d2 / (d2 - d6);
    qx = a[0] + w * ac[0]; qy = a[1] + w * ac[1]; qz = a[2] + w * ac[2];
  } else if (value_of(va) <= 0.0f && value_of(d4 - d3) >= 0.0f && value_of(d5 - d6) >= 0.0f) {
    const T w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    qx = b[0] + w * (c[0] - b[0]); qy = b[1] + w * (c[1] - b[1]); qz = b[2] + w * (c[2] - b[2]);
  } else {
    const T den = va + vb + vc;
    const T v = vb / den, w = vc / den;
    qx = a[0] + ab[0] * v + ac[0] * w; qy = a[1] + ab[1] * v + ac[1] * w; qz = a[2] + ab[2] * v + ac[2] * w;
  }
  return len3(px - qx, py - qy, pz - qz);
}

// hexagonal prism along z: apothem hx, half length hz (Quilez, "distance functions")
template <class T> UPF_HD T sd_hex_prism(T x, T y, T z, float hx, float hz, T zero) {
  const float kx = -0.8660254f, ky = 0.5f, kz = 0.57735027f;
  T px = abs_v(x), py = abs_v(y);
  const T pz = abs_v(z);
  const T dk = min_v(px * kx + py * ky, zero) * 2.0f;
  px = px - dk * kx;
  py = py - dk * ky;
  const T cx = min_v(max_v(px, lift(zero, -kz * hx)), lift(zero, kz * hx));
  const T ex = px - cx, ey = py - hx;
  const T dx = value_of(ey) < 0.0f ? -len2(ex, ey) : len2(ex, ey);
  const T dy = pz - hz;
  return min_v(max_v(dx, dy), zero) + len2(max_v(dx, zero), max_v(dy, zero));
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

template <class T> UPF_HD T smooth_max_s(T a, T b, T k, T zero) { return -smooth_min_v(-a, -b, k, zero); }

// UPF_SHAPE_SINE_NET, the architecture of sdfBunny and sdfMonkeyHead (sources/sdf.cpp:
// 252-303): a sine network of width 16 at q = (x, z, y) * scale, weights w in the layout
// documented at UPF_SINE_NET_WEIGHTS (the order of the reference's source; the Mat4
// blocks column-major), residual hidden layers divided by d1 and d2.  Outside the unit
// ball (|q| > 1) the distance is (|q| - 0.8) / scale, as in the reference.
template <class T> UPF_HD T sine_net_t(const float* w, float scale, float d1, float d2, T x, T y, T z) {
  const T qx = x * scale, qy = z * scale, qz = y * scale;
  const T len = len3(qx, qy, qz);
  if (value_of(len) > 1.0f) return (len - 0.8f) / scale;
  T h0[16], h1[16];
  for (int b = 0; b < 4; b++)
    for (int i = 0; i < 4; i++) {
      const float* L = w + 16 * b;  // A, B, C, D of block b
      h0[4 * b + i] = sin_v(fma_s(L[8 + i], qx, fma_s(L[4 + i], qz, fma_s(L[i], qy, lift(x, L[12 + i])))));
    }
  // h_out[o] = sin(sum_b M[o][b] h_in[b] + V[o]) / d + h_in[o]; M[o][b] element (col c,
  // row r) at 16 b + 4 c + r of the block's 68 floats, V[o] at 64
  for (int o = 0; o < 4; o++) {
    const float* B = w + 64 + 68 * o;
    for (int r = 0; r < 4; r++) {
      T acc = lift(x, B[64 + r]);
      for (int b = 0; b < 4; b++)
        for (int c = 0; c < 4; c++) acc = fma_s(B[16 * b + 4 * c + r], h0[4 * b + c], acc);
      h1[4 * o + r] = sin_v(acc) / d1 + h0[4 * o + r];
    }
  }
  for (int o = 0; o < 4; o++) {  // layer 2 into h0 (layer 0 is no longer needed)
    const float* B = w + 336 + 68 * o;
    for (int r = 0; r < 4; r++) {
      T acc = lift(x, B[64 + r]);
      for (int b = 0; b < 4; b++)
        for (int c = 0; c < 4; c++) acc = fma_s(B[16 * b + 4 * c + r], h1[4 * b + c], acc);
      h0[4 * o + r] = sin_v(acc) / d2 + h1[4 * o + r];
    }
  }
  T f = lift(x, w[624]);
  for (int u = 0; u < 16; u++) f = fma_s(w[608 + u], h0[u], f);
  return f / scale;
}

// The further shapes of sources/sdf.cpp (ids UPF_SHAPE_H3O .. UPF_SHAPE_TETRAHEDRON; the
// asteroid, which samples noise, is evaluated by the caller).  cage's atan2(a, b) is
// read as C's atan2(a, b) and cage's quaternion rolls as rotations about z.
template <class T> UPF_HD T shape_more_t(int shape, const float* spf, T x, T y, T z, T zero) {
  switch (shape) {
    case UPF_SHAPE_H3O: case UPF_SHAPE_H4O: {  // sources/sdf.cpp:151-178
      const T px = x * 0.75f, py = y * 0.75f, pz = z * 0.75f;
      T h;
      if (shape == UPF_SHAPE_H3O) {
        const float c = -340.0f, s = 588.89727f;  // 680 (cos, sin) of 120 degrees
        const T h0 = len3(px - 680.0f, py, pz) - 450.0f;
        const T h1 = len3(px - c, py - s, pz) - 450.0f;
        const T h2 = len3(px - c, py + s, pz) - 450.0f;
        h = min_v(min_v(h0, h1), h2);
      } else {
        const T h0 = len3(px + 550.0f, py - 400.0f, pz) - 450.0f;
        const T h1 = len3(px - 550.0f, py - 400.0f, pz) - 450.0f;
        const T h2 = len3(px, py + 400.0f, pz + 550.0f) - 450.0f;
        const T h3 = len3(px, py + 400.0f, pz - 550.0f) - 450.0f;
        h = min_v(min_v(h0, h1), min_v(h2, h3));
      }
      const T o = len3(px, py, pz) - 650.0f;
      return smooth_min_v(o, h, lift(zero, 150.0f), zero) / 0.75f;
    }
    case UPF_SHAPE_GEAR: {  // :180-204; pos = (x, z, y) / 2, the gear in its (x, z) = our (x, y)
      const float teeth = 9.0f, angle = 6.2831853f / teeth;
      const T gx = x * 0.5f, gy = y * 0.5f, gz = z * 0.5f;
      const float a = atan2f(value_of(gx), value_of(gy)) / angle + teeth + 0.5f;
      const unsigned sector = ((unsigned)a) % 9u;
      const float r = (float)sector * -angle;
      const float s = sinf(r), c = cosf(r);
      const T qx = c * gx - s * gy, qy = s * gx + c * gy;
      const T bx = max_v(abs_v(qx - 700.0f) - 150.0f, zero), by = max_v(abs_v(qy) - 70.0f, zero);
      const T b = len2(bx, by);
      const T cc = abs_v(len2(gx, gy) - 450.0f) - 150.0f;
      const T h = abs_v(gz) - 30.0f;
      return (smooth_max_s(min_v(b, cc), h, lift(zero, 50.0f), zero) - 50.0f) * 2.0f;
    }
    case UPF_SHAPE_MANDELBULB: {  // :206-232
      if (value_of(x) * value_of(x) + value_of(y) * value_of(y) + value_of(z) * value_of(z) < 1e-3f) return zero;
      const T px = x * 0.0004f, py = y * 0.0004f, pz = z * 0.0004f;
      T zx = px, zy = py, zz = pz;
      T dr = lift(zero, 1.0f), r = zero;
      for (int i = 0; i < 20; i++) {
        r = len3(zx, zy, zz);
        if (value_of(r) > 4.0f) break;
        T theta = acos_v(zz / r);
        T phi = atan2_v(zy, zx);
        dr = pow_v(r, 2.0f) * 3.0f * dr + 1.0f;
        const T zr = pow_v(r, 3.0f);
        theta = theta * 3.0f;
        phi = phi * 3.0f;
        const T st = sin_v(theta);
        zx = zr * (st * cos_v(phi)) + px;
        zy = zr * (sin_v(phi) * st) + py;
        zz = zr * cos_v(theta) + pz;
      }
      const T value = 0.5f * log_v(r) * r / dr;
      return (value + 0.1f) * 150.0f;
    }
    case UPF_SHAPE_MOBIUS: {  // :94-120 with pos = (x, z, y): the strip around the z axis
      const T l = len2(x, y);
      // dot(v, normalize(v)) of v = (x, y), taken as 0 on the axis where the reference
      // normalizes a zero vector
      const T proj = value_of(l) > 0.0f ? x * (x / l) + y * (y / l) : zero;
      const T rot = atan2_v(x, y) * 0.5f;
      const T u = proj - 1000.0f, v = z;
      const T a = atan2_v(u, v) + rot;
      const T lp = len2(u, v);
      const T qx = cos_v(a) * lp, qy = sin_v(a) * lp;
      const T dx = abs_v(qx) - 400.0f, dy = abs_v(qy) - 40.0f;
      return len2(max_v(dx, zero), max_v(dy, zero)) + min_v(max_v(dx, dy), zero) - 200.0f;
    }
    case UPF_SHAPE_FIBERS: {  // :122-139; the reference's random offset is params (ox, oy, oz)
      const float sc = 0.0007f, g = 3.23f;
      const T px = (x * sc + spf[0]) * g, py = (y * sc + spf[1]) * g, pz = (z * sc + spf[2]) * g;
      const T dot = sin_v(px) * cos_v(pz) + sin_v(py) * cos_v(px) + sin_v(pz) * cos_v(py);
      const T gy = abs_v(dot + 1.6f) / g - 0.2f;
      const T v = gy * 0.7f / sc;
      const T d = max_v(len3(x, y, z) - 1500.0f, zero) * 0.5f;
      return v + d;
    }
    case UPF_SHAPE_TWISTED_PLANE: {  // :368-373
      const T m = min_v(abs_v(z) / 1000.0f, lift(zero, 1.0f));
      const T rot = 1.0f - (1.0f - m) * (1.0f - m);
      const float sg = value_of(z) > 0.0f ? 1.0f : (value_of(z) < 0.0f ? -1.0f : 0.0f);
      const T ang = rot * (sg * 35.0f * 0.017453292f);
      return y * cos_v(ang) + x * sin_v(ang);
    }
    case UPF_SHAPE_BELT:  // :375-381
      if (value_of(x) > 0.0f)
        return -sd_cylinder(x, y, z, lift(zero, 10000.0f), lift(zero, 1800.0f), zero);
      return 1800.0f - abs_v(z);
    case UPF_SHAPE_TRIANGULAR_PRISM: {  // :62-74 with pos = (x, z, y): the prism along y
      const float a[3] = {0.0f, 900.0f, 0.0f}, b[3] = {779.42286f, -450.0f, 0.0f}, c[3] = {-779.42286f, -450.0f, 0.0f};
      const T pzz = max_v(abs_v(y) - 150.0f, zero);
      return dist_triangle(x, z, pzz, a, b, c, zero) - 100.0f;
    }
    case UPF_SHAPE_HEXAGONAL_PRISM:  // :76-79 with pos = (x, z, y): along y, apothem 400, length 300
      return sd_hex_prism(x, z, y, 400.0f, 150.0f, zero) - 100.0f;
    case UPF_SHAPE_TWISTED_HEX_PRISM: {  // :234-242: (x, z) turned by y / 1000 rad about y
      const T ang = y * 0.001f;
      const T c = cos_v(ang), s = sin_v(ang);
      const T rx = c * x - s * z, rz = s * x + c * z;
      return sd_hex_prism(rx, rz, y, 650.0f, 450.0f, zero) - 200.0f;
    }
    case UPF_SHAPE_KNOT: {  // :86-90: knot(q) / |(1, 1, 0.3)| at q = (x, y, 0.3 z)
      // the (3, 2) torus knot about z (the form documented at UPF_SHAPE_KNOT): in the
      // half-plane of q's azimuth the two strands that cross it sit 200 from the core
      // circle of radius 1600 at angles 1.5 phi and 1.5 (phi + 2 pi)
      const T qz = z * 0.3f;
      const T rho = len2(x, y);
      const T phi = atan2_v(y, x);
      T best = lift(zero, 1e30f);
      for (int j = 0; j < 2; j++) {
        const T a = (phi + 6.2831853f * (float)j) * 1.5f;
        const T dr = rho - (1600.0f + cos_v(a) * 200.0f);
        const T dz = qz - sin_v(a) * 200.0f;
        best = min_v(best, len2(dr, dz));
      }
      return (best - 100.0f) / 1.0440307f;  // |(1, 1, 0.3)| = sqrt(1.09)
    }
    case UPF_SHAPE_TETRAHEDRON: {  // :50-53: circumradius 600, vertices s (1,1,1), s (1,-1,-1), ...
      const float k = 346.41016f;  // 600 / sqrt(3)
      const float v[4][3] = {{k, k, k}, {k, -k, -k}, {-k, k, -k}, {-k, -k, k}};
      const int f[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
      // the face opposite vertex i has outward normal -v_i / |v_i| at distance k / sqrt(3)
      T inside = lift(zero, -1e30f), dmin = lift(zero, 1e30f);
      for (int i = 0; i < 4; i++) {
        const T pl = (x * v[i][0] + y * v[i][1] + z * v[i][2]) * (-1.0f / 600.0f) - 200.0f;
        inside = max_v(inside, pl);
        dmin = min_v(dmin, dist_triangle(x, y, z, v[f[i][0]], v[f[i][1]], v[f[i][2]], zero));
      }
      return (value_of(inside) > 0.0f ? dmin : inside) - 100.0f;
    }
    default:
      return zero;
  }
}

}  // namespace upf
