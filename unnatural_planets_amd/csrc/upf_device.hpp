// upf_device.hpp -- gfx950 device functions of the field path: lattice hash, noise,
// base shapes, layer stack and field composition (forward values).
//
// Everything is FP32 VALU work; the integer lattice (floor, hash, gradient selection)
// is bit-exact with the CPU restatement.  This file is compiled with -ffp-contract=off:
// every fused multiply-add is written as fmaf() on purpose, everything else rounds per
// operation.  The field kernels' default form fuses and uses the hardware square root
// (a few ulps from the restatement); the functions that take `X = true` instead round
// each operation separately, in the restatement's order, with the correctly rounded
// square root, so that their results equal the CPU restatement's bit for bit.  The texel
// bake, the material stack's elevation and the elevation modes run that form.
//
// Reference anchors: field composition sources/modesConfiguration.cpp:230-279, base
// shapes sources/sdf.cpp:20-48,81-84, layer configs sources/elevationModes.cpp:16-27.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/upf.h"
#include "upf_modes.hpp"
#include "upf_shapes.hpp"

namespace upf {

// Square root of the hot field's distances (sphere, box, torus, capsule, cellular F1):
// the hardware v_sqrt_f32 (1 ulp) instead of the correctly rounded expansion (denormal
// scaling, two Newton corrections, class fix-ups: 16 instructions), which made up 8% of
// the row field kernel's instructions on the torus + noise workload.  1 ulp is far
// inside the field's 1e-5 tolerance; the compositions of upf_shapes.hpp keep sqrt_f.
__device__ __forceinline__ float sqrt_hw(float a) { return __builtin_amdgcn_sqrtf(a); }

// a * b + c fused (X = false) or with two roundings (X = true); the square root of the
// hot distances in hardware (X = false) or correctly rounded (X = true)
template <bool X> __device__ __forceinline__ float mad_x(float a, float b, float c) { return X ? a * b + c : fmaf(a, b, c); }
template <bool X> __device__ __forceinline__ float sqrt_x(float a) { return X ? sqrt_f(a) : sqrt_hw(a); }

// Kernel-side copy of a layer: the spec's layer plus the cellular constants
// js = jitter/1023 and j0 = jitter*0.5, precomputed on the host in FP32.
struct DLayer {
  int32_t noise, fractal, octaves, blend;
  uint32_t seed;
  float weight, jitter, js, j0, pad;
  float freq[UPF_MAX_OCTAVES];
  float amp[UPF_MAX_OCTAVES];
};

struct DSpec {
  int32_t shape, compose, n_layers, all_add;
  float sp[UPF_MAX_SHAPE_PARAMS];
  float escale;
  int32_t mode;  // upf_elevation_mode_id
  const float* w;  // UPF_SHAPE_SINE_NET's weights (device memory)
  DLayer L[UPF_MAX_LAYERS];
};

struct DGrid {
  float o[3], s[3];
  int32_t n[3];
  int32_t pad;
};

// Whether a spec needs the EXT instantiations (an elevation mode or a further shape).
__host__ __device__ inline bool spec_ext(const DSpec& s) {
  return s.mode != UPF_ELEV_LAYERS || s.shape >= UPF_SHAPE_H3O;
}

// Whether a spec's shape is one of the compositions (cylinder .. hemispheres).  Kernels
// instantiated with COMPOSED = false hold only the primitives (none, sphere, box, torus,
// capsule): the compositions' code in the shape switch costs a hot kernel ~3% even when
// never taken (registers and code size), so the host picks the instantiation.
__host__ __device__ inline bool spec_composed(const DSpec& s) {
  return s.shape >= UPF_SHAPE_CYLINDER && s.shape < UPF_SHAPE_H3O;
}

// ------------------------------------------------------------------ hash and lattice
constexpr uint32_t kPX = 501125321u;
constexpr uint32_t kPY = 1136930381u;
constexpr uint32_t kPZ = 1720413743u;
constexpr uint32_t kHM = 0x27d4eb2du;

__device__ __forceinline__ float node_coord(float o, float s, int i) {
  const float m = (float)i * s;
  return o + m;
}

// Quintic fade t^3 (t (6t - 15) + 10).
template <bool X = false> __device__ __forceinline__ float fade(float t) {
  const float a = mad_x<X>(t, 6.0f, -15.0f);
  const float b = mad_x<X>(t, a, 10.0f);
  return (t * t) * t * b;
}

// d fade / dt = 30 t^2 (t - 1)^2
__device__ __forceinline__ float dfade(float t) {
  const float a = t * (t - 1.0f);
  return 30.0f * a * a;
}

template <bool X = false> __device__ __forceinline__ float lerpf(float a, float b, float t) { return mad_x<X>(t, b - a, a); }

// Dot product with gradient g = h >> 28 of Perlin's 16-entry improved-noise set:
// u = g<8 ? x : y;  v = g<4 ? y : (g==12 || g==14 ? x : z);  ±u ± v by bits 0 and 1.
__device__ __forceinline__ float gdot(uint32_t h, float x, float y, float z) {
  const uint32_t g = h >> 28;
  const float u = g < 8u ? x : y;
  const float v = g < 4u ? y : ((g == 12u || g == 14u) ? x : z);
  const float su = __uint_as_float(__float_as_uint(u) ^ ((g & 1u) << 31));
  const float sv = __uint_as_float(__float_as_uint(v) ^ ((g & 2u) << 30));
  return su + sv;
}

// Gradient vector of hash h, components in {-1, 0, 1}.
__device__ __forceinline__ void gvec(uint32_t h, float& gx, float& gy, float& gz) {
  const uint32_t g = h >> 28;
  const float su = (g & 1u) ? -1.0f : 1.0f;
  const float sv = (g & 2u) ? -1.0f : 1.0f;
  const bool ux = g < 8u;                       // u axis: x (else y)
  const bool vy = g < 4u;                       // v axis: y
  const bool vx = !vy && (g == 12u || g == 14u);  // v axis: x
  const bool vz = !vy && !vx;                   // v axis: z
  gx = (ux ? su : 0.0f) + (vx ? sv : 0.0f);
  gy = (ux ? 0.0f : su) + (vy ? sv : 0.0f);
  gz = vz ? sv : 0.0f;
}

// Corner hashes of the lattice cell (ix, iy, iz): h[c] for c = x + 2y + 4z.
__device__ __forceinline__ void cell_hashes(uint32_t seed, int ix, int iy, int iz, uint32_t h[8]) {
  const uint32_t x0 = (uint32_t)ix * kPX, x1 = x0 + kPX;
  const uint32_t y0 = (uint32_t)iy * kPY, y1 = y0 + kPY;
  const uint32_t z0 = (uint32_t)iz * kPZ, z1 = z0 + kPZ;
  const uint32_t s00 = seed ^ y0 ^ z0, s10 = seed ^ y1 ^ z0;
  const uint32_t s01 = seed ^ y0 ^ z1, s11 = seed ^ y1 ^ z1;
  h[0] = (s00 ^ x0) * kHM; h[1] = (s00 ^ x1) * kHM;
  h[2] = (s10 ^ x0) * kHM; h[3] = (s10 ^ x1) * kHM;
  h[4] = (s01 ^ x0) * kHM; h[5] = (s01 ^ x1) * kHM;
  h[6] = (s11 ^ x0) * kHM; h[7] = (s11 ^ x1) * kHM;
}

// ------------------------------------------------------------------ noise (values)
template <bool X = false>
__device__ __forceinline__ float noise_gradient(uint32_t seed, float qx, float qy, float qz) {
  const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
  uint32_t h[8];
  cell_hashes(seed, (int)fx, (int)fy, (int)fz, h);
  const float tx = qx - fx, ty = qy - fy, tz = qz - fz;
  const float sx = tx - 1.0f, sy = ty - 1.0f, sz = tz - 1.0f;
  const float ux = fade<X>(tx), uy = fade<X>(ty), uz = fade<X>(tz);
  const float x00 = lerpf<X>(gdot(h[0], tx, ty, tz), gdot(h[1], sx, ty, tz), ux);
  const float x10 = lerpf<X>(gdot(h[2], tx, sy, tz), gdot(h[3], sx, sy, tz), ux);
  const float x01 = lerpf<X>(gdot(h[4], tx, ty, sz), gdot(h[5], sx, ty, sz), ux);
  const float x11 = lerpf<X>(gdot(h[6], tx, sy, sz), gdot(h[7], sx, sy, sz), ux);
  return lerpf<X>(lerpf<X>(x00, x10, uy), lerpf<X>(x01, x11, uy), uz);
}

__device__ __forceinline__ float lattice_value(uint32_t h) { return (float)(int32_t)h * 4.656612873077393e-10f; }

template <bool X = false>
__device__ __forceinline__ float noise_value(uint32_t seed, float qx, float qy, float qz) {
  const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
  uint32_t h[8];
  cell_hashes(seed, (int)fx, (int)fy, (int)fz, h);
  const float ux = fade<X>(qx - fx), uy = fade<X>(qy - fy), uz = fade<X>(qz - fz);
  const float x00 = lerpf<X>(lattice_value(h[0]), lattice_value(h[1]), ux);
  const float x10 = lerpf<X>(lattice_value(h[2]), lattice_value(h[3]), ux);
  const float x01 = lerpf<X>(lattice_value(h[4]), lattice_value(h[5]), ux);
  const float x11 = lerpf<X>(lattice_value(h[6]), lattice_value(h[7]), ux);
  return lerpf<X>(lerpf<X>(x00, x10, uy), lerpf<X>(x01, x11, uy), uz);
}

// Cellular F1 over the 27 cells around rint(q); returns the squared distance and the
// offset (q - nearest point) for the gradient.
template <bool X = false>
__device__ __forceinline__ float cellular_d2(uint32_t seed, float qx, float qy, float qz, float js,
                                             float j0, float& ex, float& ey, float& ez) {
  const float cxf = rintf(qx), cyf = rintf(qy), czf = rintf(qz);
  const int cx = (int)cxf, cy = (int)cyf, cz = (int)czf;
  float best = __builtin_inff();
  ex = ey = ez = 0.0f;
  const uint32_t xb = (uint32_t)(cx - 1) * kPX, yb = (uint32_t)(cy - 1) * kPY, zb = (uint32_t)(cz - 1) * kPZ;
#pragma unroll
  for (int dz = 0; dz < 3; dz++) {
    const uint32_t zp = zb + (uint32_t)dz * kPZ;
    const float kz = (float)(cz - 1 + dz) - j0;
#pragma unroll
    for (int dy = 0; dy < 3; dy++) {
      const uint32_t yzs = seed ^ (yb + (uint32_t)dy * kPY) ^ zp;
      const float ky = (float)(cy - 1 + dy) - j0;
#pragma unroll
      for (int dx = 0; dx < 3; dx++) {
        const float kx = (float)(cx - 1 + dx) - j0;
        const uint32_t h = (yzs ^ (xb + (uint32_t)dx * kPX)) * kHM;
        const uint32_t m = h ^ (h >> 15);
        const float px = mad_x<X>((float)(m & 1023u), js, kx);
        const float py = mad_x<X>((float)((m >> 10) & 1023u), js, ky);
        const float pz = mad_x<X>((float)((m >> 20) & 1023u), js, kz);
        const float dxv = px - qx, dyv = py - qy, dzv = pz - qz;
        const float d2 = mad_x<X>(dzv, dzv, mad_x<X>(dyv, dyv, dxv * dxv));
        if (d2 < best) { best = d2; ex = -dxv; ey = -dyv; ez = -dzv; }
      }
    }
  }
  return best;
}

template <bool X = false>
__device__ __forceinline__ float noise_cellular(uint32_t seed, float qx, float qy, float qz, float js, float j0) {
  float ex, ey, ez;
  return sqrt_x<X>(cellular_d2<X>(seed, qx, qy, qz, js, j0, ex, ey, ez));
}

// 3D simplex noise (Perlin 2001, in the form of Gustavson's "Simplex noise demystified",
// 2005): skew by (x+y+z)/3 to find the cell, unskew by (i+j+k)/6, walk to the four
// corners of the containing simplex in the order of the offset components, and sum
// (0.5 - r^2)^4 (g . d) over the corners with r^2 < 0.5, times 72.  Radius^2 = 0.5 is
// the largest for which no other lattice point reaches into the simplex, so the noise is
// continuous; the corner gradient g and the hash are the lattice noise's.
struct SimplexCell {
  int32_t i, j, k;
  int32_t o1[3], o2[3];    // offsets of corners 1 and 2
  float d[4][3];           // position minus corner
};
__device__ __forceinline__ void simplex_cell(float x, float y, float z, SimplexCell& c) {
  const float F3 = 0.33333334f, G3 = 0.16666667f, G3x2 = 0.33333334f;
  const float s = (x + y + z) * F3;
  const float fi = floorf(x + s), fj = floorf(y + s), fk = floorf(z + s);
  const float t = (fi + fj + fk) * G3;
  const float x0 = x - (fi - t), y0 = y - (fj - t), z0 = z - (fk - t);
  c.i = (int32_t)fi; c.j = (int32_t)fj; c.k = (int32_t)fk;
  const bool xy = x0 >= y0, yz = y0 >= z0, xz = x0 >= z0;
  // corner 1: the largest offset component's axis; corner 2: the two largest
  c.o1[0] = (xy && xz) ? 1 : 0;
  c.o1[1] = (!xy && yz) ? 1 : 0;
  c.o1[2] = (!xz && !yz) ? 1 : 0;
  c.o2[0] = (xy || xz) ? 1 : 0;
  c.o2[1] = (!xy || yz) ? 1 : 0;
  c.o2[2] = (!xz || !yz) ? 1 : 0;
  c.d[0][0] = x0; c.d[0][1] = y0; c.d[0][2] = z0;
  c.d[1][0] = (x0 - (float)c.o1[0]) + G3; c.d[1][1] = (y0 - (float)c.o1[1]) + G3;
  c.d[1][2] = (z0 - (float)c.o1[2]) + G3;
  c.d[2][0] = (x0 - (float)c.o2[0]) + G3x2; c.d[2][1] = (y0 - (float)c.o2[1]) + G3x2;
  c.d[2][2] = (z0 - (float)c.o2[2]) + G3x2;
  c.d[3][0] = x0 - 0.5f; c.d[3][1] = y0 - 0.5f; c.d[3][2] = z0 - 0.5f;
}
__device__ __forceinline__ uint32_t simplex_hash(uint32_t seed, const SimplexCell& c, int corner) {
  const int32_t ox = corner == 0 ? 0 : (corner == 1 ? c.o1[0] : (corner == 2 ? c.o2[0] : 1));
  const int32_t oy = corner == 0 ? 0 : (corner == 1 ? c.o1[1] : (corner == 2 ? c.o2[1] : 1));
  const int32_t oz = corner == 0 ? 0 : (corner == 1 ? c.o1[2] : (corner == 2 ? c.o2[2] : 1));
  const uint32_t h = seed ^ ((uint32_t)(c.i + ox) * kPX) ^ ((uint32_t)(c.j + oy) * kPY) ^ ((uint32_t)(c.k + oz) * kPZ);
  return h * kHM;
}
__device__ __forceinline__ float noise_simplex(uint32_t seed, float x, float y, float z) {
  SimplexCell c;
  simplex_cell(x, y, z, c);
  float n = 0.0f;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const float dx = c.d[q][0], dy = c.d[q][1], dz = c.d[q][2];
    const float t = 0.5f - (dx * dx + dy * dy + dz * dz);
    if (t > 0.0f) {
      const float t2 = t * t;
      n = n + t2 * t2 * gdot(simplex_hash(seed, c, q), dx, dy, dz);
    }
  }
  return 72.0f * n;
}

// ------------------------------------------------------------------ shapes
// The further shapes (ids UPF_SHAPE_H3O and above): out of line, so that the field
// kernels carry one call instead of their code; parameters pass by value.
__device__ __noinline__ float shape_more(int shape, float p0, float p1, float p2, float p3, const float* w, float x,
                                         float y, float z) {
  if (shape == UPF_SHAPE_SINE_NET) return sine_net_t<float>(w, p0, p1, p2, x, y, z);
  const float sp[4] = {p0, p1, p2, p3};
  if (shape == UPF_SHAPE_ASTEROID) {  // sources/sdf.cpp:333-353
    const float l = sqrt_f(x * x + y * y + z * z);
    if (l < 1.0f) return -100.0f;
    const float dx = x / l, dy = y / l, dz = z / l;
    const uint32_t seed = (uint32_t)p1;
    const float b = 1.0f / 1.3f;  // amplitudes bounded by 1 / (1 + 0.3)
    const float n0 = noise_simplex(seed, dx * p0, dy * p0, dz * p0);
    const float f1 = p0 * 2.0f;
    const float n1 = noise_simplex(seed + 1u, dx * f1, dy * f1, dz * f1);
    const float n = n0 * b + n1 * (b * 0.3f);
    return l + n * 600.0f - 1300.0f;
  }
  return shape_more_t<float>(shape, sp, x, y, z, 0.0f);
}

template <bool EXT = true, bool COMPOSED = true, bool X = false>
__device__ __forceinline__ float shape_eval(const DSpec& s, float x, float y, float z) {
  switch (s.shape) {
    case UPF_SHAPE_SPHERE:
      return sqrt_x<X>(mad_x<X>(z, z, mad_x<X>(y, y, x * x))) - s.sp[0];
    case UPF_SHAPE_BOX: {
      const float qx = fabsf(x) - s.sp[0], qy = fabsf(y) - s.sp[1], qz = fabsf(z) - s.sp[2];
      const float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
      const float outside = sqrt_x<X>(mad_x<X>(mz, mz, mad_x<X>(my, my, mx * mx)));
      const float inside = fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
      return outside + inside - s.sp[3];
    }
    case UPF_SHAPE_TORUS: {
      const float ring = sqrt_x<X>(mad_x<X>(z, z, x * x)) - s.sp[0];
      return sqrt_x<X>(mad_x<X>(y, y, ring * ring)) - s.sp[1];
    }
    case UPF_SHAPE_CAPSULE: {
      const float h2 = s.sp[0] * 0.5f;
      const float zc = z - fminf(fmaxf(z, -h2), h2);
      return sqrt_x<X>(mad_x<X>(zc, zc, mad_x<X>(y, y, x * x))) - s.sp[1];
    }
    case UPF_SHAPE_NONE:
      return 0.0f;
    default:
      if (EXT && s.shape >= UPF_SHAPE_H3O)
        return shape_more(s.shape, s.sp[0], s.sp[1], s.sp[2], s.sp[3], s.w, x, y, z);
      if (COMPOSED) return shape_ext<float>(s.shape, s.sp, x, y, z, 0.0f);
      return 0.0f;  // not reached: the host launches COMPOSED = false for the primitives only
  }
}

// ------------------------------------------------------------------ layers
template <bool X = false>
__device__ __forceinline__ float octave_noise(const DLayer& L, int k, float x, float y, float z) {
  const float f = L.freq[k];
  const float qx = x * f, qy = y * f, qz = z * f;
  const uint32_t seed = L.seed + (uint32_t)k;
  if (L.noise == UPF_NOISE_GRADIENT) return noise_gradient<X>(seed, qx, qy, qz);
  if (L.noise == UPF_NOISE_VALUE) return noise_value<X>(seed, qx, qy, qz);
  if (L.noise == UPF_NOISE_SIMPLEX) return noise_simplex(seed, qx, qy, qz);
  return noise_cellular<X>(seed, qx, qy, qz, L.js, L.j0);
}

template <bool X = false>
__device__ __forceinline__ float layer_value(const DLayer& L, float x, float y, float z) {
  if (L.fractal == UPF_FRACTAL_NONE) return octave_noise<X>(L, 0, x, y, z);
  float acc = 0.0f;
  if (L.fractal == UPF_FRACTAL_RIDGED) {
    for (int k = 0; k < L.octaves; k++) {
      const float n = octave_noise<X>(L, k, x, y, z);
      acc = mad_x<X>(L.amp[k], mad_x<X>(-2.0f, fabsf(n), 1.0f), acc);
    }
  } else {
    for (int k = 0; k < L.octaves; k++) acc = mad_x<X>(L.amp[k], octave_noise<X>(L, k, x, y, z), acc);
  }
  return acc;
}

// Elevation: the generic layer stack, or a fixed mode (upf_modes.hpp) whose layers go
// through the same single layer_value call.  Layer 0 of the craters mode carries the
// Voronoi parameters (seed, points per cell in octaves, cell size in freq[0]).
// EXT = false compiles the modes and the further shapes out (the kernels' common
// instantiation); specs with either run the EXT = true instantiations.
template <bool EXT = true, bool X = false>
__device__ __forceinline__ float elevation_eval(const DSpec& s, float x, float y, float z) {
  const int mode = EXT ? s.mode : UPF_ELEV_LAYERS;
  ModeState<float> st;
  if (mode != UPF_ELEV_LAYERS) mode_begin(mode, s.L[0].seed, s.L[0].octaves, s.L[0].freq[0], x, y, z, 0.0f, st);
  float acc = 0.0f, m = 0.0f;
  for (int l = 0; l < s.n_layers; l++) {
    const DLayer& L = s.L[l];
    float px = x, py = y, pz = z;
    if (mode != UPF_ELEV_LAYERS) mode_position(mode, l, st, x, y, z, 0.0f, px, py, pz);
    const float raw = layer_value<X>(L, px, py, pz);
    if (mode != UPF_ELEV_LAYERS) {
      if (mode_after(mode, l, raw, 0.0f, st)) break;
      continue;
    }
    const float v = L.weight * raw;
    if (L.blend == UPF_BLEND_ADD) acc = acc + v;
    else if (L.blend == UPF_BLEND_MASK) m = __builtin_amdgcn_fmed3f(v + 0.5f, 0.0f, 1.0f);
    else acc = mad_x<X>(m, v - acc, acc);
  }
  if (mode != UPF_ELEV_LAYERS) return mode_finish(mode, st, x, y, z, 0.0f);
  return acc;
}

// X = EXT by default: the per-lane kernels of the elevation modes and the further shapes
// round like the CPU restatement (the modes' pow is upf_math.hpp's on both sides), so
// their fields equal it bit for bit; the common kernels keep the fused form.
template <bool EXT = true, bool X = EXT>
__device__ __forceinline__ float field_eval(const DSpec& s, float x, float y, float z) {
  const bool need_elev = s.compose == UPF_COMPOSE_LAND || s.compose == UPF_COMPOSE_NAVIGATION ||
                         s.compose == UPF_COMPOSE_ELEVATION_RAW;
  const float shape = s.compose == UPF_COMPOSE_ELEVATION_RAW ? 0.0f : shape_eval<EXT, true, X>(s, x, y, z);
  const float elev = need_elev ? elevation_eval<EXT, X>(s, x, y, z) : 0.0f;
  switch (s.compose) {
    case UPF_COMPOSE_LAND: return shape - s.escale * elev;
    case UPF_COMPOSE_WATER: return shape;
    case UPF_COMPOSE_NAVIGATION: return shape - s.escale * fmaxf(elev, 0.0f);
    case UPF_COMPOSE_ELEVATION: return shape / s.escale;
    default: return elev;
  }
}

}  // namespace upf
