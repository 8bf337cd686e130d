// upf_material.hpp -- the per-texel / per-tile material stack of the reference
// (coloringDefault / coloringBarren / coloringDebug, sources/coloringModes.cpp:905-966)
// on gfx950: elevation, precipitation, temperature, slope from 8 elevation probes,
// biome, type, then the layers bedrock, cliffs, mica, dirt, sand, grass, boulders,
// water, flowers, ice and snow, and the final saturation.
//
// The noise objects the reference configures are evaluated in place with their
// configurations written at the call (constants fold into the code).  Object k is seeded
// seed + 16 k and its octaves add 0..7 (the reference draws every seed from noiseSeed(),
// sources/math.cpp:73-77; objects that share `seed` in the reference share an index
// here).  cage's noise kinds map as follows (cage is not in the reference's checkout; the
// definitions are this library's, documented at upf_material_spec):
//   Perlin -> lattice gradient noise; Value -> lattice value noise; Cubic -> value noise
//   with cubic interpolation over the 4^3 lattice values around the sample; Simplex and
//   SimplexReduced -> simplex noise; Cellular -> F1 / F2 over the 27 cells around the
//   sample with a distance (Euclidean, Manhattan, Hybrid) and an operation (Cell value,
//   Subtract, Multiply, Divide).
// Everything is plain FP32 (no fused operations) and pow / atan are the library's own
// (upf_math.hpp): the CPU restatement rounds alike.
#pragma once

#include "upf_device.hpp"
#include "upf_modes.hpp"

namespace upf {

struct DMaterial {
  DSpec field;
  int32_t coloring, purpose, poles, flowers;
  uint32_t seed;
  int32_t pad[3];
};

// Tile (sources/planets.h:63-83), the fields the material stack reads and writes.
struct MTile {
  float px, py, pz, nx, ny, nz;
  float r, g, b;
  float roughness, metallic, height, elevation, slope, temperature, precipitation, opacity;
  int biome, type;
};

enum { kMNoPerlin = 0, kMNoValue = 1, kMNoCubic = 2, kMNoSimplex = 3, kMNoCellular = 4 };
enum { kDistEuclid = 0, kDistManhattan = 1, kDistHybrid = 2 };
enum { kRetCell = 0, kRetSub = 1, kRetMul = 2, kRetDiv = 3 };
enum { kBiomeBare = 0, kBiomeTundra, kBiomeTaiga, kBiomeShrubland, kBiomeGrassland, kBiomeTemperateSeasonal,
       kBiomeTemperateRain, kBiomeDesert, kBiomeSavanna, kBiomeTropicalSeasonal, kBiomeTropicalRain, kBiomeWater };
enum { kTypeRoad = 0, kTypeFlat = 1, kTypeRough = 2, kTypeCliffs = 3, kTypeWater = 4 };

// ------------------------------------------------------------------ base noises
__device__ __forceinline__ float m_lerp(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ float m_fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
__device__ __forceinline__ uint32_t m_hash(uint32_t seed, int32_t x, int32_t y, int32_t z) {
  return (seed ^ ((uint32_t)x * kPX) ^ ((uint32_t)y * kPY) ^ ((uint32_t)z * kPZ)) * kHM;
}
__device__ __forceinline__ float m_lattice(uint32_t h) { return (float)(int32_t)h * 4.656612873077393e-10f; }

__device__ __forceinline__ float m_perlin(uint32_t seed, float x, float y, float z) {
  const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  const float tx = x - fx, ty = y - fy, tz = z - fz;
  const float sx = tx - 1.0f, sy = ty - 1.0f, sz = tz - 1.0f;
  const float ux = m_fade(tx), uy = m_fade(ty), uz = m_fade(tz);
  const float n000 = gdot(m_hash(seed, ix, iy, iz), tx, ty, tz);
  const float n100 = gdot(m_hash(seed, ix + 1, iy, iz), sx, ty, tz);
  const float n010 = gdot(m_hash(seed, ix, iy + 1, iz), tx, sy, tz);
  const float n110 = gdot(m_hash(seed, ix + 1, iy + 1, iz), sx, sy, tz);
  const float n001 = gdot(m_hash(seed, ix, iy, iz + 1), tx, ty, sz);
  const float n101 = gdot(m_hash(seed, ix + 1, iy, iz + 1), sx, ty, sz);
  const float n011 = gdot(m_hash(seed, ix, iy + 1, iz + 1), tx, sy, sz);
  const float n111 = gdot(m_hash(seed, ix + 1, iy + 1, iz + 1), sx, sy, sz);
  const float y0 = m_lerp(m_lerp(n000, n100, ux), m_lerp(n010, n110, ux), uy);
  const float y1 = m_lerp(m_lerp(n001, n101, ux), m_lerp(n011, n111, ux), uy);
  return m_lerp(y0, y1, uz);
}

__device__ __forceinline__ float m_value(uint32_t seed, float x, float y, float z) {
  const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  const float ux = m_fade(x - fx), uy = m_fade(y - fy), uz = m_fade(z - fz);
  float c[2][2][2];
#pragma unroll
  for (int k = 0; k < 2; k++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int i = 0; i < 2; i++) c[k][j][i] = m_lattice(m_hash(seed, ix + i, iy + j, iz + k));
  const float y0 = m_lerp(m_lerp(c[0][0][0], c[0][0][1], ux), m_lerp(c[0][1][0], c[0][1][1], ux), uy);
  const float y1 = m_lerp(m_lerp(c[1][0][0], c[1][0][1], ux), m_lerp(c[1][1][0], c[1][1][1], ux), uy);
  return m_lerp(y0, y1, uz);
}

// cubic through b (t = 0) and c (t = 1) with neighbours a, d: t^3 p + t^2 ((a - b) - p)
// + t (c - a) + b, p = (d - c) - (a - b)
__device__ __forceinline__ float m_cubic(float a, float b, float c, float d, float t) {
  const float p = (d - c) - (a - b);
  return t * t * t * p + t * t * ((a - b) - p) + t * (c - a) + b;
}

__device__ __forceinline__ float m_value_cubic(uint32_t seed, float x, float y, float z) {
  const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  const float tx = x - fx, ty = y - fy, tz = z - fz;
  float zs[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    float ys[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      float xs[4];
#pragma unroll
      for (int i = 0; i < 4; i++) xs[i] = m_lattice(m_hash(seed, ix + i - 1, iy + j - 1, iz + k - 1));
      ys[j] = m_cubic(xs[0], xs[1], xs[2], xs[3], tx);
    }
    zs[k] = m_cubic(ys[0], ys[1], ys[2], ys[3], ty);
  }
  return m_cubic(zs[0], zs[1], zs[2], zs[3], tz) * (1.0f / 3.375f);
}

// Cellular over the 27 cells around rint(q), one point per cell at k - 0.5 + r (r from
// three 10-bit fields of h ^ (h >> 15), scaled by 1/1023): F1 d0, F2 d1 and the hash of
// the nearest cell, then the operation.
__device__ __forceinline__ float m_cellular(uint32_t seed, float x, float y, float z, int dist, int ret) {
  const int cx = (int)rintf(x), cy = (int)rintf(y), cz = (int)rintf(z);
  float d0 = __builtin_inff(), d1 = __builtin_inff();
  uint32_t best = 0;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
      for (int dx = -1; dx <= 1; dx++) {
        const int kx = cx + dx, ky = cy + dy, kz = cz + dz;
        const uint32_t h = m_hash(seed, kx, ky, kz);
        const uint32_t m = h ^ (h >> 15);
        const float px = ((float)kx - 0.5f) + (float)(m & 1023u) * (1.0f / 1023.0f);
        const float py = ((float)ky - 0.5f) + (float)((m >> 10) & 1023u) * (1.0f / 1023.0f);
        const float pz = ((float)kz - 0.5f) + (float)((m >> 20) & 1023u) * (1.0f / 1023.0f);
        const float ex = px - x, ey = py - y, ez = pz - z;
        float d;
        if (dist == kDistManhattan) d = fabsf(ex) + fabsf(ey) + fabsf(ez);
        else {
          const float e2 = ex * ex + ey * ey + ez * ez;
          d = dist == kDistHybrid ? (fabsf(ex) + fabsf(ey) + fabsf(ez)) + e2 : e2;
        }
        if (d < d0) { d1 = d0; d0 = d; best = h; }
        else if (d < d1) d1 = d;
      }
  if (ret == kRetCell) return m_lattice(best);
  if (dist == kDistEuclid) { d0 = sqrtf(d0); d1 = sqrtf(d1); }
  if (ret == kRetSub) return d1 - d0 - 1.0f;
  if (ret == kRetMul) return d1 * d0 * 0.5f - 1.0f;
  return d0 / d1 - 1.0f;
}

// One configured noise object: kind, fractal (0 none, 1 fBm, 2 ridged), octaves,
// frequency, lacunarity, gain (amplitudes bounded by 1 / sum |gain|^k), cellular
// distance and operation.  Octave k uses seed + k.
__device__ __noinline__ float mnoise(int kind, uint32_t seed, float freq, int fractal, int octaves, float lac,
                                     float gain, int dist, int ret, float x, float y, float z) {
  const int oct = fractal == 0 ? 1 : octaves;
  float bound = 1.0f;
  if (fractal != 0) {
    const float ag = fabsf(gain);
    float s = 1.0f, a = ag;
    for (int i = 1; i < oct; i++) { s = s + a; a = a * ag; }
    bound = 1.0f / s;
  }
  float sum = 0.0f, amp = bound, f = freq;
  for (int k = 0; k < oct; k++) {
    const float qx = x * f, qy = y * f, qz = z * f;
    const uint32_t sd = seed + (uint32_t)k;
    float n;
    switch (kind) {
      case kMNoPerlin: n = m_perlin(sd, qx, qy, qz); break;
      case kMNoValue: n = m_value(sd, qx, qy, qz); break;
      case kMNoCubic: n = m_value_cubic(sd, qx, qy, qz); break;
      case kMNoSimplex: n = noise_simplex(sd, qx, qy, qz); break;
      default: n = m_cellular(sd, qx, qy, qz, dist, ret); break;
    }
    if (fractal == 0) return n;
    if (fractal == 2) sum = sum + (1.0f - 2.0f * fabsf(n)) * amp;
    else sum = sum + n * amp;
    amp = amp * gain;
    f = f * lac;
  }
  return sum;
}

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ float m_sat(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float m_smoothstep(float x) { return x * x * (3.0f - 2.0f * x); }
__device__ __forceinline__ float m_smootherstep(float x) { return x * x * x * (x * (x * 6.0f - 15.0f) + 10.0f); }
__device__ __forceinline__ float m_rescale(float v, float ia, float ib, float oa, float ob) {
  return (v - ia) / (ib - ia) * (ob - oa) + oa;
}
__device__ __forceinline__ float m_sharp_edge(float v, float p = 0.05f) {  // sources/math.cpp:13-16
  return m_rescale(fminf(fmaxf(v, 0.5f - p), 0.5f + p), 0.5f - p, 0.5f + p, 0.0f, 1.0f);
}
__device__ __forceinline__ float m_range_mask(float v, float zero_at, float one_at) {  // :29-32
  return m_sat((v - zero_at) / (one_at - zero_at));
}
constexpr float kRadToDeg = 57.29577951308232f;
constexpr float kPi = 3.14159265358979f;
// slope in radians, thresholds in degrees (:22-27)
__device__ __forceinline__ float m_steepness(float slope_rad, float threshold, float smoothing) {
  return m_sat((threshold + smoothing - slope_rad * kRadToDeg) / (2.0f * smoothing));
}

// RGB <-> HSV (h, s, v in [0, 1]), the textbook hexcone model.
__device__ __forceinline__ void m_rgb_to_hsv(float r, float g, float b, float& h, float& s, float& v) {
  const float mx = fmaxf(r, fmaxf(g, b)), mn = fminf(r, fminf(g, b));
  const float d = mx - mn;
  v = mx;
  s = mx > 0.0f ? d / mx : 0.0f;
  if (d <= 0.0f) { h = 0.0f; return; }
  float hh;
  if (mx == r) hh = (g - b) / d + (g < b ? 6.0f : 0.0f);
  else if (mx == g) hh = (b - r) / d + 2.0f;
  else hh = (r - g) / d + 4.0f;
  h = hh / 6.0f;
}
__device__ __forceinline__ void m_hsv_to_rgb(float h, float s, float v, float& r, float& g, float& b) {
  const float h6 = (h - floorf(h)) * 6.0f;
  const int i = min((int)h6, 5);
  const float f = h6 - (float)i;
  const float p = v * (1.0f - s), q = v * (1.0f - s * f), t = v * (1.0f - s * (1.0f - f));
  switch (i) {
    case 0: r = v; g = t; b = p; break;
    case 1: r = q; g = v; b = p; break;
    case 2: r = p; g = v; b = t; break;
    case 3: r = p; g = q; b = v; break;
    case 4: r = t; g = p; b = v; break;
    default: r = v; g = p; b = q; break;
  }
}
// colorHueShift, sources/math.cpp:47-52
__device__ __forceinline__ void m_hue_shift(float& r, float& g, float& b, float shift) {
  float h, s, v;
  m_rgb_to_hsv(r, g, b, h, s, v);
  h = fmodf(h + shift + 1.0f, 1.0f);
  m_hsv_to_rgb(h, s, v, r, g, b);
}
// interpolateColor: HSV with the hue along the shorter arc
__device__ __forceinline__ void m_mix_color(float ar, float ag, float ab, float br, float bg, float bb, float t,
                                            float& r, float& g, float& b) {
  float h0, s0, v0, h1, s1, v1;
  m_rgb_to_hsv(ar, ag, ab, h0, s0, v0);
  m_rgb_to_hsv(br, bg, bb, h1, s1, v1);
  float dh = h1 - h0;
  if (dh > 0.5f) dh = dh - 1.0f;
  else if (dh < -0.5f) dh = dh + 1.0f;
  const float h = fmodf(h0 + t * dh + 1.0f, 1.0f);
  m_hsv_to_rgb(h, m_lerp(s0, s1, t), m_lerp(v0, v1, t), r, g, b);
}
// colorDeviation (sources/math.cpp:40-45) without the global RNG: the deviation vector
// (three 10-bit fractions minus 0.5) is hashed from the position's float bits, and the
// colour moves in HSV.
__device__ __forceinline__ void m_deviation(float& r, float& g, float& b, float dev, uint32_t seed, float px,
                                            float py, float pz) {
  const uint32_t hsh = fmix32(m_hash(seed, (int32_t)__float_as_uint(px), (int32_t)__float_as_uint(py),
                                     (int32_t)__float_as_uint(pz)));
  const float u0 = (float)(hsh & 1023u) * (1.0f / 1023.0f) - 0.5f;
  const float u1 = (float)((hsh >> 10) & 1023u) * (1.0f / 1023.0f) - 0.5f;
  const float u2 = (float)((hsh >> 20) & 1023u) * (1.0f / 1023.0f) - 0.5f;
  float h, s, v;
  m_rgb_to_hsv(r, g, b, h, s, v);
  h = fmodf(h + u0 * dev + 1.0f, 1.0f);
  s = m_sat(s + u1 * dev);
  v = m_sat(v + u2 * dev);
  m_hsv_to_rgb(h, s, v, r, g, b);
}

// Voronoi nearest with the points projected into the plane through p with normal n
// (sources/voronoi.cpp:58-66; closestPoint(plane, q) = q - dot(q - p, n) n).
__device__ __forceinline__ void m_voronoi_plane(uint32_t seed, int ppc, float cell, float px, float py, float pz, float nx,
                                float ny, float nz, float c[3]) {
  const uint32_t h0 = fmix32(seed), h1 = fmix32(h0), h2 = fmix32(h1);
  const float freq = 1.0f / cell;
  const int32_t cx = cell_index(px * freq), cy = cell_index(py * freq), cz = cell_index(pz * freq);
  float best = __builtin_inff();
  c[0] = c[1] = c[2] = 0.0f;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const int32_t kx = cx + dx, ky = cy + dy, kz = cz + dz;
        uint32_t s0 = (uint32_t)kx, s1 = (uint32_t)ky, s2 = (uint32_t)kz;
        for (int i = 0; i <= ppc; i++) {
          const uint32_t t0 = h0 + s0 + s1 + fmix32(s0);
          const uint32_t t1 = h1 + s1 + s2 + fmix32(s1);
          const uint32_t t2 = h2 + s2 + s0 + fmix32(s2);
          s0 = t0; s1 = t1; s2 = t2;
          if (i == 0) continue;
          float qx = ((float)((int32_t)s0 % 65536) / 65535.0f + (float)kx) * cell;
          float qy = ((float)((int32_t)s1 % 65536) / 65535.0f + (float)ky) * cell;
          float qz = ((float)((int32_t)s2 % 65536) / 65535.0f + (float)kz) * cell;
          const float dd = (qx - px) * nx + (qy - py) * ny + (qz - pz) * nz;
          qx = qx - dd * nx; qy = qy - dd * ny; qz = qz - dd * nz;
          const float ex = qx - px, ey = qy - py, ez = qz - pz;
          const float d2 = ex * ex + ey * ey + ez * ez;
          if (d2 < best) { best = d2; c[0] = qx; c[1] = qy; c[2] = qz; }
        }
      }
}

__device__ __forceinline__ float m_dist(float ax, float ay, float az, float bx, float by, float bz) {
  const float dx = ax - bx, dy = ay - by, dz = az - bz;
  return sqrtf(dx * dx + dy * dy + dz * dz);
}
__device__ __forceinline__ void m_blend(MTile& t, float r, float g, float b, float rough, float metal, float height,
                                        float bf) {
  t.r = m_lerp(t.r, r, bf); t.g = m_lerp(t.g, g, bf); t.b = m_lerp(t.b, b, bf);
  t.roughness = m_lerp(t.roughness, rough, bf);
  t.metallic = m_lerp(t.metallic, metal, bf);
  t.height = m_lerp(t.height, height, bf);
}

// ------------------------------------------------------------------ the stack
// seed of noise object k: octaves use seed + 16 k + octave, so objects never share seeds
__device__ __forceinline__ uint32_t ms(const DMaterial& m, uint32_t k) { return m.seed + 16u * k; }

// The base shape with separately rounded operations (the field kernels fuse them): the
// slope below differences eight probes 0.5 m apart at ~10^3 m from the origin, where the
// probes differ by little more than rounding, so the CPU restatement rounds alike.
__device__ __forceinline__ float m_shape(const DSpec& s, float x, float y, float z) {
  const float* a = s.sp;
  switch (s.shape) {
    case UPF_SHAPE_SPHERE: return sqrtf(x * x + y * y + z * z) - a[0];
    case UPF_SHAPE_BOX: {
      const float qx = fabsf(x) - a[0], qy = fabsf(y) - a[1], qz = fabsf(z) - a[2];
      const float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
      return sqrtf(mx * mx + my * my + mz * mz) + fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f) - a[3];
    }
    case UPF_SHAPE_TORUS: {
      const float ring = sqrtf(x * x + z * z) - a[0];
      return sqrtf(ring * ring + y * y) - a[1];
    }
    case UPF_SHAPE_CAPSULE: {
      const float h2 = a[0] * 0.5f;
      const float zc = z - fminf(fmaxf(z, -h2), h2);
      return sqrtf(x * x + y * y + zc * zc) - a[1];
    }
    case UPF_SHAPE_NONE: return 0.0f;
    default:
      if (s.shape >= UPF_SHAPE_H3O) return shape_more(s.shape, s.sp[0], s.sp[1], s.sp[2], s.sp[3], s.w, x, y, z);
      return shape_ext<float>(s.shape, s.sp, x, y, z, 0.0f);
  }
}

// terrainSdfElevation = shape * 10 (sources/modesConfiguration.cpp:230-237), with this
// spec's ratio 1 / elevation_scale
__device__ __forceinline__ float m_elev_sdf(const DMaterial& m, float x, float y, float z) {
  return m_shape(m.field, x, y, z) / m.field.escale;
}

__device__ __forceinline__ float m_beach(const DMaterial& m, const MTile& t) {  // :34-49
  if (t.biome == kBiomeWater) return 0.0f;
  const float n = mnoise(kMNoValue, ms(m, 0), 0.035f, 0, 1, 2.0f, 0.5f, 0, 0, t.elevation, 0.0f, 0.0f);
  return m_range_mask(t.elevation + n * 20.0f, 10.0f, 15.0f);
}

__device__ __forceinline__ void m_precipitation(const DMaterial& m, MTile& t) {  // :59-80
  float p = mnoise(kMNoCubic, ms(m, 1), 0.001f, 1, 2, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.5f + 0.5f;
  p = m_sat(p);
  p = m_smootherstep(p);
  p = m_smootherstep(p);
  p = upf_math::powf_(fmaxf(p + 0.2f, 0.0f), 3.0f) * 0.6f;
  p = p * 400.0f;
  p = p + fmaxf(120.0f - fabsf(t.elevation), 0.0f) * 0.2f;
  t.precipitation = fmaxf(p, 0.0f);
}

__device__ __forceinline__ void m_temperature(const DMaterial& m, MTile& t) {  // :82-120
  float v = mnoise(kMNoSimplex, ms(m, 2), 0.0003f, 1, 2, 2.0f, 0.4f, 0, 0, t.px, t.py, t.pz) * 0.5f + 0.5f;
  v = m_sat(v);
  v = m_smoothstep(v);
  v = v * 2.0f - 1.0f;
  if (m.poles) {
    float polar = fabsf(upf_math::atanf_(t.py / sqrtf(t.px * t.px + t.pz * t.pz))) / kPi * 2.0f;
    polar = upf_math::powf_(polar, 1.7f);
    polar = polar + mnoise(kMNoValue, ms(m, 3), 0.007f, 1, 3, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.1f;
    v = v + (0.6f - polar * 3.2f);
  }
  t.temperature = 15.0f + v * 30.0f - fmaxf(t.elevation, 0.0f) * 0.02f;
}

__device__ __forceinline__ void m_slope(const DMaterial& m, MTile& t) {  // :122-151
  // anyPerpendicular (sources/math.cpp:65-72)
  float bx = 1.0f, by = 0.0f, bz = 0.0f;
  if (fabsf(t.nx) > 0.9f) { bx = 0.0f; by = 1.0f; }
  float ax = t.ny * bz - t.nz * by, ay = t.nz * bx - t.nx * bz, az = t.nx * by - t.ny * bx;
  const float al = sqrtf(ax * ax + ay * ay + az * az);
  ax = ax / al; ay = ay / al; az = az / al;
  float cx = t.ny * az - t.nz * ay, cy = t.nz * ax - t.nx * az, cz = t.nx * ay - t.ny * ax;  // b = n x a
  const float radius = 0.5f;
  ax = ax * radius; ay = ay * radius; az = az * radius;
  cx = cx * radius; cy = cy * radius; cz = cz * radius;
  const float div = 1.0f / sqrtf(2.0f);
  const float ux = (ax + cx) * div, uy = (ay + cy) * div, uz = (az + cz) * div;
  const float vx = (ax - cx) * div, vy = (ay - cy) * div, vz = (az - cz) * div;
  const float e[8] = {m_elev_sdf(m, t.px + ax, t.py + ay, t.pz + az), m_elev_sdf(m, t.px + cx, t.py + cy, t.pz + cz),
                      m_elev_sdf(m, t.px - ax, t.py - ay, t.pz - az), m_elev_sdf(m, t.px - cx, t.py - cy, t.pz - cz),
                      m_elev_sdf(m, t.px + ux, t.py + uy, t.pz + uz), m_elev_sdf(m, t.px + vx, t.py + vy, t.pz + vz),
                      m_elev_sdf(m, t.px - ux, t.py - uy, t.pz - uz), m_elev_sdf(m, t.px - vx, t.py - vy, t.pz - vz)};
  float e1 = e[0], e2 = e[0];
#pragma unroll
  for (int i = 0; i < 8; i++) { e1 = fminf(e1, e[i]); e2 = fmaxf(e2, e[i]); }
  const float md = (e2 - e1) * 0.1f;
  t.slope = upf_math::atanf_(md / radius);
}

__device__ __forceinline__ void m_biome(MTile& t) {  // :153-190
  const float T = t.temperature, P = t.precipitation;
  if (t.elevation < 0.0f) t.biome = kBiomeWater;
  else if (T > 20.0f)
    t.biome = P < 40.0f ? kBiomeDesert : (P < 130.0f ? kBiomeSavanna : (P < 230.0f ? kBiomeTropicalSeasonal : kBiomeTropicalRain));
  else if (T > 5.0f)
    t.biome = P < 30.0f ? kBiomeShrubland : (P < 100.0f ? kBiomeGrassland : (P < 200.0f ? kBiomeTemperateSeasonal : kBiomeTemperateRain));
  else if (T > -5.0f) t.biome = P < 30.0f ? kBiomeBare : kBiomeTaiga;
  else t.biome = P < 10.0f ? kBiomeBare : kBiomeTundra;
}

__device__ __forceinline__ void m_type(MTile& t) {  // :192-200
  if (t.elevation < 0.0f) t.type = kTypeWater;
  else if (t.slope > 20.0f / kRadToDeg) t.type = kTypeCliffs;
  else t.type = kTypeFlat;
}

__device__ __forceinline__ void m_bedrock(const DMaterial& m, MTile& t) {  // :202-261
  const float dn = mnoise(kMNoCubic, ms(m, 5), 0.0031f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.5f + 0.501f;
  const float depth = dn * dn * 2.0f;
  const float freq = mnoise(kMNoCubic, ms(m, 6), 0.013f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.15f + 1.0f;
  const float fx = t.px * freq, fy = t.py * freq, fz = t.pz * freq;
  const float cr = mnoise(kMNoCellular, ms(m, 4), 0.017f, 0, 1, 2.0f, 0.5f, kDistHybrid, kRetSub, fx, fy, fz);
  const float cracks = m_sat(upf_math::powf_(cr * 0.5f + 0.5f, 0.7f));
  const float value = mnoise(kMNoCellular, ms(m, 4), 0.017f, 0, 1, 2.0f, 0.5f, kDistHybrid, kRetCell, fx, fy, fz);
  const float sat = m_sat(mnoise(kMNoCubic, ms(m, 7), 0.0022f, 1, 2, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.35f + 0.3f);
  m_hsv_to_rgb(0.07f, sat, m_sat(value * 0.1f + 0.7f) * (cracks * 0.4f + 0.6f), t.r, t.g, t.b);
  t.roughness = m_lerp(0.9f, value * -0.15f + 0.65f, cracks);
  t.height = cracks * depth;
}

__device__ __forceinline__ void m_cliffs(const DMaterial& m, MTile& t) {  // :263-293
  const float bf = m_steepness(t.slope, 19.0f, 4.0f);
  if (bf > 0.99999f) return;
  float h, s, v;
  m_rgb_to_hsv(t.r, t.g, t.b, h, s, v);
  h = m_lerp(155.0f / 255.0f, h, m_sharp_edge(bf, 0.005f));
  s = s * (bf * 0.4f + 0.6f);
  v = m_lerp(0.7f, s, bf * 0.8f + 0.2f);
  m_hsv_to_rgb(h, s, v, t.r, t.g, t.b);
  t.roughness = t.roughness * (bf * 0.2f + 0.8f);
  const float cracks = m_sharp_edge(
      m_sat(mnoise(kMNoSimplex, ms(m, 8), 0.047f, 2, 2, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 2.0f - 1.1f), 0.2f);
  t.height = t.height * m_lerp(1.0f, 0.8f, cracks);
  const float k = m_lerp(1.0f, 0.9f, cracks);
  t.r = t.r * k; t.g = t.g * k; t.b = t.b * k;
  t.roughness = m_lerp(t.roughness, 0.9f, cracks);
}

__device__ __forceinline__ void m_mica(const DMaterial& m, MTile& t) {  // :295-336
  if (!m.flowers) return;
  const float bf =
      m_sat((mnoise(kMNoCellular, ms(m, 9), 0.02f, 0, 1, 2.0f, 0.5f, kDistHybrid, kRetMul, t.px, t.py, t.pz) - 0.98f) * 10.0f);
  if (bf < 1e-7f) return;
  const float cracks = m_sharp_edge(
      m_sat(mnoise(kMNoCellular, ms(m, 10), 0.05f, 0, 1, 2.0f, 0.5f, kDistManhattan, kRetDiv, t.px, t.py, t.pz) + 0.6f));
  const float r = m_lerp(122.0f / 255.0f, 184.0f / 255.0f, cracks);
  const float g = m_lerp(90.0f / 255.0f, 209.0f / 255.0f, cracks);
  const float b = m_lerp(88.0f / 255.0f, 187.0f / 255.0f, cracks);
  m_blend(t, r, g, b, cracks * 0.3f + 0.4f, 1.0f, t.height * 0.5f, bf);
}

__device__ __forceinline__ void m_dirt(const DMaterial& m, MTile& t) {  // :338-402
  float height = mnoise(kMNoPerlin, ms(m, 11), 0.05f, 1, 3, 2.3f, 0.4f, 0, 0, t.px, t.py, t.pz) * 0.2f + 0.5f;
  const float bf = m_sharp_edge(m_sat(height - t.height + 0.4f)) * m_steepness(t.slope, 20.0f, 5.0f);
  if (bf < 1e-7f) return;
  float r = 168.0f / 255.0f, g = 94.0f / 255.0f, b = 28.0f / 255.0f;
  {
    const float saturation = m_range_mask(t.precipitation, 0.0f, 50.0f);
    float h, s, v;
    m_rgb_to_hsv(r, g, b, h, s, v);
    s = s * (saturation * 0.7f);
    v = v * 0.8f;
    m_hsv_to_rgb(h, s, v, r, g, b);
  }
  float rough = 0.8f;
  float cracks = m_sharp_edge(
      m_sat(mnoise(kMNoSimplex, ms(m, 12), 0.07f, 2, 2, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 2.0f - 1.2f), 0.15f);
  const float cm = m_smoothstep(
      m_sat(mnoise(kMNoCubic, ms(m, 13), 0.02f, 1, 4, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.5f + 0.5f));
  cracks = cracks * (cm * cm * 0.9f + 0.1f);
  cracks = cracks * (m_range_mask(t.precipitation, 70.0f, 20.0f) * 0.9f + 0.1f);
  height = m_lerp(height, height * 0.5f, cracks);
  r = m_lerp(r, 0.1f, cracks); g = m_lerp(g, 0.1f, cracks); b = m_lerp(b, 0.1f, cracks);
  rough = m_lerp(rough, 0.9f, cracks);
  m_blend(t, r, g, b, rough, 0.0f, height, bf);
}

__device__ __forceinline__ void m_sand(const DMaterial& m, MTile& t) {  // :404-465
  const float bf = m_range_mask(t.temperature, 24.0f, 28.0f) * m_steepness(t.slope, 19.0f, 10.0f);
  if (bf < 1e-7f) return;
  const float hs = mnoise(kMNoSimplex, ms(m, 14), 0.001f, 1, 3, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.3f + 1.0f;
  const float height =
      (mnoise(kMNoSimplex, ms(m, 15), 0.01f, 2, 3, 2.0f, 0.7f, 0, 0, t.px * hs, t.py * hs, t.pz * hs) * 0.2f) *
          (m_range_mask(t.precipitation, 100.0f, 50.0f) * 0.4f + 0.6f) +
      0.5f;
  const float cs = m_smootherstep(m_smootherstep(
      mnoise(kMNoPerlin, ms(m, 16), 0.0008f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.5f + 0.5f));
  const float hue = mnoise(kMNoCubic, ms(m, 17), 0.02f, 1, 2, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.1f;
  float r = 189.0f / 255.0f, g = 174.0f / 255.0f, b = 152.0f / 255.0f;
  m_hue_shift(r, g, b, hue);
  m_mix_color(r, g, b, 187.0f / 255.0f, 104.0f / 255.0f, 50.0f / 255.0f, cs, r, g, b);
  m_deviation(r, g, b, 0.08f, ms(m, 42), t.px, t.py, t.pz);
  {
    float h, s, v;
    m_rgb_to_hsv(r, g, b, h, s, v);
    m_hsv_to_rgb(h * 1.0f, s * 0.7f, v * 0.8f, r, g, b);
  }
  m_blend(t, r, g, b, (hs - 1.0f) * 0.7f + 0.55f + hue, 0.0f, height, bf);
}

__device__ __forceinline__ void m_grass(const DMaterial& m, MTile& t) {  // :467-533
  const float bf = m_range_mask(t.temperature, 35.0f, 25.0f) * m_range_mask(t.precipitation, 15.0f, 35.0f) *
                   m_steepness(t.slope, 22.0f, 3.0f) * m_beach(m, t);
  if (bf < 1e-7f) return;
  const float dry = fminf(fmaxf(t.temperature - (t.precipitation + 100.0f) * 30.0f / 400.0f, 0.0f), 5.0f) / 5.0f;
  const float hsb = mnoise(kMNoPerlin, ms(m, 19), 0.01f, 1, 2, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz);
  const float hue = hsb * 0.09f - dry * 0.1f;
  float r = 114.0f / 255.0f, g = 187.0f / 255.0f, b = 77.0f / 255.0f;
  m_hue_shift(r, g, b, hue);
  const float pm = m_sharp_edge(
      mnoise(kMNoCellular, ms(m, 18), 0.1f, 0, 1, 2.0f, 0.5f, kDistHybrid, kRetCell, t.px, t.py, t.pz) * 0.5f + 0.2f);
  const float patches =
      upf_math::powf_(mnoise(kMNoCellular, ms(m, 18), 0.1f, 0, 1, 2.0f, 0.5f, kDistHybrid, kRetSub, t.px, t.py, t.pz) * 0.5f + 0.5f,
           0.5f) *
      pm * (1.0f - dry * 0.8f);
  const float k1 = 1.0f - patches * 0.09f;
  r = r * k1; g = g * k1; b = b * k1;
  const float scr =
      m_sat(mnoise(kMNoCellular, ms(m, 20), 0.3f, 0, 1, 2.0f, 0.5f, kDistHybrid, kRetMul, t.px, t.py, t.pz) - 0.2f) *
      (1.0f - dry * 0.5f);
  const float k2 = 1.0f - scr * 0.1f;
  r = r * k2; g = g * k2; b = b * k2;
  const float rough = 0.6f + dry * 0.1f + patches * 0.1f + scr * 0.1f - hsb * 0.1f;
  const float height = m_lerp(t.height, 0.5f, 0.7f) + (0.5f - dry) * 0.2f + patches * 0.05f - scr * 0.02f + hsb * 0.1f + 0.1f;
  m_blend(t, r, g, b, rough, 0.0f, height, bf);
}

__device__ __forceinline__ void m_boulders(const DMaterial& m, MTile& t) {  // :535-615
  if (!m.flowers) return;
  if (mnoise(kMNoPerlin, ms(m, 21), 0.01f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) < 0.15f) return;
  float c[3];
  m_voronoi_plane(ms(m, 22), 2, 200.0f, t.px, t.py, t.pz, t.nx, t.ny, t.nz, c);
  const float dist = m_dist(c[0], c[1], c[2], t.px, t.py, t.pz);
  const float size = m_smootherstep(m_smootherstep(
                         m_sat(mnoise(kMNoPerlin, ms(m, 23), 0.09f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.5f + 0.5f))) *
                         3.0f +
                     3.0f;
  const float bf = m_range_mask(size - dist, 0.0f, 0.5f);
  if (bf < 1e-7f) return;
  const float hue = mnoise(kMNoPerlin, ms(m, 24), 0.3f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.12f;
  const float val = mnoise(kMNoPerlin, ms(m, 25), 0.5f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.18f;
  float h, s, v;
  m_rgb_to_hsv(154.0f / 255.0f, 154.0f / 255.0f, 154.0f / 255.0f, h, s, v);
  h = fmodf(h + hue + 1.0f, 1.0f);
  v = m_sat(v + val);
  float r, g, b;
  m_hsv_to_rgb(h, s, v, r, g, b);
  const float rough = mnoise(kMNoPerlin, ms(m, 26), 0.5f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.1f + 0.4f;
  const float q = dist / size;
  m_blend(t, r, g, b, rough, 0.0f, 1.0f - q * q * 0.5f, bf);
}

__device__ __forceinline__ void m_water(const DMaterial& m, MTile& t) {  // :617-656
  const float sw = mnoise(kMNoPerlin, ms(m, 27), 0.0015f, 1, 2, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.5f + 0.5f;
  const float hue = mnoise(kMNoCubic, ms(m, 28), 0.004f, 1, 3, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.09f;
  m_mix_color(40.0f / 255.0f, 63.0f / 255.0f, 168.0f / 255.0f, 39.0f / 255.0f, 154.0f / 255.0f, 186.0f / 255.0f, sw,
              t.r, t.g, t.b);
  m_hue_shift(t.r, t.g, t.b, hue);
  const float a = m_range_mask(t.elevation, -25.0f, 3.0f);
  const float d1 = 1.0f - a * a;
  const float d2 = m_rescale(m_range_mask(t.elevation, 0.0f, -100.0f), 0.0f, 1.0f, 0.7f, 1.0f);
  t.opacity = d1 * d2;
  t.metallic = m_smoothstep(m_range_mask(t.elevation, 3.0f, -30.0f));
}

__device__ __forceinline__ void m_flowers(const DMaterial& m, MTile& t) {  // :658-744
  if (!m.flowers) return;
  const bool lily = m.purpose != UPF_PURPOSE_LAND;
  if ((t.biome == kBiomeWater) != lily) return;
  if (m_range_mask(fabsf(t.temperature - 18.0f), 12.0f, 7.0f) < 1e-5f) return;
  if (!lily && m_range_mask(fabsf(t.precipitation - 200.0f), 100.0f, 70.0f) < 1e-5f) return;
  float c[3], k[3];
  m_voronoi_plane(ms(m, 30), 1, 35.0f, t.px, t.py, t.pz, t.nx, t.ny, t.nz, c);
  const float ex = c[0] - t.px, ey = c[1] - t.py, ez = c[2] - t.pz;
  if (ex * ex + ey * ey + ez * ez > 16.0f) return;
  m_voronoi_plane(ms(m, 29), 1, 200.0f, t.px, t.py, t.pz, t.nx, t.ny, t.nz, k);
  const float fx = c[0] - k[0], fy = c[1] - k[1], fz = c[2] - k[2];
  if (fx * fx + fy * fy + fz * fz > 400.0f) return;
  const float dist = m_dist(c[0], c[1], c[2], t.px, t.py, t.pz);
  const float size = m_smootherstep(m_smootherstep(
                         m_sat(mnoise(kMNoPerlin, ms(m, 31), 0.35f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.5f + 0.5f))) *
                         2.0f +
                     1.5f;
  const float bf = m_range_mask(size - dist, 0.0f, 0.1f);
  if (bf < 1e-7f) return;
  float r, g, b;
  if (lily) { r = 0.2f; g = 0.14f; b = 0.0f; }
  else {
    const float sel = m_sharp_edge(mnoise(kMNoPerlin, ms(m, 32), 0.02f, 0, 1, 2.0f, 0.5f, 0, 0, k[0], k[1], k[2]) + 0.4f, 0.2f);
    m_mix_color(0.5f, 0.0f, 0.35f, 0.5f, 0.4f, 0.0f, sel, r, g, b);
  }
  m_hue_shift(r, g, b,
              mnoise(kMNoPerlin, ms(m, 33), 0.005f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.1f +
                  mnoise(kMNoPerlin, ms(m, 34), 0.4f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.1f);
  const float q = dist / size;
  m_blend(t, r, g, b, 0.8f, lily ? 0.85f : 0.0f, 0.7f + q * q * 0.2f, bf);
  t.opacity = m_lerp(t.opacity, 1.0f, bf);
}

__device__ __forceinline__ void m_ice(const DMaterial& m, MTile& t) {  // :746-817
  const float off = mnoise(kMNoValue, ms(m, 35), 0.035f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 1.5f;
  const float bf = m_sharp_edge(m_range_mask(t.temperature + off, 0.0f, -3.0f)) * (1.0f - m_beach(m, t));
  if (bf < 1e-7f) return;
  if (bf > 0.1f && t.type != kTypeCliffs) t.type = kTypeRough;
  const float freq = mnoise(kMNoValue, ms(m, 36), 0.01f, 1, 3, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.08f + 1.0f;
  const float fx = t.px * freq, fy = t.py * freq, fz = t.pz * freq;
  const float thick =
      m_sharp_edge(m_sat(mnoise(kMNoCellular, ms(m, 37), 0.05f, 1, 2, 2.0f, 0.5f, kDistHybrid, kRetCell, fx, fy, fz) * 0.5f + 0.8f),
                   0.1f) * 0.6f + 0.4f;
  const float crack =
      1.0f - upf_math::powf_(mnoise(kMNoCellular, ms(m, 37), 0.05f, 1, 2, 2.0f, 0.5f, kDistHybrid, kRetSub, fx, fy, fz) * 0.5f + 0.5f, 0.3f);
  float r = fmaxf(122.0f / 255.0f - crack * 0.3f, 0.0f);
  float g = fmaxf(162.0f / 255.0f - crack * 0.3f, 0.0f);
  float b = fmaxf(164.0f / 255.0f - crack * 0.3f, 0.0f);
  m_mix_color(t.r, t.g, t.b, r, g, b, thick, r, g, b);
  float rough = 0.15f + crack * 0.6f;
  rough = m_lerp(t.roughness, rough, thick);
  const float height = thick * 0.5f * (1.0f - crack * 0.8f);
  const float op = t.opacity;
  m_blend(t, r, g, b, rough, 0.0f, height, bf);
  t.opacity = m_lerp(op, op * 0.9f + thick * 0.5f, bf);
}

__device__ __forceinline__ void m_snow(const DMaterial& m, MTile& t) {  // :819-855
  float bf = m_sharp_edge(
      m_range_mask(t.elevation + mnoise(kMNoValue, ms(m, 38), 0.055f, 0, 1, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 20.0f,
                   220.0f, 240.0f) *
      m_steepness(t.slope, 25.0f, 5.0f));
  const float thick = mnoise(kMNoCubic, ms(m, 39), 0.05f, 1, 3, 2.0f, 0.5f, 0, 0, t.px, t.py, t.pz) * 0.5f + 0.5f;
  bf = bf * m_sat(thick * 0.5f + 0.7f);
  if (bf < 1e-7f) return;
  const float c = 230.0f / 255.0f;
  m_blend(t, c, c, c, thick * 0.3f + 0.2f, 0.0f, 0.6f + t.height * 0.1f + thick * 0.3f, bf);
}

__device__ __forceinline__ void m_visualization(MTile& t) {  // :857-893
  float r, g, b;
  switch (t.biome) {
    case kBiomeBare: r = g = b = 187.0f; break;
    case kBiomeTundra: r = 175.0f; g = 226.0f; b = 255.0f; break;
    case kBiomeTaiga: r = 193.0f; g = 225.0f; b = 221.0f; break;
    case kBiomeShrubland: r = 200.0f; g = 110.0f; b = 62.0f; break;
    case kBiomeGrassland: r = 239.0f; g = 252.0f; b = 143.0f; break;
    case kBiomeTemperateSeasonal: r = 151.0f; g = 183.0f; b = 102.0f; break;
    case kBiomeTemperateRain: r = 118.0f; g = 169.0f; b = 92.0f; break;
    case kBiomeDesert: r = 221.0f; g = 187.0f; b = 76.0f; break;
    case kBiomeSavanna: r = 253.0f; g = 213.0f; b = 120.0f; break;
    case kBiomeTropicalSeasonal: r = 161.0f; g = 151.0f; b = 0.0f; break;
    case kBiomeTropicalRain: r = 50.0f; g = 122.0f; b = 30.0f; break;
    default: r = 0.0f; g = 0.0f; b = 255.0f; break;  // water
  }
  t.r = r / 255.0f; t.g = g / 255.0f; t.b = b / 255.0f;
  t.roughness = 0.7f;
}

__device__ __forceinline__ void m_finalize(MTile& t) {  // :895-902
  t.r = m_sat(t.r); t.g = m_sat(t.g); t.b = m_sat(t.b);
  t.roughness = m_sat(t.roughness);
  t.metallic = m_sat(t.metallic);
  t.height = m_sat(t.height);
  t.opacity = m_sat(t.opacity);
}

// terrainTile with the configured coloring mode (sources/coloringModes.cpp:905-966).
// The tile arrives with position and unit normal; albedo, roughness, metallic and height
// start at 0, opacity at 1.
__device__ __forceinline__ void material_tile(const DMaterial& m, MTile& t) {
  t.r = t.g = t.b = 0.0f;
  t.roughness = t.metallic = t.height = 0.0f;
  t.opacity = 1.0f;
  t.slope = t.temperature = t.precipitation = 0.0f;
  t.biome = kBiomeWater + 1;
  t.type = kTypeWater + 1;
  // generateElevation (:51-57)
  t.elevation = m.purpose == UPF_PURPOSE_LAND ? m_elev_sdf(m, t.px, t.py, t.pz)
                                               : elevation_eval<true, true>(m.field, t.px, t.py, t.pz);
  const bool water = m.purpose == UPF_PURPOSE_WATER;
  if (m.coloring == UPF_COLORING_BARREN) {  // :932-952
    m_slope(m, t);
    if (water) {
      t.biome = kBiomeWater;
      t.type = kTypeWater;
      m_water(m, t);
    } else {
      t.biome = kBiomeBare;
      t.type = kTypeFlat;
      m_bedrock(m, t);
      m_cliffs(m, t);
      m_mica(m, t);
      m_boulders(m, t);
    }
    m_finalize(t);
    return;
  }
  m_precipitation(m, t);
  m_temperature(m, t);
  m_slope(m, t);
  m_biome(t);
  m_type(t);
  if (m.coloring == UPF_COLORING_DEBUG) {  // :954-966
    m_visualization(t);
    if (water) t.opacity = 0.5f;
    m_finalize(t);
    return;
  }
  if (water) m_water(m, t);  // :905-930
  else {
    m_bedrock(m, t);
    m_cliffs(m, t);
    m_mica(m, t);
    m_dirt(m, t);
    m_sand(m, t);
    m_grass(m, t);
    m_boulders(m, t);
  }
  m_flowers(m, t);
  m_ice(m, t);
  if (!water) m_snow(m, t);
  m_finalize(t);
}

}  // namespace upf
