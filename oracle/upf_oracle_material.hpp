// upf_oracle_material.hpp -- CPU restatement of the reference's material stack
// (sources/coloringModes.cpp: coloringDefault :905-930, coloringBarren :932-952,
// coloringDebug :954-966 and the generate* functions they call), in FP32.
//
// TEST INFRASTRUCTURE ONLY (see upf_oracle.hpp).  cage's noise kinds, colour conversions
// and configuration defaults are not in the reference's checkout; the definitions
// restated here are the ones documented at upf_material_spec in include/upf.h.  pow and
// atan are the project's (upf_math.hpp), as on the device.  Noise
// object k of the stack is seeded seed + 16 k (the reference's noiseSeed() draws), with
// the index table below.
#pragma once

#include "upf_oracle.hpp"

namespace upfo {
namespace mat {

enum Kind { PERLIN, VALUE, CUBIC, SIMPLEX, CELLULAR };
enum Dist { EUCLID, MANHATTAN, HYBRID };
enum Ret { CELL, SUB, MUL, DIV };
enum Biome { BARE, TUNDRA, TAIGA, SHRUBLAND, GRASSLAND, TEMPERATE_SEASONAL, TEMPERATE_RAIN, DESERT, SAVANNA,
             TROPICAL_SEASONAL, TROPICAL_RAIN, WATER_BIOME };
enum Type { ROAD, FLAT, ROUGH, CLIFFS, WATER_TYPE };

// noise object indices: the seed of object k is seed + 16 k
enum Obj {
  O_BEACH = 0, O_PRECIP, O_TEMP, O_POLAR, O_BED_CELLS, O_BED_DEPTH, O_BED_FREQ, O_BED_SAT, O_CLIFF_CRACKS,
  O_MICA_MASK, O_MICA_CRACKS, O_DIRT_HEIGHT, O_DIRT_CRACKS, O_DIRT_MASK, O_SAND_SCALE, O_SAND_HEIGHT,
  O_SAND_COLOR, O_SAND_HUE, O_GRASS_CELLS, O_GRASS_HUE, O_GRASS_SCRATCH, O_BOULDER_THRESH, O_BOULDER_VORONOI,
  O_BOULDER_SIZE, O_BOULDER_HUE, O_BOULDER_VALUE, O_BOULDER_SCRATCH, O_WATER_SWITCH, O_WATER_HUE,
  O_FLOWER_CLUSTER, O_FLOWER_CENTER, O_FLOWER_SIZE, O_FLOWER_COLOR, O_FLOWER_HUE1, O_FLOWER_HUE2, O_ICE_OFFSET,
  O_ICE_FREQ, O_ICE_CELLS, O_SNOW_OFFSET, O_SNOW_THICK, O_UNUSED40, O_UNUSED41, O_DEVIATION
};

struct Cfg {
  Kind kind;
  float freq;
  int fractal = 0;  // 0 none, 1 fBm, 2 ridged
  int octaves = 3;
  float lac = 2.0f, gain = 0.5f;
  Dist dist = EUCLID;
  Ret ret = CELL;
};

struct Spec {
  SpecT<float> field;
  int coloring = 0, purpose = 1, poles = 0, flowers = 0;
  uint32_t seed = 0;
  uint32_t obj(int k) const { return seed + 16u * (uint32_t)k; }
};

struct Tile {
  float pos[3], nrm[3];
  float albedo[3] = {0, 0, 0};
  float roughness = 0, metallic = 0, height = 0, elevation = 0, slope = 0, temperature = 0, precipitation = 0;
  float opacity = 1;
  int biome = 12, type = 5;
};

inline float lerp(float a, float b, float t) { return a + t * (b - a); }
inline float sat(float x) { return std::min(std::max(x, 0.0f), 1.0f); }
inline float smoothstep(float x) { return x * x * (3.0f - 2.0f * x); }
inline float smootherstep(float x) { return x * x * x * (x * (x * 6.0f - 15.0f) + 10.0f); }

// ---- noise kinds
// Value noise interpolated per axis by the cubic through b (t=0) and c (t=1) with the
// outer neighbours a and d, over the 4^3 lattice values around q, scaled by 1/1.5^3.
inline float cubic(float a, float b, float c, float d, float t) {
  const float p = (d - c) - (a - b);
  return t * t * t * p + t * t * ((a - b) - p) + t * (c - a) + b;
}
inline float noise_value_cubic(uint32_t seed, const float q[3]) {
  const float f[3] = {std::floor(q[0]), std::floor(q[1]), std::floor(q[2])};
  const int32_t i0 = (int32_t)f[0], j0 = (int32_t)f[1], k0 = (int32_t)f[2];
  const float t[3] = {q[0] - f[0], q[1] - f[1], q[2] - f[2]};
  float zs[4];
  for (int k = 0; k < 4; k++) {
    float ys[4];
    for (int j = 0; j < 4; j++) {
      float xs[4];
      for (int i = 0; i < 4; i++) xs[i] = lattice_value(hash3(seed, i0 + i - 1, j0 + j - 1, k0 + k - 1));
      ys[j] = cubic(xs[0], xs[1], xs[2], xs[3], t[0]);
    }
    zs[k] = cubic(ys[0], ys[1], ys[2], ys[3], t[1]);
  }
  return cubic(zs[0], zs[1], zs[2], zs[3], t[2]) * (1.0f / 3.375f);
}

// F1 / F2 of the 27 cells around rint(q); one point per cell at k - 0.5 + r.
inline float noise_cellular2(uint32_t seed, const float q[3], Dist dist, Ret ret) {
  const int32_t c[3] = {(int32_t)std::nearbyint(q[0]), (int32_t)std::nearbyint(q[1]), (int32_t)std::nearbyint(q[2])};
  float d0 = INFINITY, d1 = INFINITY;
  uint32_t closest = 0;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const int32_t k[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
        const uint32_t h = hash3(seed, k[0], k[1], k[2]);
        const uint32_t m = h ^ (h >> 15);
        const uint32_t bits[3] = {m & 1023u, (m >> 10) & 1023u, (m >> 20) & 1023u};
        float e[3];
        for (int a = 0; a < 3; a++) e[a] = (((float)k[a] - 0.5f) + (float)bits[a] * (1.0f / 1023.0f)) - q[a];
        const float man = std::fabs(e[0]) + std::fabs(e[1]) + std::fabs(e[2]);
        const float euc = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        const float d = dist == MANHATTAN ? man : (dist == HYBRID ? man + euc : euc);
        if (d < d0) { d1 = d0; d0 = d; closest = h; }
        else if (d < d1) d1 = d;
      }
  if (ret == CELL) return lattice_value(closest);
  if (dist == EUCLID) { d0 = std::sqrt(d0); d1 = std::sqrt(d1); }
  switch (ret) {
    case SUB: return d1 - d0 - 1.0f;
    case MUL: return d1 * d0 * 0.5f - 1.0f;
    default: return d0 / d1 - 1.0f;
  }
}

inline float noise(const Cfg& c, uint32_t seed, float x, float y, float z) {
  const int oct = c.fractal == 0 ? 1 : c.octaves;
  float amp = 1.0f;
  if (c.fractal != 0) {  // fractal bounding 1 / sum |gain|^k
    const float g = std::fabs(c.gain);
    float sum = 1.0f, a = g;
    for (int i = 1; i < oct; i++) { sum = sum + a; a = a * g; }
    amp = 1.0f / sum;
  }
  float acc = 0.0f, f = c.freq;
  for (int k = 0; k < oct; k++) {
    const float q[3] = {x * f, y * f, z * f};
    const uint32_t s = seed + (uint32_t)k;
    float n;
    switch (c.kind) {
      case PERLIN: n = noise_gradient<float>(s, q); break;
      case VALUE: n = noise_value<float>(s, q); break;
      case CUBIC: n = noise_value_cubic(s, q); break;
      case SIMPLEX: n = noise_simplex<float>(s, q); break;
      default: n = noise_cellular2(s, q, c.dist, c.ret); break;
    }
    if (c.fractal == 0) return n;
    acc = acc + (c.fractal == 2 ? (1.0f - 2.0f * std::fabs(n)) : n) * amp;
    amp = amp * c.gain;
    f = f * c.lac;
  }
  return acc;
}
inline float noise(const Cfg& c, uint32_t seed, const float p[3]) { return noise(c, seed, p[0], p[1], p[2]); }

// ---- colour (hexcone HSV, h in [0, 1))
inline void rgb_to_hsv(const float c[3], float o[3]) {
  const float mx = std::max(c[0], std::max(c[1], c[2])), mn = std::min(c[0], std::min(c[1], c[2]));
  const float d = mx - mn;
  o[2] = mx;
  o[1] = mx > 0.0f ? d / mx : 0.0f;
  if (d <= 0.0f) { o[0] = 0.0f; return; }
  float h;
  if (mx == c[0]) h = (c[1] - c[2]) / d + (c[1] < c[2] ? 6.0f : 0.0f);
  else if (mx == c[1]) h = (c[2] - c[0]) / d + 2.0f;
  else h = (c[0] - c[1]) / d + 4.0f;
  o[0] = h / 6.0f;
}
inline void hsv_to_rgb(const float c[3], float o[3]) {
  const float h6 = (c[0] - std::floor(c[0])) * 6.0f;
  const int i = std::min((int)h6, 5);
  const float f = h6 - (float)i, s = c[1], v = c[2];
  const float p = v * (1.0f - s), q = v * (1.0f - s * f), t = v * (1.0f - s * (1.0f - f));
  const float tab[6][3] = {{v, t, p}, {q, v, p}, {p, v, t}, {p, q, v}, {t, p, v}, {v, p, q}};
  o[0] = tab[i][0]; o[1] = tab[i][1]; o[2] = tab[i][2];
}
inline void hue_shift(float c[3], float shift) {  // sources/math.cpp:47-52
  float h[3];
  rgb_to_hsv(c, h);
  h[0] = std::fmod(h[0] + shift + 1.0f, 1.0f);
  hsv_to_rgb(h, c);
}
inline void mix_color(const float a[3], const float b[3], float t, float o[3]) {  // HSV, shorter hue arc
  float ha[3], hb[3];
  rgb_to_hsv(a, ha);
  rgb_to_hsv(b, hb);
  float dh = hb[0] - ha[0];
  if (dh > 0.5f) dh = dh - 1.0f;
  else if (dh < -0.5f) dh = dh + 1.0f;
  const float h[3] = {std::fmod(ha[0] + t * dh + 1.0f, 1.0f), lerp(ha[1], hb[1], t), lerp(ha[2], hb[2], t)};
  hsv_to_rgb(h, o);
}
inline void deviation(float c[3], float dev, uint32_t seed, const float p[3]) {  // sources/math.cpp:40-45
  uint32_t b[3];
  std::memcpy(b, p, 12);
  const uint32_t h = hash1(hash3(seed, (int32_t)b[0], (int32_t)b[1], (int32_t)b[2]));
  const float u[3] = {(float)(h & 1023u) * (1.0f / 1023.0f) - 0.5f, (float)((h >> 10) & 1023u) * (1.0f / 1023.0f) - 0.5f,
                      (float)((h >> 20) & 1023u) * (1.0f / 1023.0f) - 0.5f};
  float hsv[3];
  rgb_to_hsv(c, hsv);
  hsv[0] = std::fmod(hsv[0] + u[0] * dev + 1.0f, 1.0f);
  hsv[1] = sat(hsv[1] + u[1] * dev);
  hsv[2] = sat(hsv[2] + u[2] * dev);
  hsv_to_rgb(hsv, c);
}

// ---- helpers of sources/coloringModes.cpp:22-49
inline float sharp_edge_f(float v, float p = 0.05f) {
  const float lo = 0.5f - p, hi = 0.5f + p;
  return (std::min(std::max(v, lo), hi) - lo) / (hi - lo) * (1.0f - 0.0f) + 0.0f;
}
inline float range_mask(float v, float zero_at, float one_at) { return sat((v - zero_at) / (one_at - zero_at)); }
constexpr float kDeg = 57.29577951308232f;
inline float steepness(float slope_rad, float threshold, float smoothing) {
  return sat((threshold + smoothing - slope_rad * kDeg) / (2.0f * smoothing));
}

// Voronoi evaluate(position, normal).points[0]: the reference's points, projected into
// the plane through the position (sources/voronoi.cpp:58-66), then the nearest.
inline void voronoi_plane(uint32_t seed, int ppc, float cell, const float pos[3], const float n[3], float out[3]) {
  const float frequency = 1.0f / cell;
  const int32_t cc[3] = {cell_index(pos[0] * frequency), cell_index(pos[1] * frequency), cell_index(pos[2] * frequency)};
  std::vector<float> pts(3 * (size_t)ppc);
  float best = INFINITY;
  out[0] = out[1] = out[2] = 0.0f;
  for (int z = -1; z < 2; z++)
    for (int y = -1; y < 2; y++)
      for (int x = -1; x < 2; x++) {
        const int32_t k[3] = {cc[0] + x, cc[1] + y, cc[2] + z};
        voronoi_cell_points(seed, ppc, cell, k, pts.data());
        for (int i = 0; i < ppc; i++) {
          float q[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
          const float dd = (q[0] - pos[0]) * n[0] + (q[1] - pos[1]) * n[1] + (q[2] - pos[2]) * n[2];
          for (int a = 0; a < 3; a++) q[a] = q[a] - dd * n[a];
          const float e[3] = {q[0] - pos[0], q[1] - pos[1], q[2] - pos[2]};
          const float d = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
          if (d < best) { best = d; out[0] = q[0]; out[1] = q[1]; out[2] = q[2]; }
        }
      }
}
inline float dist3(const float a[3], const float b[3]) {
  const float d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}
inline void apply(Tile& t, const float c[3], float rough, float metal, float height, float bf) {
  for (int a = 0; a < 3; a++) t.albedo[a] = lerp(t.albedo[a], c[a], bf);
  t.roughness = lerp(t.roughness, rough, bf);
  t.metallic = lerp(t.metallic, metal, bf);
  t.height = lerp(t.height, height, bf);
}

// terrainSdfElevation (sources/modesConfiguration.cpp:230-237): shape * ratio
inline float elev_sdf(const Spec& s, const float p[3]) { return shape_eval(s.field, p) / s.field.elevation_scale; }

inline float beach(const Spec& s, const Tile& t) {  // :34-49
  if (t.biome == WATER_BIOME) return 0.0f;
  return range_mask(t.elevation + noise({VALUE, 0.035f}, s.obj(O_BEACH), t.elevation, 0.0f, 0.0f) * 20.0f, 10.0f, 15.0f);
}

inline void precipitation(const Spec& s, Tile& t) {  // :59-80
  Cfg c{CUBIC, 0.001f, 1, 2};
  float p = noise(c, s.obj(O_PRECIP), t.pos) * 0.5f + 0.5f;
  p = sat(p);
  p = smootherstep(p);
  p = smootherstep(p);
  p = upf_math::powf_(std::max(p + 0.2f, 0.0f), 3.0f) * 0.6f;
  p *= 400.0f;
  p += std::max(120.0f - std::fabs(t.elevation), 0.0f) * 0.2f;
  t.precipitation = std::max(p, 0.0f);
}

inline void temperature(const Spec& s, Tile& t) {  // :82-120
  Cfg c{SIMPLEX, 0.0003f, 1, 2, 2.0f, 0.4f};
  float v = noise(c, s.obj(O_TEMP), t.pos) * 0.5f + 0.5f;
  v = sat(v);
  v = smoothstep(v);
  v = v * 2.0f - 1.0f;
  if (s.poles) {
    const float len = std::sqrt(t.pos[0] * t.pos[0] + t.pos[2] * t.pos[2]);
    float polar = std::fabs(upf_math::atanf_(t.pos[1] / len)) / 3.14159265358979f * 2.0f;
    polar = upf_math::powf_(polar, 1.7f);
    polar += noise({VALUE, 0.007f, 1, 3}, s.obj(O_POLAR), t.pos) * 0.1f;
    v += 0.6f - polar * 3.2f;
  }
  t.temperature = 15.0f + v * 30.0f - std::max(t.elevation, 0.0f) * 0.02f;
}

inline void slope(const Spec& s, Tile& t) {  // :122-151
  const float* n = t.nrm;
  float b0[3] = {1.0f, 0.0f, 0.0f};  // anyPerpendicular, sources/math.cpp:65-72
  if (std::fabs(n[0] * b0[0] + n[1] * b0[1] + n[2] * b0[2]) > 0.9f) { b0[0] = 0.0f; b0[1] = 1.0f; }
  float a[3] = {n[1] * b0[2] - n[2] * b0[1], n[2] * b0[0] - n[0] * b0[2], n[0] * b0[1] - n[1] * b0[0]};
  const float al = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  for (int i = 0; i < 3; i++) a[i] = a[i] / al;
  float b[3] = {n[1] * a[2] - n[2] * a[1], n[2] * a[0] - n[0] * a[2], n[0] * a[1] - n[1] * a[0]};
  const float radius = 0.5f;
  for (int i = 0; i < 3; i++) { a[i] *= radius; b[i] *= radius; }
  const float div = 1.0f / std::sqrt(2.0f);
  float c[3], d[3];
  for (int i = 0; i < 3; i++) { c[i] = (a[i] + b[i]) * div; d[i] = (a[i] - b[i]) * div; }
  const float* dirs[4] = {a, b, c, d};
  float e[8];
  for (int k = 0; k < 8; k++) {
    const float* v = dirs[k < 4 ? (k % 2) : 2 + (k % 2)];
    const float sg = (k % 4) < 2 ? 1.0f : -1.0f;  // +a +b -a -b +c +d -c -d
    const float p[3] = {sg > 0 ? t.pos[0] + v[0] : t.pos[0] - v[0], sg > 0 ? t.pos[1] + v[1] : t.pos[1] - v[1],
                        sg > 0 ? t.pos[2] + v[2] : t.pos[2] - v[2]};
    e[k] = elev_sdf(s, p);
  }
  float e1 = e[0], e2 = e[0];
  for (float v : e) { e1 = std::min(e1, v); e2 = std::max(e2, v); }
  const float md = (e2 - e1) * 0.1f;
  t.slope = upf_math::atanf_(md / radius);
}

inline void biome(Tile& t) {  // :153-190
  const float T = t.temperature, P = t.precipitation;
  if (t.elevation < 0) t.biome = WATER_BIOME;
  else if (T > 20) t.biome = P < 40 ? DESERT : P < 130 ? SAVANNA : P < 230 ? TROPICAL_SEASONAL : TROPICAL_RAIN;
  else if (T > 5) t.biome = P < 30 ? SHRUBLAND : P < 100 ? GRASSLAND : P < 200 ? TEMPERATE_SEASONAL : TEMPERATE_RAIN;
  else if (T > -5) t.biome = P < 30 ? BARE : TAIGA;
  else t.biome = P < 10 ? BARE : TUNDRA;
}

inline void type(Tile& t) {  // :192-200
  if (t.elevation < 0) t.type = WATER_TYPE;
  else if (t.slope > 20.0f / kDeg) t.type = CLIFFS;
  else t.type = FLAT;
}

inline void bedrock(const Spec& s, Tile& t) {  // :202-261
  const float dn = noise({CUBIC, 0.0031f}, s.obj(O_BED_DEPTH), t.pos) * 0.5f + 0.501f;
  const float depth = dn * dn * 2.0f;
  const float freq = noise({CUBIC, 0.013f}, s.obj(O_BED_FREQ), t.pos) * 0.15f + 1.0f;
  const float pf[3] = {t.pos[0] * freq, t.pos[1] * freq, t.pos[2] * freq};
  Cfg cr{CELLULAR, 0.017f};
  cr.dist = HYBRID;
  cr.ret = SUB;
  const float cracks = sat(upf_math::powf_(noise(cr, s.obj(O_BED_CELLS), pf) * 0.5f + 0.5f, 0.7f));
  cr.ret = CELL;
  const float value = noise(cr, s.obj(O_BED_CELLS), pf);
  const float satn = sat(noise({CUBIC, 0.0022f, 1, 2}, s.obj(O_BED_SAT), t.pos) * 0.35f + 0.3f);
  const float hsv[3] = {0.07f, satn, sat(value * 0.1f + 0.7f) * (cracks * 0.4f + 0.6f)};
  hsv_to_rgb(hsv, t.albedo);
  t.roughness = lerp(0.9f, value * -0.15f + 0.65f, cracks);
  t.height = cracks * depth;
}

inline void cliffs(const Spec& s, Tile& t) {  // :263-293
  const float bf = steepness(t.slope, 19.0f, 4.0f);
  if (bf > 0.99999f) return;
  float hsv[3];
  rgb_to_hsv(t.albedo, hsv);
  hsv[0] = lerp(155.0f / 255.0f, hsv[0], sharp_edge_f(bf, 0.005f));
  hsv[1] *= bf * 0.4f + 0.6f;
  hsv[2] = lerp(0.7f, hsv[1], bf * 0.8f + 0.2f);
  hsv_to_rgb(hsv, t.albedo);
  t.roughness *= bf * 0.2f + 0.8f;
  const float cracks = sharp_edge_f(sat(noise({SIMPLEX, 0.047f, 2, 2}, s.obj(O_CLIFF_CRACKS), t.pos) * 2.0f - 1.1f), 0.2f);
  t.height *= lerp(1.0f, 0.8f, cracks);
  const float k = lerp(1.0f, 0.9f, cracks);
  for (float& c : t.albedo) c *= k;
  t.roughness = lerp(t.roughness, 0.9f, cracks);
}

inline void mica(const Spec& s, Tile& t) {  // :295-336
  if (!s.flowers) return;
  Cfg m{CELLULAR, 0.02f};
  m.dist = HYBRID;
  m.ret = MUL;
  const float bf = sat((noise(m, s.obj(O_MICA_MASK), t.pos) - 0.98f) * 10.0f);
  if (bf < 1e-7f) return;
  Cfg c{CELLULAR, 0.05f};
  c.dist = MANHATTAN;
  c.ret = DIV;
  const float cracks = sharp_edge_f(sat(noise(c, s.obj(O_MICA_CRACKS), t.pos) + 0.6f));
  const float col[3] = {lerp(122.0f / 255, 184.0f / 255, cracks), lerp(90.0f / 255, 209.0f / 255, cracks),
                        lerp(88.0f / 255, 187.0f / 255, cracks)};
  apply(t, col, cracks * 0.3f + 0.4f, 1.0f, t.height * 0.5f, bf);
}

inline void dirt(const Spec& s, Tile& t) {  // :338-402
  float height = noise({PERLIN, 0.05f, 1, 3, 2.3f, 0.4f}, s.obj(O_DIRT_HEIGHT), t.pos) * 0.2f + 0.5f;
  const float bf = sharp_edge_f(sat(height - t.height + 0.4f)) * steepness(t.slope, 20.0f, 5.0f);
  if (bf < 1e-7f) return;
  float col[3] = {168.0f / 255, 94.0f / 255, 28.0f / 255};
  {
    const float saturation = range_mask(t.precipitation, 0, 50);
    float hsv[3];
    rgb_to_hsv(col, hsv);
    hsv[1] *= saturation * 0.7f;
    hsv[2] *= 0.8f;
    hsv_to_rgb(hsv, col);
  }
  float rough = 0.8f;
  float cracks = sharp_edge_f(sat(noise({SIMPLEX, 0.07f, 2, 2}, s.obj(O_DIRT_CRACKS), t.pos) * 2.0f - 1.2f), 0.15f);
  const float cm = smoothstep(sat(noise({CUBIC, 0.02f, 1, 4}, s.obj(O_DIRT_MASK), t.pos) * 0.5f + 0.5f));
  cracks *= cm * cm * 0.9f + 0.1f;
  cracks *= range_mask(t.precipitation, 70, 20) * 0.9f + 0.1f;
  height = lerp(height, height * 0.5f, cracks);
  for (float& c : col) c = lerp(c, 0.1f, cracks);
  rough = lerp(rough, 0.9f, cracks);
  apply(t, col, rough, 0.0f, height, bf);
}

inline void sand(const Spec& s, Tile& t) {  // :404-465
  const float bf = range_mask(t.temperature, 24, 28) * steepness(t.slope, 19.0f, 10.0f);
  if (bf < 1e-7f) return;
  const float hs = noise({SIMPLEX, 0.001f, 1, 3}, s.obj(O_SAND_SCALE), t.pos) * 0.3f + 1.0f;
  const float ph[3] = {t.pos[0] * hs, t.pos[1] * hs, t.pos[2] * hs};
  const float height = (noise({SIMPLEX, 0.01f, 2, 3, 2.0f, 0.7f}, s.obj(O_SAND_HEIGHT), ph) * 0.2f) *
                           (range_mask(t.precipitation, 100, 50) * 0.4f + 0.6f) +
                       0.5f;
  const float cs = smootherstep(smootherstep(noise({PERLIN, 0.0008f}, s.obj(O_SAND_COLOR), t.pos) * 0.5f + 0.5f));
  const float hue = noise({CUBIC, 0.02f, 1, 2}, s.obj(O_SAND_HUE), t.pos) * 0.1f;
  float col[3] = {189.0f / 255, 174.0f / 255, 152.0f / 255};
  hue_shift(col, hue);
  const float other[3] = {187.0f / 255, 104.0f / 255, 50.0f / 255};
  mix_color(col, other, cs, col);
  deviation(col, 0.08f, s.obj(O_DEVIATION), t.pos);
  float hsv[3];
  rgb_to_hsv(col, hsv);
  hsv[0] *= 1.0f;
  hsv[1] *= 0.7f;
  hsv[2] *= 0.8f;
  hsv_to_rgb(hsv, col);
  apply(t, col, (hs - 1.0f) * 0.7f + 0.55f + hue, 0.0f, height, bf);
}

inline void grass(const Spec& s, Tile& t) {  // :467-533
  const float bf = range_mask(t.temperature, 35, 25) * range_mask(t.precipitation, 15, 35) *
                   steepness(t.slope, 22.0f, 3.0f) * beach(s, t);
  if (bf < 1e-7f) return;
  const float dry = std::min(std::max(t.temperature - (t.precipitation + 100.0f) * 30.0f / 400.0f, 0.0f), 5.0f) / 5.0f;
  const float hsb = noise({PERLIN, 0.01f, 1, 2}, s.obj(O_GRASS_HUE), t.pos);
  const float hue = hsb * 0.09f - dry * 0.1f;
  float col[3] = {114.0f / 255, 187.0f / 255, 77.0f / 255};
  hue_shift(col, hue);
  Cfg cells{CELLULAR, 0.1f};
  cells.dist = HYBRID;
  cells.ret = CELL;
  const float pm = sharp_edge_f(noise(cells, s.obj(O_GRASS_CELLS), t.pos) * 0.5f + 0.2f);
  cells.ret = SUB;
  const float patches = upf_math::powf_(noise(cells, s.obj(O_GRASS_CELLS), t.pos) * 0.5f + 0.5f, 0.5f) * pm * (1.0f - dry * 0.8f);
  for (float& c : col) c *= 1.0f - patches * 0.09f;
  Cfg scr{CELLULAR, 0.3f};
  scr.dist = HYBRID;
  scr.ret = MUL;
  const float scratches = sat(noise(scr, s.obj(O_GRASS_SCRATCH), t.pos) - 0.2f) * (1.0f - dry * 0.5f);
  for (float& c : col) c *= 1.0f - scratches * 0.1f;
  const float rough = 0.6f + dry * 0.1f + patches * 0.1f + scratches * 0.1f - hsb * 0.1f;
  const float height =
      lerp(t.height, 0.5f, 0.7f) + (0.5f - dry) * 0.2f + patches * 0.05f - scratches * 0.02f + hsb * 0.1f + 0.1f;
  apply(t, col, rough, 0.0f, height, bf);
}

inline void boulders(const Spec& s, Tile& t) {  // :535-615
  if (!s.flowers) return;
  if (noise({PERLIN, 0.01f}, s.obj(O_BOULDER_THRESH), t.pos) < 0.15f) return;
  float center[3];
  voronoi_plane(s.obj(O_BOULDER_VORONOI), 2, 200.0f, t.pos, t.nrm, center);
  const float dist = dist3(center, t.pos);
  const float size =
      smootherstep(smootherstep(sat(noise({PERLIN, 0.09f}, s.obj(O_BOULDER_SIZE), t.pos) * 0.5f + 0.5f))) * 3.0f + 3.0f;
  const float bf = range_mask(size - dist, 0, 0.5f);
  if (bf < 1e-7f) return;
  const float hue = noise({PERLIN, 0.3f}, s.obj(O_BOULDER_HUE), t.pos) * 0.12f;
  const float val = noise({PERLIN, 0.5f}, s.obj(O_BOULDER_VALUE), t.pos) * 0.18f;
  const float grey[3] = {154.0f / 255, 154.0f / 255, 154.0f / 255};
  float hsv[3];
  rgb_to_hsv(grey, hsv);
  hsv[0] = std::fmod(hsv[0] + hue + 1.0f, 1.0f);
  hsv[2] = sat(hsv[2] + val);
  float col[3];
  hsv_to_rgb(hsv, col);
  const float rough = noise({PERLIN, 0.5f}, s.obj(O_BOULDER_SCRATCH), t.pos) * 0.1f + 0.4f;
  const float q = dist / size;
  apply(t, col, rough, 0.0f, 1.0f - q * q * 0.5f, bf);
}

inline void water(const Spec& s, Tile& t) {  // :617-656
  const float sw = noise({PERLIN, 0.0015f, 1, 2}, s.obj(O_WATER_SWITCH), t.pos) * 0.5f + 0.5f;
  const float hue = noise({CUBIC, 0.004f, 1, 3}, s.obj(O_WATER_HUE), t.pos) * 0.09f;
  const float a[3] = {40.0f / 255, 63.0f / 255, 168.0f / 255}, b[3] = {39.0f / 255, 154.0f / 255, 186.0f / 255};
  mix_color(a, b, sw, t.albedo);
  hue_shift(t.albedo, hue);
  const float r = range_mask(t.elevation, -25, 3);
  const float d1 = 1.0f - r * r;
  const float d2 = rescale(range_mask(t.elevation, 0, -100), 0.0f, 1.0f, 0.7f, 1.0f);
  t.opacity = d1 * d2;
  t.metallic = smoothstep(range_mask(t.elevation, 3, -30));
}

inline void flowers(const Spec& s, Tile& t) {  // :658-744
  if (!s.flowers) return;
  const bool lily = s.purpose != UPF_PURPOSE_LAND;
  if ((t.biome == WATER_BIOME) != lily) return;
  if (range_mask(std::fabs(t.temperature - 18.0f), 12, 7) < 1e-5f) return;
  if (!lily && range_mask(std::fabs(t.precipitation - 200.0f), 100, 70) < 1e-5f) return;
  float center[3], cluster[3];
  voronoi_plane(s.obj(O_FLOWER_CENTER), 1, 35.0f, t.pos, t.nrm, center);
  const float ce[3] = {center[0] - t.pos[0], center[1] - t.pos[1], center[2] - t.pos[2]};
  if (ce[0] * ce[0] + ce[1] * ce[1] + ce[2] * ce[2] > 16.0f) return;
  voronoi_plane(s.obj(O_FLOWER_CLUSTER), 1, 200.0f, t.pos, t.nrm, cluster);
  const float cc[3] = {center[0] - cluster[0], center[1] - cluster[1], center[2] - cluster[2]};
  if (cc[0] * cc[0] + cc[1] * cc[1] + cc[2] * cc[2] > 400.0f) return;
  const float dist = dist3(center, t.pos);
  const float size =
      smootherstep(smootherstep(sat(noise({PERLIN, 0.35f}, s.obj(O_FLOWER_SIZE), t.pos) * 0.5f + 0.5f))) * 2.0f + 1.5f;
  const float bf = range_mask(size - dist, 0, 0.1f);
  if (bf < 1e-7f) return;
  float col[3];
  if (lily) { col[0] = 0.2f; col[1] = 0.14f; col[2] = 0.0f; }
  else {
    const float a[3] = {0.5f, 0.0f, 0.35f}, b[3] = {0.5f, 0.4f, 0.0f};
    mix_color(a, b, sharp_edge_f(noise({PERLIN, 0.02f}, s.obj(O_FLOWER_COLOR), cluster) + 0.4f, 0.2f), col);
  }
  hue_shift(col, noise({PERLIN, 0.005f}, s.obj(O_FLOWER_HUE1), t.pos) * 0.1f +
                     noise({PERLIN, 0.4f}, s.obj(O_FLOWER_HUE2), t.pos) * 0.1f);
  const float q = dist / size;
  apply(t, col, 0.8f, lily ? 0.85f : 0.0f, 0.7f + q * q * 0.2f, bf);
  t.opacity = lerp(t.opacity, 1.0f, bf);
}

inline void ice(const Spec& s, Tile& t) {  // :746-817
  const float off = noise({VALUE, 0.035f}, s.obj(O_ICE_OFFSET), t.pos) * 1.5f;
  const float bf = sharp_edge_f(range_mask(t.temperature + off, 0, -3)) * (1.0f - beach(s, t));
  if (bf < 1e-7f) return;
  if (bf > 0.1f && t.type != CLIFFS) t.type = ROUGH;
  const float freq = noise({VALUE, 0.01f, 1, 3}, s.obj(O_ICE_FREQ), t.pos) * 0.08f + 1.0f;
  const float pf[3] = {t.pos[0] * freq, t.pos[1] * freq, t.pos[2] * freq};
  Cfg cc{CELLULAR, 0.05f, 1, 2};
  cc.dist = HYBRID;
  cc.ret = CELL;
  const float thick = sharp_edge_f(sat(noise(cc, s.obj(O_ICE_CELLS), pf) * 0.5f + 0.8f), 0.1f) * 0.6f + 0.4f;
  cc.ret = SUB;
  const float crack = 1.0f - upf_math::powf_(noise(cc, s.obj(O_ICE_CELLS), pf) * 0.5f + 0.5f, 0.3f);
  float col[3] = {std::max(122.0f / 255 - crack * 0.3f, 0.0f), std::max(162.0f / 255 - crack * 0.3f, 0.0f),
                  std::max(164.0f / 255 - crack * 0.3f, 0.0f)};
  mix_color(t.albedo, col, thick, col);
  float rough = 0.15f + crack * 0.6f;
  rough = lerp(t.roughness, rough, thick);
  const float height = thick * 0.5f * (1.0f - crack * 0.8f);
  const float op = t.opacity;
  apply(t, col, rough, 0.0f, height, bf);
  t.opacity = lerp(op, op * 0.9f + thick * 0.5f, bf);
}

inline void snow(const Spec& s, Tile& t) {  // :819-855
  float bf = sharp_edge_f(
      range_mask(t.elevation + noise({VALUE, 0.055f}, s.obj(O_SNOW_OFFSET), t.pos) * 20.0f, 220, 240) *
      steepness(t.slope, 25.0f, 5.0f));
  const float thick = noise({CUBIC, 0.05f, 1, 3}, s.obj(O_SNOW_THICK), t.pos) * 0.5f + 0.5f;
  bf *= sat(thick * 0.5f + 0.7f);
  if (bf < 1e-7f) return;
  const float col[3] = {230.0f / 255, 230.0f / 255, 230.0f / 255};
  apply(t, col, thick * 0.3f + 0.2f, 0.0f, 0.6f + t.height * 0.1f + thick * 0.3f, bf);
}

inline void visualization(Tile& t) {  // :857-893
  static const float tab[12][3] = {{187, 187, 187}, {175, 226, 255}, {193, 225, 221}, {200, 110, 62},
                                   {239, 252, 143}, {151, 183, 102}, {118, 169, 92},  {221, 187, 76},
                                   {253, 213, 120}, {161, 151, 0},   {50, 122, 30},   {0, 0, 255}};
  const int b = t.biome < 0 || t.biome > 11 ? 11 : t.biome;
  for (int a = 0; a < 3; a++) t.albedo[a] = tab[b][a] / 255.0f;
  t.roughness = 0.7f;
}

inline void finalize(Tile& t) {  // :895-902
  for (float& c : t.albedo) c = sat(c);
  t.roughness = sat(t.roughness);
  t.metallic = sat(t.metallic);
  t.height = sat(t.height);
  t.opacity = sat(t.opacity);
}

// terrainTile (sources/modesConfiguration.cpp:281-286) with the configured coloring.
inline void tile(const Spec& s, Tile& t) {
  t.elevation = s.purpose == UPF_PURPOSE_LAND ? elev_sdf(s, t.pos) : elevation_eval(s.field, t.pos);  // :51-57
  const bool wat = s.purpose == UPF_PURPOSE_WATER;
  if (s.coloring == UPF_COLORING_BARREN) {
    slope(s, t);
    if (wat) { t.biome = WATER_BIOME; t.type = WATER_TYPE; water(s, t); }
    else { t.biome = BARE; t.type = FLAT; bedrock(s, t); cliffs(s, t); mica(s, t); boulders(s, t); }
    finalize(t);
    return;
  }
  precipitation(s, t);
  temperature(s, t);
  slope(s, t);
  biome(t);
  type(t);
  if (s.coloring == UPF_COLORING_DEBUG) {
    visualization(t);
    if (wat) t.opacity = 0.5f;
    finalize(t);
    return;
  }
  if (wat) water(s, t);
  else { bedrock(s, t); cliffs(s, t); mica(s, t); dirt(s, t); sand(s, t); grass(s, t); boulders(s, t); }
  flowers(s, t);
  ice(s, t);
  if (!wat) snow(s, t);
  finalize(t);
}

inline Spec spec_from(const upf_material_spec& m) {
  Spec s;
  s.field = spec_to<float>(m.field);
  s.coloring = m.coloring;
  s.purpose = m.purpose;
  s.poles = m.poles ? 1 : 0;
  s.flowers = m.flowers ? 1 : 0;
  s.seed = m.seed;
  return s;
}

}  // namespace mat
}  // namespace upfo
