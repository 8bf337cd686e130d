// upf_oracle_bake.hpp -- CPU oracle for the texel bake (TEST INFRASTRUCTURE ONLY).
//
// Restates what libupf documents at upf_bake_spec / upf_atlas_per_triangle /
// upf_bake_rasterize: the per-triangle atlas, UV rasterization of texel centres (ties to
// the larger triangle id) and the reduced coloringDefault material
// (reference sources/textures.cpp:24-82, sources/coloringModes.cpp:905-930).  cage's
// meshGenerateTexture rasterizer is not in the reference checkout; its texel-centre and
// coverage rules are unpinned.
#pragma once

#include <cmath>
#include <cstdint>
#include <algorithm>
#include <vector>

#include "upf_oracle.hpp"

namespace upfo {

inline float edge_fn(float ax, float ay, float bx, float by, float px, float py) {
  const float u = (bx - ax) * (py - ay);
  const float v = (by - ay) * (px - ax);
  return u - v;
}

inline int atlas_side(int64_t n, int w, int h) {
  const int64_t cells = (n + 1) / 2;
  int best = 0;
  for (int s = 2; s <= std::min(w, h); s++)
    if ((int64_t)(w / s) * (h / s) >= cells) best = s;
  return best;
}

inline void atlas(int64_t n, int w, int h, float* uv) {
  const int side = atlas_side(n, w, h);
  if (side < 2) return;  // no cell of 2 x 2 texels fits (the library reports UPF_ERR_CAPACITY)
  const int cols = w / side;
  const float iw = 1.0f / (float)w, ih = 1.0f / (float)h;
  for (int64_t t = 0; t < n; t++) {
    const int64_t c = t / 2;
    const float x0 = (float)((int)(c % cols) * side), y0 = (float)((int)(c / cols) * side), s = (float)side;
    const float lo[6] = {x0, y0, x0 + s, y0, x0, y0 + s};
    const float hi[6] = {x0 + s, y0 + s, x0, y0 + s, x0 + s, y0};
    const float* q = (t % 2 == 0) ? lo : hi;
    for (int v = 0; v < 3; v++) {
      uv[6 * t + 2 * v] = q[2 * v] * iw;
      uv[6 * t + 2 * v + 1] = q[2 * v + 1] * ih;
    }
  }
}

// Brute force: every texel centre tested against every triangle whose box covers it.
inline void rasterize(const float* uv, int64_t n, int w, int h, int32_t* tri) {
  for (int64_t i = 0; i < (int64_t)w * h; i++) tri[i] = -1;
  for (int64_t t = 0; t < n; t++) {
    const float ax = uv[6 * t] * (float)w, ay = uv[6 * t + 1] * (float)h;
    const float bx = uv[6 * t + 2] * (float)w, by = uv[6 * t + 3] * (float)h;
    const float cx = uv[6 * t + 4] * (float)w, cy = uv[6 * t + 5] * (float)h;
    const float area = edge_fn(ax, ay, bx, by, cx, cy);
    if (!(area != 0.0f) || !std::isfinite(area)) continue;
    const float sgn = area > 0 ? 1.0f : -1.0f;
    for (int y = 0; y < h; y++) {
      const float py = (float)y + 0.5f;
      if (py < std::min(ay, std::min(by, cy)) || py > std::max(ay, std::max(by, cy))) continue;
      for (int x = 0; x < w; x++) {
        const float px = (float)x + 0.5f;
        if (px < std::min(ax, std::min(bx, cx)) || px > std::max(ax, std::max(bx, cx))) continue;
        if (sgn * edge_fn(bx, by, cx, cy, px, py) >= 0 && sgn * edge_fn(cx, cy, ax, ay, px, py) >= 0 &&
            sgn * edge_fn(ax, ay, bx, by, px, py) >= 0)
          tri[(int64_t)y * w + x] = std::max<int32_t>(tri[(int64_t)y * w + x], (int32_t)t);
      }
    }
  }
}

inline float sat(float x) { return std::min(std::max(x, 0.0f), 1.0f); }
inline float smoothstep(float e0, float e1, float x) {
  const float t = sat((x - e0) / (e1 - e0));
  return t * t * (3.0f - 2.0f * t);
}
inline float mixf(float a, float b, float t) { return a + t * (b - a); }

// One texel of the material, given the interpolated position and normal.
inline void material(const upf_bake_spec& b, const float p[3], const float n[3], float albedo[3], float& bump,
                     float& rough) {
  const SpecT<float> f = spec_to<float>(b.field);
  upf_field_spec tmp = b.field;
  tmp.n_layers = 1;
  tmp.layers[0] = b.detail;
  const SpecT<float> dl = spec_to<float>(tmp);
  const float nl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  const float pl = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  float cosang = 1.0f;
  if (nl > 0 && pl > 0) cosang = (n[0] * p[0] + n[1] * p[1] + n[2] * p[2]) / (nl * pl);
  const float slope = 1.0f - cosang;
  const float elev = shape_eval(f, p) / b.field.elevation_scale;
  const float d = dl.layers[0].weight * layer_value(dl.layers[0], p);
  const float w = b.band_width;
  const float s1 = smoothstep(b.bands[0] - w, b.bands[0] + w, elev);
  const float s2 = smoothstep(b.bands[1] - w, b.bands[1] + w, elev);
  const float s3 = smoothstep(b.bands[2] - w, b.bands[2] + w, elev);
  const float cl = smoothstep(b.cliff[0], b.cliff[1], slope);
  const float m = 1.0f + b.detail_albedo * d;
  for (int a = 0; a < 3; a++) {
    float c = mixf(b.colors[0][a], b.colors[1][a], s1);
    c = mixf(c, b.colors[2][a], s2);
    c = mixf(c, b.colors[3][a], s3);
    c = mixf(c, b.colors[4][a], cl);
    albedo[a] = sat(c * m);
  }
  rough = sat(b.roughness[0] + b.roughness[1] * slope + b.roughness[2] * d + b.roughness[3] * s3);
  bump = sat(0.5f + b.bump[0] * d + b.bump[1] * elev);
}

inline void bake(const upf_bake_spec& b, const float* P, const float* N, const uint32_t* I, const float* uv,
                 const int32_t* tri, int w, int h, float* albedo, float* bump, float* rough) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const int64_t o = (int64_t)y * w + x;
      const int32_t t = tri[o];
      if (t < 0) {
        albedo[3 * o] = albedo[3 * o + 1] = albedo[3 * o + 2] = NAN;
        bump[o] = rough[o] = NAN;
        continue;
      }
      const float ax = uv[6 * (int64_t)t] * (float)w, ay = uv[6 * (int64_t)t + 1] * (float)h;
      const float bx = uv[6 * (int64_t)t + 2] * (float)w, by = uv[6 * (int64_t)t + 3] * (float)h;
      const float cx = uv[6 * (int64_t)t + 4] * (float)w, cy = uv[6 * (int64_t)t + 5] * (float)h;
      const float px = (float)x + 0.5f, py = (float)y + 0.5f;
      const float area = edge_fn(ax, ay, bx, by, cx, cy);
      const float wt[3] = {edge_fn(bx, by, cx, cy, px, py) / area, edge_fn(cx, cy, ax, ay, px, py) / area,
                           edge_fn(ax, ay, bx, by, px, py) / area};
      const uint32_t iv[3] = {I[3 * (int64_t)t], I[3 * (int64_t)t + 1], I[3 * (int64_t)t + 2]};
      float p[3], n[3];
      for (int a = 0; a < 3; a++) {
        p[a] = wt[0] * P[3 * iv[0] + a] + wt[1] * P[3 * iv[1] + a] + wt[2] * P[3 * iv[2] + a];
        n[a] = wt[0] * N[3 * iv[0] + a] + wt[1] * N[3 * iv[1] + a] + wt[2] * N[3 * iv[2] + a];
      }
      material(b, p, n, albedo + 3 * o, bump[o], rough[o]);
    }
}

// Dilation pass and U8 conversion as documented at upf_image_dilate / upf_image_to_u8
// (the steps of sources/textures.cpp:73-81).
inline void dilate(const float* in, float* out, int w, int h, int C) {
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const int64_t o = ((int64_t)y * w + x) * C;
      if (!std::isnan(in[o])) {
        for (int c = 0; c < C; c++) out[o + c] = in[o + c];
        continue;
      }
      float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      int cnt = 0;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int yy = y + dy, xx = x + dx;
          if ((dx == 0 && dy == 0) || yy < 0 || yy >= h || xx < 0 || xx >= w) continue;
          const int64_t q = ((int64_t)yy * w + xx) * C;
          if (std::isnan(in[q])) continue;
          for (int c = 0; c < C; c++) sum[c] = sum[c] + in[q + c];
          cnt++;
        }
      for (int c = 0; c < C; c++) out[o + c] = cnt ? sum[c] / (float)cnt : NAN;
    }
}

inline void dilate_rounds(float* img, int w, int h, int C, int rounds) {
  std::vector<float> tmp((size_t)w * h * C);
  for (int r = 0; r < rounds; r++) {
    dilate(img, tmp.data(), w, h, C);
    std::copy(tmp.begin(), tmp.end(), img);
  }
}

inline uint8_t to_u8(float v) {
  if (std::isnan(v)) return 0;
  return (uint8_t)(std::min(std::max(v, 0.0f), 1.0f) * 255.0f + 0.5f);
}

// The export conversions of a land chunk's U8 maps (sources/generator.cpp:136-139) as
// documented at upf_image_height_to_normal / upf_image_special_to_gltf_pbr.  cage's
// imageConvertHeigthToNormal and imageConvertSpecialToGltfPbr are not in the checkout.
inline uint8_t unit_to_u8(float n) { return (uint8_t)((n * 0.5f + 0.5f) * 255.0f + 0.5f); }

inline void height_to_normal(const uint8_t* in, int w, int h, float strength, uint8_t* out) {
  auto at = [&](int x, int y) {
    x = std::min(std::max(x, 0), w - 1);
    y = std::min(std::max(y, 0), h - 1);
    return (float)in[(int64_t)y * w + x] * (1.0f / 255.0f);
  };
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const float a = at(x - 1, y - 1), b = at(x, y - 1), c = at(x + 1, y - 1);
      const float d = at(x - 1, y), f = at(x + 1, y);
      const float g = at(x - 1, y + 1), hh = at(x, y + 1), i = at(x + 1, y + 1);
      const float gx = (c + 2.0f * f + i) - (a + 2.0f * d + g);
      const float gy = (g + 2.0f * hh + i) - (a + 2.0f * b + c);
      const float vx = -strength * gx, vy = -strength * gy;
      const float inv = 1.0f / std::sqrt(vx * vx + vy * vy + 1.0f);
      uint8_t* o = out + ((int64_t)y * w + x) * 3;
      o[0] = unit_to_u8(vx * inv);
      o[1] = unit_to_u8(vy * inv);
      o[2] = unit_to_u8(inv);
    }
}

inline void special_to_pbr(const uint8_t* in, int64_t n, uint8_t* out) {
  for (int64_t i = 0; i < n; i++) {
    out[3 * i] = 255;
    out[3 * i + 1] = in[2 * i];
    out[3 * i + 2] = in[2 * i + 1];
  }
}

}  // namespace upfo
