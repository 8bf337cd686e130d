// upf_modes.hpp -- the reference's fixed elevation functions (sources/elevationModes.cpp)
// as one template over the scalar type, and the impact Voronoi of the craters mode
// (sources/voronoi.cpp:10-80).
//
// The template is compiled for float (the field kernels), Grad (value and position
// gradient: normals) and the host operation counter (the flop table).  The noise layers
// are evaluated by a caller-supplied functor layer(l, x, y, z) -> T, so the same formulas
// serve the per-lane noise, its forward-mode derivative and the counter.  The operation
// order follows the reference statement by statement (`a *= b * c` is a * (b * c)).
#pragma once

#include <math.h>
#include <stdint.h>

#include "upf_shapes.hpp"

namespace upf {

// ------------------------------------------------------------------ scalar helpers
template <class T> UPF_HD T sat_v(T a, T zero) { return clamp01(a, zero, lift(zero, 1.0f)); }
// cage's one-argument smoothstep / smootherstep: x^2 (3 - 2x) and x^3 (x (6x - 15) + 10)
template <class T> UPF_HD T smoothstep1(T x) { return x * x * (3.0f - 2.0f * x); }
template <class T> UPF_HD T smootherstep1(T x) { return x * x * x * (x * (x * 6.0f - 15.0f) + 10.0f); }
template <class T> UPF_HD T lerp_v(T a, T b, T t) { return a + t * (b - a); }
// terrace, sources/math.cpp:18-26
template <class T> UPF_HD T terrace_v(T x, float steepness, T zero) {
  const T f = floor_v(x);
  const float s1 = steepness > 1.0f ? steepness : 1.0f;
  const float s01 = steepness < 0.0f ? 0.0f : (steepness > 1.0f ? 1.0f : steepness);
  const T t = smootherstep1(sat_v((x - f) * s1, zero)) + f;
  return lerp_v(x, t, lift(zero, s01));
}
// smoothMax, sources/math.cpp:35-38
template <class T> UPF_HD T smooth_max_v(T a, T b, T k, T zero) { return -smooth_min_v(-a, -b, k, zero); }

// ------------------------------------------------------------------ impact Voronoi
// murmur3 32-bit finaliser (Appleby, public domain): the integer hash in place of cage's
// hash(), which the reference's checkout does not contain.
UPF_HD uint32_t fmix32(uint32_t k) {
  k ^= k >> 16;
  k *= 0x85ebca6bu;
  k ^= k >> 13;
  k *= 0xc2b2ae35u;
  k ^= k >> 16;
  return k;
}

// Nearest of the points of the 27 cells around trunc(p / cell_size), each cell holding
// `ppc` points (sources/voronoi.cpp:17-56, 67-77).  Integer arithmetic wraps mod 2^32;
// the point of a state s is ((s % 65536) / 65535 + cell) * cell_size with C's truncating
// %; ties keep the first point in generation order (z, y, x outer to inner).
// Truncation of a cell coordinate to int32, clamped to +-(2^31 - 128) first, so that the
// conversion (and the neighbours' +-1) stays defined for any finite position and cell size
// (identical for every position whose cell index fits).
UPF_HD int32_t cell_index(float v) {
  const float m = 2147483520.0f;
  return (int32_t)(v < -m ? -m : (v > m ? m : v));
}

UPF_HD void voronoi_nearest(uint32_t seed, int ppc, float cell_size, float px, float py, float pz, float c[3]) {
  const uint32_t h0 = fmix32(seed), h1 = fmix32(h0), h2 = fmix32(h1);
  const float freq = 1.0f / cell_size;
  const int32_t cx = cell_index(px * freq), cy = cell_index(py * freq), cz = cell_index(pz * freq);
  float best = __builtin_inff();
  c[0] = c[1] = c[2] = 0.0f;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const int32_t kx = cx + dx, ky = cy + dy, kz = cz + dz;
        uint32_t s0 = (uint32_t)kx, s1 = (uint32_t)ky, s2 = (uint32_t)kz;
        for (int i = 0; i <= ppc; i++) {  // the cell's own mix, then one per point
          const uint32_t t0 = h0 + s0 + s1 + fmix32(s0);
          const uint32_t t1 = h1 + s1 + s2 + fmix32(s1);
          const uint32_t t2 = h2 + s2 + s0 + fmix32(s2);
          s0 = t0; s1 = t1; s2 = t2;
          if (i == 0) continue;
          const float qx = ((float)((int32_t)s0 % 65536) / 65535.0f + (float)kx) * cell_size;
          const float qy = ((float)((int32_t)s1 % 65536) / 65535.0f + (float)ky) * cell_size;
          const float qz = ((float)((int32_t)s2 % 65536) / 65535.0f + (float)kz) * cell_size;
          const float ex = qx - px, ey = qy - py, ez = qz - pz;
          const float d2 = ex * ex + ey * ey + ez * ez;
          if (d2 < best) {
            best = d2;
            c[0] = qx; c[1] = qy; c[2] = qz;
          }
        }
      }
}

// ------------------------------------------------------------------ modes
// A mode runs as begin -> for each layer l in order: position(l), layer value,
// after(l) (which may stop the loop early) -> finish.  This lets a kernel keep ONE call
// site of its layer evaluation for the generic stack and every mode alike.
template <class T>
struct ModeState {
  T v0, v1, v2, v3;     // layer values in order
  T land, cover;        // lakes / islands: shaped land and mountain cover
  float c[3];           // craters: the nearest Voronoi point
  bool flat;            // lakes / islands: cover < 1e-7, the mountains are skipped
};

template <class T> UPF_HD T mode_v(const ModeState<T>& st, int l) {
  return l == 0 ? st.v0 : (l == 1 ? st.v1 : (l == 2 ? st.v2 : st.v3));
}

template <class T>
UPF_HD void mode_begin(int mode, uint32_t vseed, int vppc, float vcell, T x, T y, T z, T zero, ModeState<T>& st) {
  st.v0 = st.v1 = st.v2 = st.v3 = zero;
  st.land = st.cover = zero;
  st.c[0] = st.c[1] = st.c[2] = 0.0f;
  st.flat = false;
  if (mode == UPF_ELEV_CRATERS)  // sources/elevationModes.cpp:194
    voronoi_nearest(vseed, vppc, vcell, value_of(x), value_of(y), value_of(z), st.c);
}

// Where layer l is sampled: the sample itself, except the legacy mode's second layer
// (p * scale, :57-58) and the craters' scale noise (at the Voronoi point, :195).
template <class T>
UPF_HD void mode_position(int mode, int l, const ModeState<T>& st, T x, T y, T z, T zero, T& px, T& py, T& pz) {
  px = x; py = y; pz = z;
  if (mode == UPF_ELEV_LEGACY && l == 1) {
    const T scale = st.v0 * 0.0005f + 0.0015f;
    px = x * scale; py = y * scale; pz = z * scale;
  } else if (mode == UPF_ELEV_CRATERS && l == 1) {
    px = lift(zero, st.c[0]); py = lift(zero, st.c[1]); pz = lift(zero, st.c[2]);
  }
}

// Records layer l's value; returns true when the remaining layers are not needed.
template <class T> UPF_HD bool mode_after(int mode, int l, T v, T zero, ModeState<T>& st) {
  if (l == 0) st.v0 = v;
  else if (l == 1) st.v1 = v;
  else if (l == 2) st.v2 = v;
  else st.v3 = v;
  if ((mode == UPF_ELEV_LAKES || mode == UPF_ELEV_ISLANDS) && l == 0) {
    // elevationLakes / elevationIslands, :143-148 / :165-170
    T land = v * 0.5f + 0.5f;
    land = sat_v(land, zero);
    land = 1.0f - pow_v(land, mode == UPF_ELEV_LAKES ? 1.24f : 0.83f);
    land = land * 2.0f - 1.0f;
    land = land / (abs_v(land) + 0.17f) + 0.15f;
    land = land * 150.0f;
    st.land = land;
    st.cover = 1.0f - sat_v(land * -0.1f, zero);  // commonElevationMountains, :102-104
    st.flat = value_of(st.cover) < 1e-7f;
    return st.flat;
  }
  return false;
}

// Elevation (metres) of mode `mode` (not UPF_ELEV_LAYERS) from the recorded layer values.
template <class T> UPF_HD T mode_finish(int mode, const ModeState<T>& st, T x, T y, T z, T zero) {
  switch (mode) {
    case UPF_ELEV_SIMPLE: {  // :29-31
      T a = st.v0;
      a = -a + 0.3f;
      a = pow_v(a * 1.3f - 0.35f, 3.0f) + 0.1f;
      return 100.0f - a * 1000.0f;
    }
    case UPF_ELEV_LEGACY: {  // :58-63
      T a = st.v1;
      a = a + 0.11f;
      if (value_of(a) < 0.0f) a = -pow_v(-a, 0.85f);
      else a = pow_v(a, 1.7f);
      return a * 2500.0f;
    }
    case UPF_ELEV_LAKES:
    case UPF_ELEV_ISLANDS: {  // commonElevationMountains, :102-122
      if (st.flat) return st.land;
      const T mask = st.v1;
      const T rm = smoothstep1(sat_v(mask * 7.0f - 0.3f, zero));
      const T tm = smoothstep1(sat_v(mask * -7.0f - 1.5f, zero));
      T ridge = st.v2;
      ridge = max_v(ridge - 0.1f, zero);
      ridge = pow_v(ridge, 1.6f);
      ridge = ridge * (rm * st.cover);
      ridge = ridge * 1000.0f;
      T terraces = st.v3;
      terraces = max_v(terraces + 0.1f, zero) * 2.5f;
      terraces = terrace_v(terraces, 4.0f, zero);
      terraces = terraces * (tm * st.cover);
      terraces = terraces * 250.0f;
      return st.land + smooth_max_v(zero, max_v(ridge, terraces), lift(zero, 50.0f), zero);
    }
    case UPF_ELEV_CRATERS: {  // :194-197
      const T cx = lift(zero, st.c[0]), cy = lift(zero, st.c[1]), cz = lift(zero, st.c[2]);
      const T s = 100.0f + st.v1 * 50.0f;
      const T d = len3(x - cx, y - cy, z - cz);
      const T u = sat_v(d / s, zero);
      const T h = (pow_v(u, 4.0f) - 0.9f * pow_v(u, 5.0f) - 0.1f) * 10.0f;
      return 200.0f + h * 190.0f;
    }
    default:  // UPF_ELEV_NONE, :9-12
      return lift(zero, 100.0f);
  }
}

// The whole mode with a layer functor layer(l, x, y, z) -> T (one call site).
template <class T, class LF>
UPF_HD T elevation_mode_v(int mode, uint32_t vseed, int vppc, float vcell, int n_layers, T x, T y, T z,
                          const LF& layer, T zero) {
  ModeState<T> st;
  mode_begin(mode, vseed, vppc, vcell, x, y, z, zero, st);
  for (int l = 0; l < n_layers; l++) {
    T px, py, pz;
    mode_position(mode, l, st, x, y, z, zero, px, py, pz);
    if (mode_after(mode, l, layer(l, px, py, pz), zero, st)) break;
  }
  return mode_finish(mode, st, x, y, z, zero);
}

// Layers each mode reads (0 for the generic stack: any count).
UPF_HD int elevation_mode_layers(int mode) {
  switch (mode) {
    case UPF_ELEV_SIMPLE: return 1;
    case UPF_ELEV_LEGACY: return 2;
    case UPF_ELEV_LAKES: case UPF_ELEV_ISLANDS: return 4;
    case UPF_ELEV_CRATERS: return 2;
    default: return 0;
  }
}

}  // namespace upf
