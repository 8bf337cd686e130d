// upf_device_row.hpp -- wave-cooperative noise for grid rows.
//
// A wavefront evaluates 64 consecutive x nodes of one grid row, so y and z are the
// same in every lane.  Per octave the lattice points the row touches are few (the
// corners of consecutive lanes coincide), so instead of every lane hashing its own 8
// corners (or 27 Voronoi cells), the lanes of the wave hash each lattice point once,
// write its gradient vector / lattice value / feature point into the wave's slice of
// LDS, and every lane then gathers its corners with immediate-offset LDS reads.
// Results are bit-identical to the per-lane functions of upf_device.hpp: the same
// hashes, and dot products with {-1,0,1} gradient vectors round exactly like the
// per-lane signed sums.  Rows whose lattice span does not fit the slice fall back to
// the per-lane functions (a wave-uniform branch).
#pragma once

#include "upf_device.hpp"

namespace upf {

constexpr int kRowSlice = 256;  // float4 entries of LDS per wave

// Gradient vectors of Perlin's 16-entry set, indexed by h >> 28 (see gdot()).
__device__ __forceinline__ float4 gradient_entry(int g) {
  float gx, gy, gz;
  gvec((uint32_t)g << 28, gx, gy, gz);
  return make_float4(gx, gy, gz, 0.0f);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float dot_g(const float4 g, float x, float y, float z) {
  return fmaf(g.z, z, fmaf(g.y, y, g.x * x));
}

struct RowCtx {
  float4* slice;        // this wave's LDS slice (kRowSlice entries)
  const float4* gtab;   // 16 gradient vectors in LDS
  int lane;
};

// Range of an integer that is monotone across the lanes of the wave.
__device__ __forceinline__ void lane_range(int v, int& lo, int& hi) {
  const int a = __builtin_amdgcn_readfirstlane(v);
  const int b = __builtin_amdgcn_readlane(v, 63);
  lo = min(a, b);
  hi = max(a, b);
}

// Per-octave lattice data of one layer lives in the wave's slice in regions packed one
// after another; a batch is as many consecutive octaves as fit.  Writers and readers
// recompute the same (ixmin, nl) from lanes 0 and 63, so the region bases agree.
//   gradient: region of 4*nl entries, entry base + combo*nl + xo holds the gradient
//             vector of lattice point (ixmin + xo, iy + dy, iz + dz), combo = dy + 2 dz
//             (consecutive x in consecutive 16-byte entries: conflict-free reads);
//   value:    region of nl entries, entry base + xo holds the 4 combos' lattice values;
//   cellular: region of 72 entries, entry base + lx + 8*(dy + 3 dz) holds the feature
//             point of cell (cxmin - 1 + lx, cy - 1 + dy, cz - 1 + dz), lx < 7.
struct RowOct {
  int i0, lo, n;  // lane's lattice index, wave minimum, entries of x span
};

__device__ __forceinline__ RowOct row_span(float q, bool cell) {
  RowOct o;
  o.i0 = (int)(cell ? rintf(q) : floorf(q));
  int hi;
  lane_range(o.i0, o.lo, hi);
  o.n = hi - o.lo + (cell ? 3 : 2);
  return o;
}

__device__ __forceinline__ int region_size(int noise, int n) {
  return noise == UPF_NOISE_GRADIENT ? 4 * n : (noise == UPF_NOISE_VALUE ? n : (n <= 7 ? 72 : kRowSlice + 1));
}

__device__ __forceinline__ void row_write(const RowCtx& c, const DLayer& L, uint32_t seed, float qx, float qy,
                                          float qz, const RowOct& o, int base) {
  if (L.noise == UPF_NOISE_CELLULAR) {
    const int cy = (int)rintf(qy), cz = (int)rintf(qz);
    const int lx = c.lane % 7, c9 = c.lane / 7;  // lane 63: c9 = 9, idle
    const int dy = c9 % 3, dz = c9 / 3;
    const int kx = o.lo - 1 + lx, ky = cy - 1 + dy, kz = cz - 1 + dz;
    const uint32_t h = (seed ^ ((uint32_t)kx * kPX) ^ ((uint32_t)ky * kPY) ^ ((uint32_t)kz * kPZ)) * kHM;
    const uint32_t m = h ^ (h >> 15);
    const float px = fmaf((float)(m & 1023u), L.js, (float)kx - L.j0);
    const float py = fmaf((float)((m >> 10) & 1023u), L.js, (float)ky - L.j0);
    const float pz = fmaf((float)((m >> 20) & 1023u), L.js, (float)kz - L.j0);
    if (c9 < 9) c.slice[base + lx + 8 * c9] = make_float4(px, py, pz, 0.0f);
    return;
  }
  const int iy = (int)floorf(qy), iz = (int)floorf(qz);
  const uint32_t y0 = (uint32_t)iy * kPY, y1 = y0 + kPY;
  const uint32_t z0 = (uint32_t)iz * kPZ, z1 = z0 + kPZ;
  const int combo = c.lane >> 4;  // dy + 2 dz
  const uint32_t s = ((combo & 1) ? y1 : y0) ^ ((combo & 2) ? z1 : z0) ^ seed;
  const uint32_t xbase = (uint32_t)o.lo * kPX;
  for (int r = 0; r * 16 < o.n; r++) {
    const int xo = (c.lane & 15) + 16 * r;
    if (xo < o.n) {
      const uint32_t h = (s ^ (xbase + (uint32_t)xo * kPX)) * kHM;
      if (L.noise == UPF_NOISE_GRADIENT) c.slice[base + combo * o.n + xo] = c.gtab[h >> 28];
      else reinterpret_cast<float*>(c.slice + base + xo)[combo] = lattice_value(h);
    }
  }
}

__device__ __forceinline__ float row_read(const RowCtx& c, const DLayer& L, float qx, float qy, float qz,
                                          const RowOct& o, int base) {
  if (L.noise == UPF_NOISE_CELLULAR) {
    const float4* p = c.slice + base + (o.i0 - o.lo);  // cell (cx-1+dx, cy-1+dy, cz-1+dz) at p[dx + 8(dy + 3dz)]
    float best = __builtin_inff();
#pragma unroll
    for (int dz = 0; dz < 3; dz++)
#pragma unroll
      for (int dy = 0; dy < 3; dy++)
#pragma unroll
        for (int dx = 0; dx < 3; dx++) {
          const float4 q = p[dx + 8 * (dy + 3 * dz)];
          const float ex = q.x - qx, ey = q.y - qy, ez = q.z - qz;
          best = fminf(best, fmaf(ez, ez, fmaf(ey, ey, ex * ex)));
        }
    return __builtin_sqrtf(best);
  }
  const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
  const float tx = qx - fx, ty = qy - fy, tz = qz - fz;
  const float ux = fade(tx), uy = fade(ty), uz = fade(tz);
  const float4* e = c.slice + base + (o.i0 - o.lo);
  if (L.noise == UPF_NOISE_GRADIENT) {
    const float sx = tx - 1.0f, sy = ty - 1.0f, sz = tz - 1.0f;
    const int n = o.n;
    // e[combo * n + dx]
    const float x00 = lerpf(dot_g(e[0], tx, ty, tz), dot_g(e[1], sx, ty, tz), ux);
    const float x10 = lerpf(dot_g(e[n], tx, sy, tz), dot_g(e[n + 1], sx, sy, tz), ux);
    const float x01 = lerpf(dot_g(e[2 * n], tx, ty, sz), dot_g(e[2 * n + 1], sx, ty, sz), ux);
    const float x11 = lerpf(dot_g(e[3 * n], tx, sy, sz), dot_g(e[3 * n + 1], sx, sy, sz), ux);
    return lerpf(lerpf(x00, x10, uy), lerpf(x01, x11, uy), uz);
  }
  const float4 v0 = e[0], v1 = e[1];  // .x .y .z .w = combos (dy, dz) = 00 10 01 11
  const float x00 = lerpf(v0.x, v1.x, ux), x10 = lerpf(v0.y, v1.y, ux);
  const float x01 = lerpf(v0.z, v1.z, ux), x11 = lerpf(v0.w, v1.w, ux);
  return lerpf(lerpf(x00, x10, uy), lerpf(x01, x11, uy), uz);
}

__device__ __forceinline__ float row_layer_value(const RowCtx& c, const DLayer& L, float x, float y, float z) {
  const int oct = L.fractal == UPF_FRACTAL_NONE ? 1 : L.octaves;
  const bool cell = L.noise == UPF_NOISE_CELLULAR;
  float acc = 0.0f;
  int k = 0;
  while (k < oct) {
    // writers: as many octaves as fit in the slice
    const int k0 = k;
    int base = 0;
    while (k < oct) {
      const float f = L.freq[k];
      const float qx = x * f, qy = y * f, qz = z * f;
      const RowOct o = row_span(qx, cell);
      const int sz = region_size(L.noise, o.n);
      if (base + sz > kRowSlice) break;
      row_write(c, L, L.seed + (uint32_t)k, qx, qy, qz, o, base);
      base += sz;
      k++;
    }
    if (k == k0) {  // one octave too wide for the slice: per-lane evaluation
      const float n = octave_noise(L, k, x, y, z);
      acc = L.fractal == UPF_FRACTAL_NONE ? n
            : (L.fractal == UPF_FRACTAL_RIDGED ? fmaf(L.amp[k], fmaf(-2.0f, fabsf(n), 1.0f), acc)
                                               : fmaf(L.amp[k], n, acc));
      k++;
      continue;
    }
    wave_sync();
    base = 0;
    for (int kk = k0; kk < k; kk++) {
      const float f = L.freq[kk];
      const float qx = x * f, qy = y * f, qz = z * f;
      const RowOct o = row_span(qx, cell);
      const float n = row_read(c, L, qx, qy, qz, o, base);
      base += region_size(L.noise, o.n);
      acc = L.fractal == UPF_FRACTAL_NONE ? n
            : (L.fractal == UPF_FRACTAL_RIDGED ? fmaf(L.amp[kk], fmaf(-2.0f, fabsf(n), 1.0f), acc)
                                               : fmaf(L.amp[kk], n, acc));
    }
    wave_sync();
  }
  return acc;
}

__device__ __forceinline__ float row_elevation(const RowCtx& c, const DSpec& s, float x, float y, float z) {
  float acc = 0.0f, m = 0.0f;
  for (int l = 0; l < s.n_layers; l++) {
    const DLayer& L = s.L[l];
    const float v = L.weight * row_layer_value(c, L, x, y, z);
    if (L.blend == UPF_BLEND_ADD) acc = acc + v;
    else if (L.blend == UPF_BLEND_MASK) m = __builtin_amdgcn_fmed3f(v + 0.5f, 0.0f, 1.0f);
    else acc = fmaf(m, v - acc, acc);
  }
  return acc;
}

__device__ __forceinline__ float row_field(const RowCtx& c, const DSpec& s, float x, float y, float z) {
  const bool need_elev = s.compose == UPF_COMPOSE_LAND || s.compose == UPF_COMPOSE_NAVIGATION ||
                         s.compose == UPF_COMPOSE_ELEVATION_RAW;
  const float shape = s.compose == UPF_COMPOSE_ELEVATION_RAW ? 0.0f : shape_eval(s, x, y, z);
  const float elev = need_elev ? row_elevation(c, s, x, y, z) : 0.0f;
  switch (s.compose) {
    case UPF_COMPOSE_LAND: return shape - s.escale * elev;
    case UPF_COMPOSE_WATER: return shape;
    case UPF_COMPOSE_NAVIGATION: return shape - s.escale * fmaxf(elev, 0.0f);
    case UPF_COMPOSE_ELEVATION: return shape / s.escale;
    default: return elev;
  }
}

}  // namespace upf
