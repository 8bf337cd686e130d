// upf_device_grad.hpp -- hand-written adjoint of the field stack on gfx950.
//
// For one sample p it returns f(p), df/dp (exact normals) and, through a wave-uniform
// accumulator, dLdf * df/dθ for θ = shape parameters, layer weights and per-octave
// frequency/amplitude (the upf_theta_grad layout).  Nothing is stored per octave:
//   pass 1 (only when a MASK/LERP blend is present) evaluates the layer values and
//          the blend chain forward, then runs the chain backward to get the scalar
//          adjoint c_l = d elev / d v_l of every layer;
//   pass 2 recomputes every octave with its analytic lattice gradient and scatters
//          into df/dp and the θ accumulator.
// With only ADD blends c_l = weight_l is known up front and pass 1 is skipped.
#pragma once

#include "upf_device.hpp"

namespace upf {

// ------------------------------------------------------------------ noise + gradient
// Returns n(q) and dn/dq.
__device__ __forceinline__ float noise_gradient_d(uint32_t seed, float qx, float qy, float qz,
                                                  float& dx, float& dy, float& dz) {
  const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
  uint32_t h[8];
  cell_hashes(seed, (int)fx, (int)fy, (int)fz, h);
  const float tx = qx - fx, ty = qy - fy, tz = qz - fz;
  const float sx = tx - 1.0f, sy = ty - 1.0f, sz = tz - 1.0f;
  const float ux = fade(tx), uy = fade(ty), uz = fade(tz);
  float g[8][3];
#pragma unroll
  for (int c = 0; c < 8; c++) gvec(h[c], g[c][0], g[c][1], g[c][2]);
  float d[8];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const float ax = (c & 1) ? sx : tx, ay = (c & 2) ? sy : ty, az = (c & 4) ? sz : tz;
    d[c] = fmaf(g[c][2], az, fmaf(g[c][1], ay, g[c][0] * ax));
  }
  const float x00 = lerpf(d[0], d[1], ux), x10 = lerpf(d[2], d[3], ux);
  const float x01 = lerpf(d[4], d[5], ux), x11 = lerpf(d[6], d[7], ux);
  const float y0 = lerpf(x00, x10, uy), y1 = lerpf(x01, x11, uy);
  // trilinear blend of the corner gradients
  float gb[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float b00 = lerpf(g[0][a], g[1][a], ux), b10 = lerpf(g[2][a], g[3][a], ux);
    const float b01 = lerpf(g[4][a], g[5][a], ux), b11 = lerpf(g[6][a], g[7][a], ux);
    gb[a] = lerpf(lerpf(b00, b10, uy), lerpf(b01, b11, uy), uz);
  }
  const float dnx = lerpf(lerpf(d[1] - d[0], d[3] - d[2], uy), lerpf(d[5] - d[4], d[7] - d[6], uy), uz);
  const float dny = lerpf(x10 - x00, x11 - x01, uz);
  const float dnz = y1 - y0;
  dx = fmaf(dfade(tx), dnx, gb[0]);
  dy = fmaf(dfade(ty), dny, gb[1]);
  dz = fmaf(dfade(tz), dnz, gb[2]);
  return lerpf(y0, y1, uz);
}

__device__ __forceinline__ float noise_value_d(uint32_t seed, float qx, float qy, float qz,
                                               float& dx, float& dy, float& dz) {
  const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
  uint32_t h[8];
  cell_hashes(seed, (int)fx, (int)fy, (int)fz, h);
  const float tx = qx - fx, ty = qy - fy, tz = qz - fz;
  const float ux = fade(tx), uy = fade(ty), uz = fade(tz);
  float v[8];
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = lattice_value(h[c]);
  const float x00 = lerpf(v[0], v[1], ux), x10 = lerpf(v[2], v[3], ux);
  const float x01 = lerpf(v[4], v[5], ux), x11 = lerpf(v[6], v[7], ux);
  const float y0 = lerpf(x00, x10, uy), y1 = lerpf(x01, x11, uy);
  dx = dfade(tx) * lerpf(lerpf(v[1] - v[0], v[3] - v[2], uy), lerpf(v[5] - v[4], v[7] - v[6], uy), uz);
  dy = dfade(ty) * lerpf(x10 - x00, x11 - x01, uz);
  dz = dfade(tz) * (y1 - y0);
  return lerpf(y0, y1, uz);
}

__device__ __forceinline__ float noise_cellular_d(uint32_t seed, float qx, float qy, float qz, float js,
                                                  float j0, float& dx, float& dy, float& dz) {
  float ex, ey, ez;
  const float d2 = cellular_d2(seed, qx, qy, qz, js, j0, ex, ey, ez);
  const float r = sqrt_hw(d2);
  const float inv = r > 0.0f ? 1.0f / r : 0.0f;
  dx = ex * inv; dy = ey * inv; dz = ez * inv;
  return r;
}

// Simplex noise and its gradient: d/dq of t^4 (g . d) with t = 0.5 - |d|^2 is
// t^4 g - 8 t^3 (g . d) d per corner.
__device__ __forceinline__ float noise_simplex_d(uint32_t seed, float qx, float qy, float qz, float& dx, float& dy,
                                                 float& dz) {
  SimplexCell c;
  simplex_cell(qx, qy, qz, c);
  float n = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const float ex = c.d[q][0], ey = c.d[q][1], ez = c.d[q][2];
    const float t = 0.5f - (ex * ex + ey * ey + ez * ez);
    if (t > 0.0f) {
      const uint32_t h = simplex_hash(seed, c, q);
      float vx, vy, vz;
      gvec(h, vx, vy, vz);
      const float dot = gdot(h, ex, ey, ez);
      const float t2 = t * t, t4 = t2 * t2;
      n = n + t4 * dot;
      const float k = -8.0f * (t2 * t) * dot;
      gx = gx + (t4 * vx + k * ex);
      gy = gy + (t4 * vy + k * ey);
      gz = gz + (t4 * vz + k * ez);
    }
  }
  dx = 72.0f * gx; dy = 72.0f * gy; dz = 72.0f * gz;
  return 72.0f * n;
}

__device__ __forceinline__ float octave_noise_d(const DLayer& L, int k, float x, float y, float z,
                                                float& dx, float& dy, float& dz) {
  const float f = L.freq[k];
  const float qx = x * f, qy = y * f, qz = z * f;
  const uint32_t seed = L.seed + (uint32_t)k;
  if (L.noise == UPF_NOISE_GRADIENT) return noise_gradient_d(seed, qx, qy, qz, dx, dy, dz);
  if (L.noise == UPF_NOISE_VALUE) return noise_value_d(seed, qx, qy, qz, dx, dy, dz);
  if (L.noise == UPF_NOISE_SIMPLEX) return noise_simplex_d(seed, qx, qy, qz, dx, dy, dz);
  return noise_cellular_d(seed, qx, qy, qz, L.js, L.j0, dx, dy, dz);
}

// ------------------------------------------------------------------ shapes + gradient
// The further shapes with their position gradient (forward mode), out of line.
__device__ __noinline__ Grad shape_more_d(int shape, float p0, float p1, float p2, float p3, const float* w, float x,
                                          float y, float z) {
  if (shape == UPF_SHAPE_SINE_NET) return sine_net_t<Grad>(w, p0, p1, p2, gvar(x, 0), gvar(y, 1), gvar(z, 2));
  const float sp[4] = {p0, p1, p2, p3};
  if (shape == UPF_SHAPE_ASTEROID) {
    const float l = sqrt_f(x * x + y * y + z * z);
    if (l < 1.0f) return gc(-100.0f);
    const Grad gx = gvar(x, 0), gy = gvar(y, 1), gz = gvar(z, 2);
    const Grad gl = sq_root(gx * gx + gy * gy + gz * gz);
    const Grad dx = gx / gl, dy = gy / gl, dz = gz / gl;
    const uint32_t seed = (uint32_t)p1;
    const float b = 1.0f / 1.3f;
    float a0, a1, a2, c0, c1, c2;
    const float n0 = noise_simplex_d(seed, dx.v * p0, dy.v * p0, dz.v * p0, a0, a1, a2);
    const float f1 = p0 * 2.0f;
    const float n1 = noise_simplex_d(seed + 1u, dx.v * f1, dy.v * f1, dz.v * f1, c0, c1, c2);
    // dn/dd = b (f a + 0.3 (2 f) c), chained through d = p / |p|
    const float k0 = b * p0, k1 = b * 0.3f * f1;
    const float ex = k0 * a0 + k1 * c0, ey = k0 * a1 + k1 * c1, ez = k0 * a2 + k1 * c2;
    const Grad n = {n0 * b + n1 * (b * 0.3f), ex * dx.dx + ey * dy.dx + ez * dz.dx, ex * dx.dy + ey * dy.dy + ez * dz.dy,
                    ex * dx.dz + ey * dy.dz + ez * dz.dz};
    return gl + n * 600.0f - 1300.0f;
  }
  return shape_more_t<Grad>(shape, sp, gvar(x, 0), gvar(y, 1), gvar(z, 2), gc(0.0f));
}

// Returns shape(p); writes d shape/dp and d shape/d params[0..3].
template <bool EXT = true>
__device__ __forceinline__ float shape_eval_d(const DSpec& s, float x, float y, float z, float g[3],
                                              float dp[4]) {
  g[0] = g[1] = g[2] = 0.0f;
  dp[0] = dp[1] = dp[2] = dp[3] = 0.0f;
  switch (s.shape) {
    case UPF_SHAPE_SPHERE: {
      const float r = sqrt_hw(fmaf(z, z, fmaf(y, y, x * x)));
      const float inv = r > 0.0f ? 1.0f / r : 0.0f;
      g[0] = x * inv; g[1] = y * inv; g[2] = z * inv;
      dp[0] = -1.0f;
      return r - s.sp[0];
    }
    case UPF_SHAPE_BOX: {
      const float q[3] = {fabsf(x) - s.sp[0], fabsf(y) - s.sp[1], fabsf(z) - s.sp[2]};
      const float sg[3] = {x < 0.0f ? -1.0f : 1.0f, y < 0.0f ? -1.0f : 1.0f, z < 0.0f ? -1.0f : 1.0f};
      const float m[3] = {fmaxf(q[0], 0.0f), fmaxf(q[1], 0.0f), fmaxf(q[2], 0.0f)};
      const float outside = sqrt_hw(fmaf(m[2], m[2], fmaf(m[1], m[1], m[0] * m[0])));
      const float qm = fmaxf(q[0], fmaxf(q[1], q[2]));
      const float inside = fminf(qm, 0.0f);
      float dq[3];
      if (outside > 0.0f) {
        const float inv = 1.0f / outside;
        for (int a = 0; a < 3; a++) dq[a] = m[a] * inv;
      } else {
        const int a = (q[0] >= q[1] && q[0] >= q[2]) ? 0 : (q[1] >= q[2] ? 1 : 2);
        dq[0] = dq[1] = dq[2] = 0.0f;
        if (qm < 0.0f) dq[a] = 1.0f;
      }
      for (int a = 0; a < 3; a++) { g[a] = sg[a] * dq[a]; dp[a] = -dq[a]; }
      dp[3] = -1.0f;
      return outside + inside - s.sp[3];
    }
    case UPF_SHAPE_TORUS: {
      const float rho = sqrt_hw(fmaf(z, z, x * x));
      const float ring = rho - s.sp[0];
      const float d = sqrt_hw(fmaf(y, y, ring * ring));
      const float invd = d > 0.0f ? 1.0f / d : 0.0f;
      const float invr = rho > 0.0f ? 1.0f / rho : 0.0f;
      const float a = ring * invd;
      g[0] = a * x * invr; g[1] = y * invd; g[2] = a * z * invr;
      dp[0] = -a;
      dp[1] = -1.0f;
      return d - s.sp[1];
    }
    case UPF_SHAPE_CAPSULE: {
      const float h2 = s.sp[0] * 0.5f;
      const float zc = z - fminf(fmaxf(z, -h2), h2);
      const float d = sqrt_hw(fmaf(zc, zc, fmaf(y, y, x * x)));
      const float inv = d > 0.0f ? 1.0f / d : 0.0f;
      g[0] = x * inv; g[1] = y * inv; g[2] = zc * inv;
      dp[0] = z > h2 ? -0.5f * zc * inv : (z < -h2 ? 0.5f * zc * inv : 0.0f);
      dp[1] = -1.0f;
      return d - s.sp[1];
    }
    case UPF_SHAPE_NONE:
      return 0.0f;
    default: {  // extended shapes: forward-mode position gradient, no parameter gradient
      const Grad r = EXT && s.shape >= UPF_SHAPE_H3O ? shape_more_d(s.shape, s.sp[0], s.sp[1], s.sp[2], s.sp[3], s.w, x, y, z)
                                              : shape_ext<Grad>(s.shape, s.sp, gvar(x, 0), gvar(y, 1), gvar(z, 2),
                                                                gc(0.0f));
      g[0] = r.dx; g[1] = r.dy; g[2] = r.dz;
      return r.v;
    }
  }
}

// ------------------------------------------------------------------ θ accumulator
// Per-wave accumulation of dLdf * df/dθ: every call site is wave-uniform; each call
// reduces across the 64 lanes and lane 0 adds into this wave's slice of LDS.
// v plus the value of the lane that DPP control CTRL names, in the rows ROWS enables
// (the other rows add 0).
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_add(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}

// Sum of v over the 64 lanes, complete in lane 63: pair and quad sums (quad_perm 1032,
// 2301), 8- and 16-lane sums (half-row and row mirrors), then rows 1 and 3 add lane 15
// of the row before (row_bcast:15) and rows 2 and 3 add lane 31 (row_bcast:31).  Six DPP
// adds; the butterfly over __shfl_xor was six dependent LDS permutes.
__device__ __forceinline__ float wave_sum_to_lane63(float v) {
  v = dpp_add<0xB1, 0xF>(v);
  v = dpp_add<0x4E, 0xF>(v);
  v = dpp_add<0x141, 0xF>(v);
  v = dpp_add<0x140, 0xF>(v);
  v = dpp_add<0x142, 0xA>(v);
  return dpp_add<0x143, 0xC>(v);
}

struct ThetaAcc {
  float* slice;  // LDS, UPF_THETA_COUNT floats for this wave
  __device__ __forceinline__ void add(int idx, float v) {
    v = wave_sum_to_lane63(v);
    if ((threadIdx.x & 63) == 63) slice[idx] += v;
  }
};

constexpr int kThShape = 0;
constexpr int kThWeight = UPF_MAX_SHAPE_PARAMS;
constexpr int kThFreq = kThWeight + UPF_MAX_LAYERS;
constexpr int kThAmp = kThFreq + UPF_MAX_LAYERS * UPF_MAX_OCTAVES;

__device__ __forceinline__ int shape_param_count(int shape) {
  switch (shape) {
    case UPF_SHAPE_SPHERE: return 1;
    case UPF_SHAPE_BOX: return 4;
    case UPF_SHAPE_TORUS: return 2;
    case UPF_SHAPE_CAPSULE: return 2;
    default: return 0;
  }
}

// Layer contribution with gradient: adds coeff * dv/dp into g, and (THETA) the
// per-octave dθ terms scaled by coeff_theta = dLdf * d f / d v.  Returns v (unweighted).
template <bool THETA>
__device__ __forceinline__ float layer_grad(const DLayer& L, int l, float x, float y, float z, float coeff,
                                            float coeff_theta, float g[3], ThetaAcc& acc) {
  const int oct = L.fractal == UPF_FRACTAL_NONE ? 1 : L.octaves;
  float v = 0.0f;
  for (int k = 0; k < oct; k++) {
    float nx, ny, nz;
    const float n = octave_noise_d(L, k, x, y, z, nx, ny, nz);
    float term, dterm;  // term(n) and d term / d n
    if (L.fractal == UPF_FRACTAL_NONE) { term = n; dterm = 1.0f; }
    else if (L.fractal == UPF_FRACTAL_RIDGED) {
      term = fmaf(-2.0f, fabsf(n), 1.0f);
      dterm = (n < 0.0f ? 2.0f : -2.0f) * L.amp[k];
    } else { term = n; dterm = L.amp[k]; }
    v = L.fractal == UPF_FRACTAL_NONE ? term : fmaf(L.amp[k], term, v);
    const float f = L.freq[k];
    const float s = coeff * dterm * f;  // d/dp = dterm * f * grad_q n
    g[0] = fmaf(s, nx, g[0]); g[1] = fmaf(s, ny, g[1]); g[2] = fmaf(s, nz, g[2]);
    if (THETA) {
      // d term / d freq = dterm * (grad_q n . p)
      const float qdot = fmaf(nz, z, fmaf(ny, y, nx * x));
      acc.add(kThFreq + l * UPF_MAX_OCTAVES + k, coeff_theta * dterm * qdot);
      if (L.fractal != UPF_FRACTAL_NONE) acc.add(kThAmp + l * UPF_MAX_OCTAVES + k, coeff_theta * term);
    }
  }
  return v;
}

// Elevation of a fixed mode with its position gradient (forward mode).  A layer sampled
// at a position that is itself a function of the sample (the legacy mode's p * scale)
// chains its lattice gradient through that position's derivatives.
__device__ __forceinline__ Grad elevation_mode_grad(const DSpec& s, float x, float y, float z) {
  const int mode = s.mode;
  const Grad gx = gvar(x, 0), gy = gvar(y, 1), gz = gvar(z, 2), zero = gc(0.0f);
  ModeState<Grad> st;
  mode_begin(mode, s.L[0].seed, s.L[0].octaves, s.L[0].freq[0], gx, gy, gz, zero, st);
  for (int l = 0; l < s.n_layers; l++) {
    Grad px, py, pz;
    mode_position(mode, l, st, gx, gy, gz, zero, px, py, pz);
    float g[3] = {0.0f, 0.0f, 0.0f};
    ThetaAcc none{nullptr};
    const float v = layer_grad<false>(s.L[l], l, px.v, py.v, pz.v, 1.0f, 0.0f, g, none);
    const Grad lv = {v, g[0] * px.dx + g[1] * py.dx + g[2] * pz.dx, g[0] * px.dy + g[1] * py.dy + g[2] * pz.dy,
                     g[0] * px.dz + g[1] * py.dz + g[2] * pz.dz};
    if (mode_after(mode, l, lv, zero, st)) break;
  }
  return mode_finish(mode, st, gx, gy, gz, zero);
}

// f, df/dp and (THETA) dLdf * df/dθ for one sample.  EXT compiles in the modes and the
// further shapes; with an elevation mode only the position gradient is formed (the
// parameter gradient is defined for the layer stack).
template <bool THETA, bool EXT = false>
__device__ __forceinline__ float field_grad(const DSpec& s, float x, float y, float z, float dLdf, float g[3],
                                            ThetaAcc& acc) {
  float gs[3], dps[4];
  const bool need_shape = s.compose != UPF_COMPOSE_ELEVATION_RAW;
  const bool need_elev = s.compose == UPF_COMPOSE_LAND || s.compose == UPF_COMPOSE_NAVIGATION ||
                         s.compose == UPF_COMPOSE_ELEVATION_RAW;
  float shape = 0.0f;
  float dshape = 0.0f;  // d f / d shape
  if (need_shape) {
    shape = shape_eval_d<EXT>(s, x, y, z, gs, dps);
    dshape = s.compose == UPF_COMPOSE_ELEVATION ? 1.0f / s.escale : 1.0f;
  } else {
    gs[0] = gs[1] = gs[2] = 0.0f;
    dps[0] = dps[1] = dps[2] = dps[3] = 0.0f;
  }
  g[0] = dshape * gs[0]; g[1] = dshape * gs[1]; g[2] = dshape * gs[2];
  if (THETA) {
    const int np = shape_param_count(s.shape);
    for (int i = 0; i < np; i++) acc.add(kThShape + i, dLdf * dshape * dps[i]);
  }
  if (EXT && !THETA && need_elev && s.mode != UPF_ELEV_LAYERS) {
    const Grad e = elevation_mode_grad(s, x, y, z);
    if (s.compose == UPF_COMPOSE_ELEVATION_RAW) {
      g[0] = e.dx; g[1] = e.dy; g[2] = e.dz;
      return e.v;
    }
    const bool on = s.compose == UPF_COMPOSE_LAND || e.v > 0.0f;
    if (on) {
      g[0] = g[0] - s.escale * e.dx; g[1] = g[1] - s.escale * e.dy; g[2] = g[2] - s.escale * e.dz;
    }
    return s.compose == UPF_COMPOSE_LAND ? shape - s.escale * e.v : shape - s.escale * fmaxf(e.v, 0.0f);
  }
  if (!need_elev || s.n_layers == 0) {
    return s.compose == UPF_COMPOSE_ELEVATION ? shape / s.escale
           : (s.compose == UPF_COMPOSE_ELEVATION_RAW ? 0.0f : shape);
  }

  // pass 1: forward values and the backward blend chain.  Skipped when every blend is
  // ADD and the composition is linear in elev (then c_l = weight_l is known up front).
  float c[UPF_MAX_LAYERS];   // d elev / d (unweighted layer value)
  float cw[UPF_MAX_LAYERS];  // d elev / d (weighted layer value)
  float elev = 0.0f;
  const bool pass1 = !s.all_add || s.compose == UPF_COMPOSE_NAVIGATION;
  if (pass1) {
    float vv[UPF_MAX_LAYERS], accin[UPF_MAX_LAYERS], min_[UPF_MAX_LAYERS];
    float acc0 = 0.0f, m = 0.0f;
    for (int l = 0; l < s.n_layers; l++) {
      const DLayer& L = s.L[l];
      const float v = L.weight * layer_value(L, x, y, z);
      vv[l] = v; accin[l] = acc0; min_[l] = m;
      if (L.blend == UPF_BLEND_ADD) acc0 = acc0 + v;
      else if (L.blend == UPF_BLEND_MASK) m = __builtin_amdgcn_fmed3f(v + 0.5f, 0.0f, 1.0f);
      else acc0 = fmaf(m, v - acc0, acc0);
    }
    elev = acc0;
    float abar = 1.0f, mbar = 0.0f;
    for (int l = s.n_layers - 1; l >= 0; l--) {
      const DLayer& L = s.L[l];
      if (L.blend == UPF_BLEND_ADD) cw[l] = abar;
      else if (L.blend == UPF_BLEND_MASK) {
        const float t = vv[l] + 0.5f;
        cw[l] = (t > 0.0f && t < 1.0f) ? mbar : 0.0f;
        mbar = 0.0f;
      } else {
        const float mm = min_[l];
        cw[l] = abar * mm;
        mbar = fmaf(abar, vv[l] - accin[l], mbar);
        abar = abar * (1.0f - mm);
      }
      c[l] = cw[l] * L.weight;
    }
  } else {
    for (int l = 0; l < s.n_layers; l++) { cw[l] = 1.0f; c[l] = s.L[l].weight; }
  }

  // d f / d elev
  float fe = 1.0f;
  if (s.compose == UPF_COMPOSE_LAND) fe = -s.escale;
  else if (s.compose == UPF_COMPOSE_NAVIGATION) fe = elev > 0.0f ? -s.escale : 0.0f;

  // pass 2: recompute every octave with its lattice gradient
  float elev2 = 0.0f;
  for (int l = 0; l < s.n_layers; l++) {
    const DLayer& L = s.L[l];
    const float k = fe * c[l];
    const float v = layer_grad<THETA>(L, l, x, y, z, k, dLdf * k, g, acc);
    elev2 = fmaf(L.weight, v, elev2);
    if (THETA) acc.add(kThWeight + l, dLdf * fe * cw[l] * v);
  }
  if (!pass1) elev = elev2;
  if (s.compose == UPF_COMPOSE_LAND) return shape - s.escale * elev;
  if (s.compose == UPF_COMPOSE_NAVIGATION) return shape - s.escale * fmaxf(elev, 0.0f);
  return elev;  // ELEVATION_RAW
}

}  // namespace upf
