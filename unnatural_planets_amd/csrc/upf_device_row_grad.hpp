// upf_device_row_grad.hpp -- the fused forward + adjoint of upf_device_grad.hpp for grid
// rows: a wavefront evaluates 64 consecutive x nodes of one row, so the lattice work of
// every octave is shared the way upf_device_row.hpp shares it for the forward field,
// now with derivatives.
//
// Collapsed columns with derivatives.  For a column X of a gradient-noise octave (the 4
// lattice points (X, iy + dy, iz + dz), corner c = dy + 2 dz, a_c = g_x, b_c = g_y (ty -
// dy) + g_z (tz - dz)) the writers store three planes of (A, B) pairs:
//   plane 0: A = fold(a), B = fold(b)                                (the forward pair)
//   plane 1: dA/dty = uy' lerp_z(a1 - a0, a3 - a2),
//            dB/dty = fold(g_y) + uy' lerp_z(b1 - b0, b3 - b2)
//   plane 2: dA/dtz = uz' (lerp_y(a2, a3) - lerp_y(a0, a1)),
//            dB/dtz = fold(g_z) + uz' (lerp_y(b2, b3) - lerp_y(b0, b1))
// with fold(v) = lerp_z(lerp_y(v0, v1, uy), lerp_y(v2, v3, uy), uz) and uy' = fade'(ty).
// A lane's octave is n = lerp(a0, a1, ux) with a0 = A0 tx + B0, a1 = A1 (tx - 1) + B1, and
//   dn/dtx = lerp(A0, A1, ux) + ux' (a1 - a0),
//   dn/dty = lerp(A0y tx + B0y, A1y (tx - 1) + B1y, ux)   (and the same for tz),
// which is the per-lane trilinear sum and its gradient with the sums regrouped.  Value
// noise has a_c = the lattice values and B = 0.  Cellular: the writers store per feature
// point (p_x, e, q_y - p_y, q_z - p_z), and a lane keeps the first nearest point in the
// per-lane order, whose offset is the distance gradient.
#pragma once

#include "upf_device_grad.hpp"
#include "upf_device_row.hpp"

namespace upf {

template <int NOISE>
__device__ __forceinline__ void col_write_d(const RowCtx& c, const DLayer& L, int oct, float y, float z,
                                            const int* lo, const int* base, int total, float2* slice) {
  const int lane = lane_of(c);
  if (lane < oct) {
    float f = L.freq[0];
    int off = lo[0] - base[0];
#pragma unroll
    for (int j = 1; j < UPF_MAX_OCTAVES; j++)
      if (j < oct && lane == j) { f = L.freq[j]; off = lo[j] - base[j]; }
    const float qy = y * f, qz = z * f;
    const float fy = floorf(qy), fz = floorf(qz);
    const float ty = qy - fy, tz = qz - fz;
    const uint32_t y0 = (uint32_t)(int)fy * kPY, z0 = (uint32_t)(int)fz * kPZ;
    const uint32_t sk = L.seed + (uint32_t)lane;
    c.oct[3 * lane] = make_float4(ty, tz, fade(ty), fade(tz));
    c.oct[3 * lane + 1] = make_float4(__uint_as_float(sk ^ y0 ^ z0), __uint_as_float(sk ^ (y0 + kPY) ^ z0),
                                        __uint_as_float(sk ^ y0 ^ (z0 + kPZ)),
                                        __uint_as_float(sk ^ (y0 + kPY) ^ (z0 + kPZ)));
    c.oct[3 * lane + 2] = make_float4(__int_as_float(off), dfade(ty), dfade(tz), 0.0f);
  }
  wave_sync();
  for (int rd = 0; rd * 64 < total; rd++) {
    const int t = 64 * rd + lane;
    int k = 0;
#pragma unroll
    for (int j = 1; j < UPF_MAX_OCTAVES; j++)
      if (j < oct) k += t >= base[j];
    const float4 p0 = c.oct[3 * k], p1 = c.oct[3 * k + 1], p2 = c.oct[3 * k + 2];
    const uint32_t hx = (uint32_t)(t + __float_as_int(p2.x)) * kPX;
    const uint32_t h0 = (__float_as_uint(p1.x) ^ hx) * kHM, h1 = (__float_as_uint(p1.y) ^ hx) * kHM;
    const uint32_t h2 = (__float_as_uint(p1.z) ^ hx) * kHM, h3 = (__float_as_uint(p1.w) ^ hx) * kHM;
    const float ty = p0.x, tz = p0.y, uy = p0.z, uz = p0.w, duy = p2.y, duz = p2.z;
    float a0, a1, a2, a3, b0 = 0.0f, b1 = 0.0f, b2 = 0.0f, b3 = 0.0f;
    float gy = 0.0f, gz = 0.0f;  // fold(g_y), fold(g_z)
    if (NOISE == UPF_NOISE_GRADIENT) {
      const float4 g0 = c.gtab[h0 >> 28], g1 = c.gtab[h1 >> 28], g2 = c.gtab[h2 >> 28], g3 = c.gtab[h3 >> 28];
      const float ty1 = ty - 1.0f, tz1 = tz - 1.0f;
      a0 = g0.x; a1 = g1.x; a2 = g2.x; a3 = g3.x;
      b0 = fmaf(g0.z, tz, g0.y * ty); b1 = fmaf(g1.z, tz, g1.y * ty1);
      b2 = fmaf(g2.z, tz1, g2.y * ty); b3 = fmaf(g3.z, tz1, g3.y * ty1);
      gy = lerpf(lerpf(g0.y, g1.y, uy), lerpf(g2.y, g3.y, uy), uz);
      gz = lerpf(lerpf(g0.z, g1.z, uy), lerpf(g2.z, g3.z, uy), uz);
    } else {
      a0 = lattice_value(h0); a1 = lattice_value(h1); a2 = lattice_value(h2); a3 = lattice_value(h3);
    }
    const float ay0 = lerpf(a0, a1, uy), ay1 = lerpf(a2, a3, uy);
    const float by0 = lerpf(b0, b1, uy), by1 = lerpf(b2, b3, uy);
    const float A = lerpf(ay0, ay1, uz), B = lerpf(by0, by1, uz);
    const float Ay = duy * lerpf(a1 - a0, a3 - a2, uz);
    const float By = fmaf(duy, lerpf(b1 - b0, b3 - b2, uz), gy);
    const float Az = duz * (ay1 - ay0);
    const float Bz = fmaf(duz, by1 - by0, gz);
    if (t < total) {
      slice[t] = make_float2(A, B);
      slice[total + t] = make_float2(Ay, By);
      slice[2 * total + t] = make_float2(Az, Bz);
    }
  }
}

// n and dn/dq of one octave from the three planes (`total` entries apart).
template <int NOISE>
__device__ __forceinline__ float col_read_d(float qx, int lo, const float2* region, int total, float& dx, float& dy,
                                            float& dz) {
  const float fx = floorf(qx);
  const float tx = qx - fx, sx = tx - 1.0f;
  const float ux = fade(tx), dux = dfade(tx);
  const int o = (int)fx - lo;
  const float2 e0 = region[o], e1 = region[o + 1];
  const float2 y0 = region[total + o], y1 = region[total + o + 1];
  const float2 z0 = region[2 * total + o], z1 = region[2 * total + o + 1];
  if (NOISE == UPF_NOISE_VALUE) {
    dx = dux * (e1.x - e0.x);
    dy = lerpf(y0.x, y1.x, ux);
    dz = lerpf(z0.x, z1.x, ux);
    return lerpf(e0.x, e1.x, ux);
  }
  const float a0 = fmaf(e0.x, tx, e0.y), a1 = fmaf(e1.x, sx, e1.y);
  dx = fmaf(dux, a1 - a0, lerpf(e0.x, e1.x, ux));
  dy = lerpf(fmaf(y0.x, tx, y0.y), fmaf(y1.x, sx, y1.y), ux);
  dz = lerpf(fmaf(z0.x, tx, z0.y), fmaf(z1.x, sx, z1.y), ux);
  return lerpf(a0, a1, ux);
}

// Cellular with derivatives: 63 lanes write one feature point each as float4.
__device__ __forceinline__ void cell_write_d(const RowCtx& c, const DLayer& L, uint32_t seed, float qy, float qz,
                                             int lo, float4* region) {
  const int cy = __builtin_amdgcn_readfirstlane((int)rintf(qy));
  const int cz = __builtin_amdgcn_readfirstlane((int)rintf(qz));
  const CellLane cl = cell_lane(c.lane);
  const int lx = cl.lx, c9 = cl.c9;
  const int kx = lo - 1 + lx, ky = cy - 1 + cl.dy, kz = cz - 1 + cl.dz;
  const uint32_t h = (seed ^ ((uint32_t)kx * kPX) ^ ((uint32_t)ky * kPY) ^ ((uint32_t)kz * kPZ)) * kHM;
  const uint32_t m = h ^ (h >> 15);
  const float px = fmaf((float)(m & 1023u), L.js, (float)kx - L.j0);
  const float py = fmaf((float)((m >> 10) & 1023u), L.js, (float)ky - L.j0);
  const float pz = fmaf((float)((m >> 20) & 1023u), L.js, (float)kz - L.j0);
  const float ey = py - qy, ez = pz - qz;
  if (c9 < 9) region[lx + 8 * c9] = make_float4(px, fmaf(ez, ez, ey * ey), -ey, -ez);
}

__device__ __forceinline__ float cell_read_d(float qx, const float4* p, float& dx, float& dy, float& dz) {
  float best = __builtin_inff();
  float4 bp = make_float4(qx, 0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int dzc = 0; dzc < 3; dzc++)
#pragma unroll
    for (int dyc = 0; dyc < 3; dyc++)
#pragma unroll
      for (int dxc = 0; dxc < 3; dxc++) {
        const float4 v = p[dxc + 8 * (dyc + 3 * dzc)];
        const float d = v.x - qx;
        const float d2 = fmaf(d, d, v.y);
        if (d2 < best) { best = d2; bp = v; }
      }
  const float r = sqrt_hw(best);
  const float inv = r > 0.0f ? 1.0f / r : 0.0f;
  dx = (qx - bp.x) * inv;
  dy = bp.z * inv;
  dz = bp.w * inv;
  return r;
}

// Row version of layer_grad (same per-octave accumulation); falls back to it per lane when
// the layer's columns do not fit the slice.
template <bool THETA, int NOISE>
__device__ __forceinline__ float row_layer_grad_t(const RowCtx& c, const DLayer& L, int l, float x, float y, float z,
                                                  float coeff, float coeff_theta, float g[3], ThetaAcc& acc) {
  constexpr bool CELL = NOISE == UPF_NOISE_CELLULAR;
  const int oct = L.fractal == UPF_FRACTAL_NONE ? 1 : L.octaves;
  int lo[UPF_MAX_OCTAVES], base[UPF_MAX_OCTAVES];
  int total = 0;
  bool fits = true;
#pragma unroll
  for (int k = 0; k < UPF_MAX_OCTAVES; k++) {
    if (k < oct) {
      const float q = x * L.freq[k];
      const int i0 = CELL ? (int)rintf(q) : (int)floorf(q);
      int hi;
      lane_range(i0, lo[k], hi);
      const int n = hi - lo[k] + (CELL ? 3 : 2);
      base[k] = total;
      total += CELL ? 72 : n;
      fits = fits && (!CELL || n <= 7);
    }
  }
  if (!fits || (CELL ? total > kRowSlice : 3 * total > kRowSlice2))
    return layer_grad<THETA>(L, l, x, y, z, coeff, coeff_theta, g, acc);
  float2* slice2 = reinterpret_cast<float2*>(c.slice);
  if (CELL) {
#pragma unroll
    for (int k = 0; k < UPF_MAX_OCTAVES; k++)
      if (k < oct) {
        const float f = L.freq[k];
        cell_write_d(c, L, L.seed + (uint32_t)k, y * f, z * f, lo[k], c.slice + base[k]);
      }
  } else {
    col_write_d<NOISE>(c, L, oct, y, z, lo, base, total, slice2);
  }
  wave_sync();
  float v = 0.0f;
#pragma unroll
  for (int k = 0; k < UPF_MAX_OCTAVES; k++) {
    if (k < oct) {
      const float f = L.freq[k];
      const float q = x * f;
      float nx, ny, nz;
      const float n = CELL ? cell_read_d(q, c.slice + base[k] + ((int)rintf(q) - lo[k]), nx, ny, nz)
                           : col_read_d<NOISE>(q, lo[k], slice2 + base[k], total, nx, ny, nz);
      float term, dterm;
      if (L.fractal == UPF_FRACTAL_NONE) { term = n; dterm = 1.0f; }
      else if (L.fractal == UPF_FRACTAL_RIDGED) {
        term = fmaf(-2.0f, fabsf(n), 1.0f);
        dterm = (n < 0.0f ? 2.0f : -2.0f) * L.amp[k];
      } else { term = n; dterm = L.amp[k]; }
      v = L.fractal == UPF_FRACTAL_NONE ? term : fmaf(L.amp[k], term, v);
      const float s = coeff * dterm * f;
      g[0] = fmaf(s, nx, g[0]); g[1] = fmaf(s, ny, g[1]); g[2] = fmaf(s, nz, g[2]);
      if (THETA) {
        const float qdot = fmaf(nz, z, fmaf(ny, y, nx * x));
        acc.add(kThFreq + l * UPF_MAX_OCTAVES + k, coeff_theta * dterm * qdot);
        if (L.fractal != UPF_FRACTAL_NONE) acc.add(kThAmp + l * UPF_MAX_OCTAVES + k, coeff_theta * term);
      }
    }
  }
  wave_sync();
  return v;
}

template <bool THETA>
__device__ __forceinline__ float row_layer_grad(const RowCtx& c, const DLayer& L, int l, float x, float y, float z,
                                                float coeff, float coeff_theta, float g[3], ThetaAcc& acc) {
  if (L.noise == UPF_NOISE_GRADIENT)
    return row_layer_grad_t<THETA, UPF_NOISE_GRADIENT>(c, L, l, x, y, z, coeff, coeff_theta, g, acc);
  if (L.noise == UPF_NOISE_VALUE)
    return row_layer_grad_t<THETA, UPF_NOISE_VALUE>(c, L, l, x, y, z, coeff, coeff_theta, g, acc);
  if (L.noise == UPF_NOISE_CELLULAR)
    return row_layer_grad_t<THETA, UPF_NOISE_CELLULAR>(c, L, l, x, y, z, coeff, coeff_theta, g, acc);
  return layer_grad<THETA>(L, l, x, y, z, coeff, coeff_theta, g, acc);  // simplex: per lane
}

// Row version of field_grad with the layer passes ordered so that no layer is evaluated
// twice except a mask's value.  elev = sum over layers of coefficient * weight * value.
// The coefficient of an add layer is abar(l), of a lerp layer abar(l) * m_l, where m_l
// is the mask in effect at l and abar(l) the product of (1 - m) over the lerp layers
// after l: both depend only on the masks' values.  A mask's coefficient is the sum over
// the lerp layers it drives of abar(q) * (v_q - acc before q), where acc is the running
// blend.  So: the masks' values first (one octave each in practice), then the gradients
// of the add and lerp layers with their coefficients known, whose values complete the
// masks' coefficients, then the masks' gradients.  NAVIGATION needs the sign of elev
// before any coefficient, so it takes every layer's value first, as field_grad does.
// Per-layer scalars live in float4 registers (lget / lset): indexing a local array with
// the (uniform, not constant) layer index would put it in scratch memory.
static_assert(UPF_MAX_LAYERS == 4, "per-layer scalars live in a float4");

__device__ __forceinline__ float lget(const float4& v, int l) {
  return l == 0 ? v.x : (l == 1 ? v.y : (l == 2 ? v.z : v.w));
}
__device__ __forceinline__ void lset(float4& v, int l, float x) {
  v.x = l == 0 ? x : v.x;
  v.y = l == 1 ? x : v.y;
  v.z = l == 2 ? x : v.z;
  v.w = l == 3 ? x : v.w;
}

// abar(l): the product of (1 - m) over the lerp layers after l.
__device__ __forceinline__ float lerp_tail(const DSpec& s, const float4& min_, int l) {
  float a = 1.0f;
  for (int q = l + 1; q < s.n_layers; q++)
    if (s.L[q].blend == UPF_BLEND_LERP) a *= 1.0f - lget(min_, q);
  return a;
}

template <bool THETA>
__device__ __forceinline__ float row_field_grad(const RowCtx& c, const DSpec& s, float x, float y, float z,
                                                float dLdf, float g[3], ThetaAcc& acc) {
  float gs[3], dps[4];
  const bool need_shape = s.compose != UPF_COMPOSE_ELEVATION_RAW;
  const bool need_elev = s.compose == UPF_COMPOSE_LAND || s.compose == UPF_COMPOSE_NAVIGATION ||
                         s.compose == UPF_COMPOSE_ELEVATION_RAW;
  float shape = 0.0f, dshape = 0.0f;
  if (need_shape) {
    shape = shape_eval_d<false>(s, x, y, z, gs, dps);
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
  if (!need_elev || s.n_layers == 0) {
    return s.compose == UPF_COMPOSE_ELEVATION ? shape / s.escale
           : (s.compose == UPF_COMPOSE_ELEVATION_RAW ? 0.0f : shape);
  }
  const bool nav = s.compose == UPF_COMPOSE_NAVIGATION;
  const bool blended = !s.all_add;
  // min_: the mask in effect at each layer; mcoef: the masks' coefficients (before fe
  // and the weight); inband: bit l set when mask l is strictly inside its clamp
  float4 min_ = make_float4(0.0f, 0.0f, 0.0f, 0.0f), mcoef = min_;
  uint32_t inband = 0u;
  float elev = 0.0f;
  if (nav || blended) {
    float acc0 = 0.0f, m = 0.0f;
    float4 dl = min_;  // NAVIGATION: v - acc before, per lerp layer
    for (int l = 0; l < s.n_layers; l++) {
      const DLayer& L = s.L[l];
      lset(min_, l, m);
      if (nav || L.blend == UPF_BLEND_MASK) {
        float lv;
        row_layer_value<1>(c, L, x, &y, z, &lv);
        const float wv = L.weight * lv;
        if (L.blend == UPF_BLEND_ADD) acc0 = acc0 + wv;
        else if (L.blend == UPF_BLEND_MASK) {
          m = __builtin_amdgcn_fmed3f(wv + 0.5f, 0.0f, 1.0f);
          inband |= (wv + 0.5f > 0.0f && wv + 0.5f < 1.0f) ? 1u << l : 0u;
        } else {
          lset(dl, l, wv - acc0);
          acc0 = fmaf(m, wv - acc0, acc0);
        }
      }
    }
    elev = acc0;  // complete under NAVIGATION only
    if (nav && blended) {
      int drv = -1;
      for (int l = 0; l < s.n_layers; l++) {
        const int bl = s.L[l].blend;
        if (bl == UPF_BLEND_MASK) drv = l;
        else if (bl == UPF_BLEND_LERP && drv >= 0)
          lset(mcoef, drv, fmaf(lerp_tail(s, min_, l), lget(dl, l), lget(mcoef, drv)));
      }
    }
  }
  float fe = 1.0f;
  if (s.compose == UPF_COMPOSE_LAND) fe = -s.escale;
  else if (nav) fe = elev > 0.0f ? -s.escale : 0.0f;
  // one loop with a single call site of the layer gradient (inlined once): first the add
  // and lerp layers, then the masks
  float acc0 = 0.0f, elev2 = 0.0f;
  int drv = -1;
  const int passes = blended ? 2 * s.n_layers : s.n_layers;
  for (int it = 0; it < passes; it++) {
    const bool second = it >= s.n_layers;
    const int l = second ? it - s.n_layers : it;
    const DLayer& L = s.L[l];
    if (L.blend == UPF_BLEND_MASK) drv = l;
    if ((L.blend == UPF_BLEND_MASK) != second) continue;
    float cwl;
    if (second) cwl = (inband >> l) & 1u ? lget(mcoef, l) : 0.0f;
    else if (!blended) cwl = 1.0f;
    else cwl = lerp_tail(s, min_, l) * (L.blend == UPF_BLEND_LERP ? lget(min_, l) : 1.0f);
    const float k = fe * cwl * L.weight;
    const float v = row_layer_grad<THETA>(c, L, l, x, y, z, k, dLdf * k, g, acc);
    if (THETA) acc.add(kThWeight + l, dLdf * fe * cwl * v);
    if (second) continue;
    const float wv = L.weight * v;
    elev2 = fmaf(L.weight, v, elev2);
    if (blended) {
      if (L.blend == UPF_BLEND_LERP) {
        if (!nav && drv >= 0) lset(mcoef, drv, fmaf(lerp_tail(s, min_, l), wv - acc0, lget(mcoef, drv)));
        acc0 = fmaf(lget(min_, l), wv - acc0, acc0);
      } else {
        acc0 = acc0 + wv;
      }
    }
  }
  if (!nav) elev = blended ? acc0 : elev2;
  if (s.compose == UPF_COMPOSE_LAND) return shape - s.escale * elev;
  if (nav) return shape - s.escale * fmaxf(elev, 0.0f);
  return elev;  // ELEVATION_RAW
}

}  // namespace upf
