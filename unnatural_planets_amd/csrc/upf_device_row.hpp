// upf_device_row.hpp -- wave-cooperative noise for grid rows.
//
// A wavefront evaluates 64 consecutive x nodes of one grid row, so y and z are the
// same in every lane.  Per octave the lattice points the row touches are few (the
// corners of consecutive lanes coincide), so instead of every lane hashing its own 8
// corners (or 27 Voronoi cells), the lanes of the wave hash each lattice point once,
// write its gradient vector / lattice value / feature point into the wave's slice of
// LDS, and every lane then gathers its corners with immediate-offset LDS reads.
// The lattice (floors, hashes, gradient vectors, feature points) is exactly the
// per-lane one of upf_device.hpp; the interpolation sums are regrouped (see "Collapsed
// columns" below) and agree with the per-lane results to a few ulps.  Rows whose
// lattice span does not fit the slice fall back to the per-lane functions (a
// wave-uniform branch).
#pragma once

#include "upf_device.hpp"

namespace upf {

constexpr int kRowSlice = 256;  // float4 entries of LDS per wave
constexpr int kRowSlice2 = 2 * kRowSlice;  // the same slice as float2 entries

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

constexpr int kRowMaxRows = 2;                                 // rows per wave at most
constexpr int kRowOctTab = 3 * kRowMaxRows * UPF_MAX_OCTAVES;  // float4 entries of the per-octave table

struct RowCtx {
  float4* slice;        // this wave's LDS slice (kRowSlice entries)
  const float4* gtab;   // 16 gradient vectors in LDS
  float4* oct;          // this wave's per-octave table (kRowOctTab entries)
  int lane;
};

// The lane index, read where it is used.  The empty asm keeps the compiler from hoisting
// lane-derived values (LDS addresses, lane masks, hash offsets) to the kernel's start,
// where, live across the whole layer loop, they were spilled to scratch; recomputed at
// each layer call they cost one or two VALU instructions each.
__device__ __forceinline__ int lane_of(const RowCtx& c) {
  int l = c.lane;
  asm volatile("" : "+v"(l));
  return l;
}

// The lane's feature point in the cellular writers: (lx, c9) = (lane % 7, lane / 7),
// c9 = dy + 3 dz.  The lane index passes through an empty asm so that the compiler cannot
// hoist these values to the kernel's start: held in registers across the whole layer
// loop they were spilled to scratch (1.8 KB per wave, ~450 MB of HBM writes per 512^3
// launch); recomputed at the writer they cost a few VALU instructions per call.
struct CellLane {
  int lx, c9, dy, dz;
};
__device__ __forceinline__ CellLane cell_lane(int lane) {
  asm volatile("" : "+v"(lane));
  const int c9 = lane / 7;
  return {lane % 7, c9, c9 % 3, c9 / 3};
}

__device__ __forceinline__ RowCtx make_row_ctx(float4* slice, const float4* gtab, float4* oct, int lane) {
  RowCtx c;
  c.slice = slice;
  c.gtab = gtab;
  c.oct = oct;
  c.lane = lane;
  return c;
}

// Range of an integer that is monotone across the lanes of the wave.
__device__ __forceinline__ void lane_range(int v, int& lo, int& hi) {
  const int a = __builtin_amdgcn_readfirstlane(v);
  const int b = __builtin_amdgcn_readlane(v, 63);
  lo = min(a, b);
  hi = max(a, b);
}

// Collapsed columns.  In a row y and z are the same in every lane, so the y and z part
// of the trilinear interpolation is the same for every sample that uses a given lattice
// column X (the 4 lattice points (X, iy + dy, iz + dz)).  The writers fold it once per
// column:
//   gradient: A(X) = fold(g_x),  B(X) = fold(g_y (ty - dy) + g_z (tz - dz))
//   value:    A(X) = fold(v),    B(X) = 0
//   fold(a)   = lerp(lerp(a_00, a_10, uy), lerp(a_01, a_11, uy), uz)   (index dy dz)
// and a lane's octave becomes lerp(A(ix) tx + B(ix), A(ix+1) (tx - 1) + B(ix+1), ux):
// the same trilinear sum with the x interpolation done last.  The lattice (floors,
// hashes, gradients) is exactly the per-lane one; the sums round in another order, a
// few ulps apart.  Cellular: the writers store per feature point (p_x, e) with
// e = (p_y - q_y)^2 + (p_z - q_z)^2, so a lane's squared distance is one fma.
//
// Layout: the wave's slice as float2 entries, a region per octave of the layer packed
// one after the other (a layer whose regions do not all fit is evaluated per lane):
//   gradient, value: nl entries, entry xo = column ixmin + xo;
//   cellular:        72 entries, entry lx + 8 (dy + 3 dz) = cell (cxmin-1+lx, cy-1+dy, cz-1+dz).

// quad-permute (DPP) partners inside groups of 4 lanes: lane ^ 1 and lane ^ 2
__device__ __forceinline__ float quad_xor1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}
__device__ __forceinline__ float quad_xor2(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));
}

__device__ __forceinline__ float fractal_acc(const DLayer& L, int k, float n, float acc) {
  return L.fractal == UPF_FRACTAL_NONE ? n
         : (L.fractal == UPF_FRACTAL_RIDGED ? fmaf(L.amp[k], fmaf(-2.0f, fabsf(n), 1.0f), acc)
                                            : fmaf(L.amp[k], n, acc));
}

// Gradient / value columns of a single octave (one-octave layers, where the table set-up
// of col_write_all would cost more than it saves).  Lane = 4 * column + combo (combo =
// dy + 2 dz), 16 columns per round; the y and z folds take the partner lanes lane ^ 1 and
// lane ^ 2, in the same order as col_write_all, so the entries are the same.
template <int NOISE>
__device__ __forceinline__ void col_write(const RowCtx& c, uint32_t seed, float qy, float qz, int lo, int n,
                                          float2* region) {
  const float fy = floorf(qy), fz = floorf(qz);
  const float ty = qy - fy, tz = qz - fz;
  const float uy = fade(ty), uz = fade(tz);
  const int iy = __builtin_amdgcn_readfirstlane((int)fy), iz = __builtin_amdgcn_readfirstlane((int)fz);
  const uint32_t y0 = (uint32_t)iy * kPY, z0 = (uint32_t)iz * kPZ;
  const int lane = lane_of(c);
  const int combo = lane & 3;
  const uint32_t s = seed ^ ((combo & 1) ? y0 + kPY : y0) ^ ((combo & 2) ? z0 + kPZ : z0);
  const float ay = (combo & 1) ? ty - 1.0f : ty, az = (combo & 2) ? tz - 1.0f : tz;
  const uint32_t xbase = (uint32_t)lo * kPX + (uint32_t)(lane >> 2) * kPX;
  for (int r = 0; r * 16 < n; r++) {
    const uint32_t h = (s ^ (xbase + (uint32_t)(16 * r) * kPX)) * kHM;
    float a, b;
    if (NOISE == UPF_NOISE_GRADIENT) {
      const float4 g = c.gtab[h >> 28];
      a = g.x;
      b = fmaf(g.z, az, g.y * ay);
    } else {
      a = lattice_value(h);
      b = 0.0f;
    }
    const float ya = lerpf(a, quad_xor1(a), uy);
    const float za = lerpf(ya, quad_xor2(ya), uz);
    float zb = 0.0f;
    if (NOISE == UPF_NOISE_GRADIENT) {
      const float yb = lerpf(b, quad_xor1(b), uy);
      zb = lerpf(yb, quad_xor2(yb), uz);
    }
    const int xo = (lane >> 2) + 16 * r;
    if (combo == 0 && xo < n) region[xo] = make_float2(za, zb);
  }
}

// The single-octave columns of R rows (qy[R]; shared z and x) in one round when each row
// needs at most 16 / R: lanes (64 / R) h + 4 col + combo, the same lattice points, hashes
// and folds as col_write (so the same entries), row h's at region + h * stride.
template <int NOISE, int R>
__device__ __forceinline__ void col_write_group(const RowCtx& c, uint32_t seed, const float* qyr, float qz, int lo,
                                                int n, float2* region, int stride) {
  const int lane = lane_of(c);
  const int h = lane / (64 / R);
  float qy = qyr[0];
#pragma unroll
  for (int q = 1; q < R; q++)
    if (h == q) qy = qyr[q];
  const float fy = floorf(qy), fz = floorf(qz);
  const float ty = qy - fy, tz = qz - fz;
  const float uy = fade(ty), uz = fade(tz);
  const int iz = __builtin_amdgcn_readfirstlane((int)fz);
  const uint32_t y0 = (uint32_t)(int)fy * kPY, z0 = (uint32_t)iz * kPZ;
  const int combo = lane & 3;
  const uint32_t s = seed ^ ((combo & 1) ? y0 + kPY : y0) ^ ((combo & 2) ? z0 + kPZ : z0);
  const float ay = (combo & 1) ? ty - 1.0f : ty, az = (combo & 2) ? tz - 1.0f : tz;
  const int xo = (lane >> 2) & (16 / R - 1);
  const uint32_t hh = (s ^ ((uint32_t)lo * kPX + (uint32_t)xo * kPX)) * kHM;
  float a, b;
  if (NOISE == UPF_NOISE_GRADIENT) {
    const float4 g = c.gtab[hh >> 28];
    a = g.x;
    b = fmaf(g.z, az, g.y * ay);
  } else {
    a = lattice_value(hh);
    b = 0.0f;
  }
  const float ya = lerpf(a, quad_xor1(a), uy);
  const float za = lerpf(ya, quad_xor2(ya), uz);
  float zb = 0.0f;
  if (NOISE == UPF_NOISE_GRADIENT) {
    const float yb = lerpf(b, quad_xor1(b), uy);
    zb = lerpf(yb, quad_xor2(yb), uz);
  }
  if (combo == 0 && xo < n) region[h * stride + xo] = make_float2(za, zb);
}

// One lattice column X of a (row, octave): p0 = (ty, tz, uy, uz), p1 = the 4 combo seeds
// (layout below); its 4 lattice points hashed and folded in y, then z.  col_fold: column
// task t of a packed task list with the table entry e (written by the task list's first
// lanes), into the task's entry of the slice.
template <int NOISE>
__device__ __forceinline__ float2 fold_column(const RowCtx& c, const float4 p0, const float4 p1, uint32_t X) {
  const uint32_t hx = X * kPX;
  const uint32_t h0 = (__float_as_uint(p1.x) ^ hx) * kHM, h1 = (__float_as_uint(p1.y) ^ hx) * kHM;
  const uint32_t h2 = (__float_as_uint(p1.z) ^ hx) * kHM, h3 = (__float_as_uint(p1.w) ^ hx) * kHM;
  const float ty = p0.x, tz = p0.y, uy = p0.z, uz = p0.w;
  float A, B = 0.0f;
  if (NOISE == UPF_NOISE_GRADIENT) {
    const float4 g0 = c.gtab[h0 >> 28], g1 = c.gtab[h1 >> 28], g2 = c.gtab[h2 >> 28], g3 = c.gtab[h3 >> 28];
    const float ty1 = ty - 1.0f, tz1 = tz - 1.0f;
    A = lerpf(lerpf(g0.x, g1.x, uy), lerpf(g2.x, g3.x, uy), uz);
    const float b0 = fmaf(g0.z, tz, g0.y * ty), b1 = fmaf(g1.z, tz, g1.y * ty1);
    const float b2 = fmaf(g2.z, tz1, g2.y * ty), b3 = fmaf(g3.z, tz1, g3.y * ty1);
    B = lerpf(lerpf(b0, b1, uy), lerpf(b2, b3, uy), uz);
  } else {
    A = lerpf(lerpf(lattice_value(h0), lattice_value(h1), uy), lerpf(lattice_value(h2), lattice_value(h3), uy), uz);
  }
  return make_float2(A, B);
}

template <int NOISE>
__device__ __forceinline__ void col_fold(const RowCtx& c, int e, int t, bool store, float2* slice) {
  const float4 p0 = c.oct[3 * e], p1 = c.oct[3 * e + 1];
  const int off = __float_as_int(c.oct[3 * e + 2].x);
  const float2 v = fold_column<NOISE>(c, p0, p1, (uint32_t)(t + off));
  if (store) slice[t] = v;
}

// Columns of every octave of a gradient / value layer at once, for R rows that share x
// (same column spans) and z and differ in y: column task t (region entry t of the packed
// regions, row after row) goes to lane t % 64 of round t / 64, whatever its row and
// octave, and that lane hashes and folds the column's 4 lattice points itself.  Lanes
// 0..R*oct-1 (lane = r * oct + k) first put their row's octave constants into the wave's
// table:
//   [3e] = (ty, tz, uy, uz),  [3e+1] = combo seeds s_c = (seed+k) ^ (iy+dy) PY ^ (iz+dz) PZ
//   (c = dy + 2 dz, as bits),  [3e+2].x = lo_k - base_k - r * total (as bits).
// The lattice points, hashes and gradient vectors are those of the per-lane noise; each
// column folds y first (uy), then z (uz).  The rounds are the packed task count / 64, not
// one or more per octave, and the y/z constants are computed once per octave and row.
template <int NOISE, int R>
__device__ __forceinline__ void col_write_all(const RowCtx& c, const DLayer& L, int oct, const float* y, float z,
                                              const int* lo, const int* base, int total, float2* slice) {
  const int lane = lane_of(c);
  if (lane < R * oct) {
    const int r = R == 1 ? 0 : (R == 2 ? (lane >= oct ? 1 : 0) : lane / oct);
    const int k = lane - r * oct;
    float f = L.freq[0];
    int off = lo[0] - base[0];
#pragma unroll
    for (int j = 1; j < UPF_MAX_OCTAVES; j++)
      if (j < oct && k == j) { f = L.freq[j]; off = lo[j] - base[j]; }
    off -= r * total;
    float yr = y[0];
#pragma unroll
    for (int q = 1; q < R; q++)
      if (r == q) yr = y[q];
    const float qy = yr * f, qz = z * f;
    const float fy = floorf(qy), fz = floorf(qz);
    const float ty = qy - fy, tz = qz - fz;
    const uint32_t y0 = (uint32_t)(int)fy * kPY, z0 = (uint32_t)(int)fz * kPZ;
    const uint32_t sk = L.seed + (uint32_t)k;
    c.oct[3 * lane] = make_float4(ty, tz, fade(ty), fade(tz));
    c.oct[3 * lane + 1] = make_float4(__uint_as_float(sk ^ y0 ^ z0), __uint_as_float(sk ^ (y0 + kPY) ^ z0),
                                        __uint_as_float(sk ^ y0 ^ (z0 + kPZ)),
                                        __uint_as_float(sk ^ (y0 + kPY) ^ (z0 + kPZ)));
    c.oct[3 * lane + 2] = make_float4(__int_as_float(off), 0.0f, 0.0f, 0.0f);
  }
  wave_sync();
  for (int rd = 0; rd * 64 < R * total; rd++) {
    const int t = 64 * rd + lane;
    int r = 0;
#pragma unroll
    for (int q = 1; q < R; q++) r += t >= q * total;
    const int tt = t - r * total;
    int k = 0;
#pragma unroll
    for (int j = 1; j < UPF_MAX_OCTAVES; j++)
      if (j < oct) k += tt >= base[j];
    col_fold<NOISE>(c, r * oct + k, t, t < R * total, slice);
  }
}

// A lane's octave from the columns of R rows: the x part (floor, t, fade, entry offset)
// is computed once and each row reads its own region (the rows' regions are `stride`
// entries apart).
template <int NOISE, int R>
__device__ __forceinline__ void col_read(float qx, int lo, const float2* region, int stride, float* out) {
  const float fx = floorf(qx);
  const float tx = qx - fx;
  const float ux = fade(tx);
  const int o = (int)fx - lo;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const float2* e = region + r * stride + o;
    const float2 e0 = e[0], e1 = e[1];
    if (NOISE == UPF_NOISE_VALUE) {
      out[r] = lerpf(e0.x, e1.x, ux);
    } else {
      const float d0 = fmaf(e0.x, tx, e0.y);
      const float d1 = fmaf(e1.x, tx - 1.0f, e1.y);
      out[r] = lerpf(d0, d1, ux);
    }
  }
}

// Cellular: 63 lanes generate one feature point each, (lx, c9) = (lane % 7, lane / 7).
__device__ __forceinline__ void cell_write(const RowCtx& c, const DLayer& L, uint32_t seed, float qy, float qz,
                                           int lo, float2* region) {
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
  if (c9 < 9) region[lx + 8 * c9] = make_float2(px, fmaf(ez, ez, ey * ey));
}

__device__ __forceinline__ float cell_read(float qx, const float2* p) {
  // cell (cx-1+dx, cy-1+dy, cz-1+dz) at p[dx + 8(dy+3dz)], p = region + rint(qx) - lo
  float best = __builtin_inff();
#pragma unroll
  for (int dz = 0; dz < 3; dz++)
#pragma unroll
    for (int dy = 0; dy < 3; dy++)
#pragma unroll
      for (int dx = 0; dx < 3; dx++) {
        const float2 v = p[dx + 8 * (dy + 3 * dz)];
        const float d = v.x - qx;
        best = fminf(best, fmaf(d, d, v.y));
      }
  return sqrt_hw(best);
}

// One layer over R rows that share x and z (y[r] per row).  The octave loops are
// unrolled, so each octave's span (lo, n) and region base are wave-uniform scalars
// computed once; row r's regions follow row r-1's (total entries apart).  When the
// layer's regions do not fit the slice the layer is evaluated per lane.
template <int NOISE, int R>
__device__ __forceinline__ void row_layer_t(const RowCtx& c, const DLayer& L, float x, const float* y, float z,
                                            float* out) {
  constexpr bool CELL = NOISE == UPF_NOISE_CELLULAR;
  const int oct = L.fractal == UPF_FRACTAL_NONE ? 1 : L.octaves;
  float2* slice = reinterpret_cast<float2*>(c.slice);
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
  if (!fits || total > kRowSlice2) {
#pragma unroll
    for (int r = 0; r < R; r++) out[r] = layer_value(L, x, y[r], z);
    return;
  }
  if (R > 1 && R * total > kRowSlice2) {  // one row at a time: the same results as R = 1
#pragma unroll
    for (int r = 0; r < R; r++) row_layer_t<NOISE, 1>(c, L, x, y + r, z, out + r);
    return;
  }
  if (CELL) {
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int k = 0; k < UPF_MAX_OCTAVES; k++) {
        if (k < oct) {
          const float f = L.freq[k];
          cell_write(c, L, L.seed + (uint32_t)k, y[r] * f, z * f, lo[k], slice + r * total + base[k]);
        }
      }
  } else if (oct == 1 && R > 1 && total <= 16 / R) {
    float qy[R];
#pragma unroll
    for (int r = 0; r < R; r++) qy[r] = y[r] * L.freq[0];
    col_write_group<NOISE, R>(c, L.seed, qy, z * L.freq[0], lo[0], total, slice, total);
  } else if (oct == 1) {
#pragma unroll
    for (int r = 0; r < R; r++) col_write<NOISE>(c, L.seed, y[r] * L.freq[0], z * L.freq[0], lo[0], total, slice + r * total);
  } else {
    col_write_all<NOISE, R>(c, L, oct, y, z, lo, base, total, slice);
  }
  wave_sync();
  float acc[R];
#pragma unroll
  for (int r = 0; r < R; r++) acc[r] = 0.0f;
#pragma unroll
  for (int k = 0; k < UPF_MAX_OCTAVES; k++) {
    if (k < oct) {
      const float q = x * L.freq[k];
      float n[R];
      if (CELL) {
        const float2* p = slice + base[k] + ((int)rintf(q) - lo[k]);
#pragma unroll
        for (int r = 0; r < R; r++) n[r] = cell_read(q, p + r * total);
      } else {
        col_read<NOISE, R>(q, lo[k], slice + base[k], total, n);
      }
#pragma unroll
      for (int r = 0; r < R; r++) acc[r] = fractal_acc(L, k, n[r], acc[r]);
    }
  }
  wave_sync();
#pragma unroll
  for (int r = 0; r < R; r++) out[r] = acc[r];
}

template <int R>
__device__ __forceinline__ void row_layer_value(const RowCtx& c, const DLayer& L, float x, const float* y, float z,
                                                float* out) {
  if (L.noise == UPF_NOISE_GRADIENT) row_layer_t<UPF_NOISE_GRADIENT, R>(c, L, x, y, z, out);
  else if (L.noise == UPF_NOISE_VALUE) row_layer_t<UPF_NOISE_VALUE, R>(c, L, x, y, z, out);
  else if (L.noise == UPF_NOISE_CELLULAR) row_layer_t<UPF_NOISE_CELLULAR, R>(c, L, x, y, z, out);
  else {  // simplex: its lattice is skewed, not axis-aligned columns; per lane
#pragma unroll
    for (int r = 0; r < R; r++) out[r] = layer_value(L, x, y[r], z);
  }
}

template <int R>
__device__ __forceinline__ void row_elevation(const RowCtx& c, const DSpec& s, float x, const float* y, float z,
                                              float* out) {
  float acc[R], m[R];
#pragma unroll
  for (int r = 0; r < R; r++) acc[r] = m[r] = 0.0f;
  for (int l = 0; l < s.n_layers; l++) {
    const DLayer& L = s.L[l];
    float v[R];
    row_layer_value<R>(c, L, x, y, z, v);
#pragma unroll
    for (int r = 0; r < R; r++) {
      const float w = L.weight * v[r];
      if (L.blend == UPF_BLEND_ADD) acc[r] = acc[r] + w;
      else if (L.blend == UPF_BLEND_MASK) m[r] = __builtin_amdgcn_fmed3f(w + 0.5f, 0.0f, 1.0f);
      else acc[r] = fmaf(m[r], w - acc[r], acc[r]);
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) out[r] = acc[r];
}

// f at (x, y[r], z) for the R rows of the wave.
template <int R, bool COMPOSED = true>
__device__ __forceinline__ void row_field(const RowCtx& c, const DSpec& s, float x, const float* y, float z,
                                          float* out) {
  const bool need_elev = s.compose == UPF_COMPOSE_LAND || s.compose == UPF_COMPOSE_NAVIGATION ||
                         s.compose == UPF_COMPOSE_ELEVATION_RAW;
  float elev[R];
  if (need_elev) row_elevation<R>(c, s, x, y, z, elev);
#pragma unroll
  for (int r = 0; r < R; r++) {
    const float shape = s.compose == UPF_COMPOSE_ELEVATION_RAW ? 0.0f : shape_eval<false, COMPOSED>(s, x, y[r], z);
    const float e = need_elev ? elev[r] : 0.0f;
    switch (s.compose) {
      case UPF_COMPOSE_LAND: out[r] = shape - s.escale * e; break;
      case UPF_COMPOSE_WATER: out[r] = shape; break;
      case UPF_COMPOSE_NAVIGATION: out[r] = shape - s.escale * fmaxf(e, 0.0f); break;
      case UPF_COMPOSE_ELEVATION: out[r] = shape / s.escale; break;
      default: out[r] = e; break;
    }
  }
}

// ------------------------------------------------------------------ brick rows
// A wave over the rows of a brick that share their x nodes but not y and z (the rows of
// 9 nodes of a narrow-band brick): M groups of kBrickRows rows, row rho = kBrickRows m + r,
// lane r * nxr + i holding node i of row r of every group (the lanes past kBrickRows *
// nxr stand in for the last node of row kBrickRows - 1, so that lanes 0 and 63 hold the
// extreme x).  The collapsed columns of every (row, octave) go into the wave's slice, row
// rho's regions at rho * total, and each lane reads its M rows with one x part.  The
// lattice and the column folds and reads are those of the row kernels, so brick values
// equal the row kernels' at the same nodes bit for bit (while neither falls back to the
// per-lane functions).  Simplex layers, and layers whose regions do not fit the slice,
// run per lane.
constexpr int kBrickRows = 7;
constexpr int kBrickGroups = 3;  // M: groups of rows per wave

// Quotient a / b of small non-negative integers (a < 4096, 1 <= b <= 512; here a task or
// lane index and a region size) with the divisor given as rb = 1 / b in float32:
// (a + 1/2) / b lies at least 1 / (2 b) >= 2^-10 away from an integer, and the roundings
// of rb and of the product move it by less than 4096 * 2^-23 = 2^-11, so truncation
// gives the exact quotient (an integer division by a variable is a long correction
// sequence).
__device__ __forceinline__ int small_div(int a, float rb) { return (int)(((float)a + 0.5f) * rb); }

// One gradient / value layer at this lane's M nodes (x, y[m], z[m]) of rows
// rho = kBrickRows m + rl; yz(rho, y, z) gives row rho's coordinates.  The (row, octave)
// table has two float4 per entry (ty, tz, uy, uz; the combo seeds), entry rho * oct + k.
template <int NOISE, int M, class YZ>
__device__ __forceinline__ void brick_layer_m(const RowCtx& c, const DLayer& L, int rl, YZ yz, float x,
                                              const float* y, const float* z, float* out) {
  constexpr int NR = kBrickRows * M;
  const int oct = L.fractal == UPF_FRACTAL_NONE ? 1 : L.octaves;
  float2* slice = reinterpret_cast<float2*>(c.slice);
  int lo[UPF_MAX_OCTAVES], base[UPF_MAX_OCTAVES];
  int total = 0;
#pragma unroll
  for (int k = 0; k < UPF_MAX_OCTAVES; k++) {
    if (k < oct) {
      int hi;
      lane_range((int)floorf(x * L.freq[k]), lo[k], hi);
      base[k] = total;
      total += hi - lo[k] + 2;
    }
  }
  if (NR * total > kRowSlice2) {
#pragma unroll
    for (int m = 0; m < M; m++) out[m] = layer_value(L, x, y[m], z[m]);
    return;
  }
  const float roct = 1.0f / (float)oct;
  const int lane = lane_of(c);
  for (int e = lane; e < NR * oct; e += 64) {
    const int rho = small_div(e, roct);
    const int k = e - rho * oct;
    float f = L.freq[0];
#pragma unroll
    for (int j = 1; j < UPF_MAX_OCTAVES; j++)
      if (j < oct && k == j) f = L.freq[j];
    float yr, zr;
    yz(rho, yr, zr);
    const float qy = yr * f, qz = zr * f;
    const float fy = floorf(qy), fz = floorf(qz);
    const float ty = qy - fy, tz = qz - fz;
    const uint32_t y0 = (uint32_t)(int)fy * kPY, z0 = (uint32_t)(int)fz * kPZ;
    const uint32_t sk = L.seed + (uint32_t)k;
    c.oct[2 * e] = make_float4(ty, tz, fade(ty), fade(tz));
    c.oct[2 * e + 1] = make_float4(__uint_as_float(sk ^ y0 ^ z0), __uint_as_float(sk ^ (y0 + kPY) ^ z0),
                                   __uint_as_float(sk ^ y0 ^ (z0 + kPZ)), __uint_as_float(sk ^ (y0 + kPY) ^ (z0 + kPZ)));
  }
  wave_sync();
  const float rtotal = 1.0f / (float)total;
  for (int rd = 0; rd * 64 < NR * total; rd++) {
    const int t = 64 * rd + lane;
    const int rho = min(small_div(t, rtotal), NR - 1);
    const int tt = t - rho * total;
    int k = 0, off = lo[0] - base[0];  // column X = tt + off of octave k
#pragma unroll
    for (int j = 1; j < UPF_MAX_OCTAVES; j++)
      if (j < oct && tt >= base[j]) { k = j; off = lo[j] - base[j]; }
    const int e = rho * oct + k;
    const float2 v = fold_column<NOISE>(c, c.oct[2 * e], c.oct[2 * e + 1], (uint32_t)(tt + off));
    if (t < NR * total) slice[t] = v;
  }
  wave_sync();
  float acc[M];
#pragma unroll
  for (int m = 0; m < M; m++) acc[m] = 0.0f;
#pragma unroll
  for (int k = 0; k < UPF_MAX_OCTAVES; k++) {
    if (k < oct) {
      float n[M];
      col_read<NOISE, M>(x * L.freq[k], lo[k], slice + base[k] + rl * total, kBrickRows * total, n);
#pragma unroll
      for (int m = 0; m < M; m++) acc[m] = fractal_acc(L, k, n[m], acc[m]);
    }
  }
  wave_sync();
#pragma unroll
  for (int m = 0; m < M; m++) out[m] = acc[m];
}

// One cellular layer at this lane's node of row rl of a group of kBrickRows rows
// (yz(r, y, z) for r < kBrickRows): the feature points of each row written once (63
// lanes), each lane reading its own row's; as row_layer_t.
template <class YZ>
__device__ __forceinline__ float brick_cell_layer(const RowCtx& c, const DLayer& L, int rl, YZ yz, float x, float y,
                                                  float z) {
  constexpr int R = kBrickRows;
  const int oct = L.fractal == UPF_FRACTAL_NONE ? 1 : L.octaves;
  float2* slice = reinterpret_cast<float2*>(c.slice);
  int lo[UPF_MAX_OCTAVES], base[UPF_MAX_OCTAVES];
  int total = 0;
  bool fits = true;
#pragma unroll
  for (int k = 0; k < UPF_MAX_OCTAVES; k++) {
    if (k < oct) {
      int hi;
      lane_range((int)rintf(x * L.freq[k]), lo[k], hi);
      base[k] = total;
      total += 72;
      fits = fits && hi - lo[k] + 3 <= 7;
    }
  }
  if (!fits || R * total > kRowSlice2) return layer_value(L, x, y, z);
  for (int r = 0; r < R; r++) {
    float yr, zr;
    yz(r, yr, zr);
#pragma unroll
    for (int k = 0; k < UPF_MAX_OCTAVES; k++) {
      if (k < oct) {
        const float f = L.freq[k];
        cell_write(c, L, L.seed + (uint32_t)k, yr * f, zr * f, lo[k], slice + r * total + base[k]);
      }
    }
  }
  wave_sync();
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < UPF_MAX_OCTAVES; k++) {
    if (k < oct) {
      const float q = x * L.freq[k];
      acc = fractal_acc(L, k, cell_read(q, slice + rl * total + base[k] + ((int)rintf(q) - lo[k])), acc);
    }
  }
  wave_sync();
  return acc;
}

// f at this lane's M nodes (field_eval<false> with the layers from the brick rows).
template <int M, bool COMPOSED = true, class YZ>
__device__ __forceinline__ void brick_field(const RowCtx& c, const DSpec& s, int rl, YZ yz, float x, const float* y,
                                            const float* z, float* out) {
  const bool need_elev = s.compose == UPF_COMPOSE_LAND || s.compose == UPF_COMPOSE_NAVIGATION ||
                         s.compose == UPF_COMPOSE_ELEVATION_RAW;
  float elev[M];
#pragma unroll
  for (int m = 0; m < M; m++) elev[m] = 0.0f;
  if (need_elev) {
    float acc[M], msk[M];
#pragma unroll
    for (int m = 0; m < M; m++) acc[m] = msk[m] = 0.0f;
    for (int l = 0; l < s.n_layers; l++) {
      const DLayer& L = s.L[l];
      float raw[M];
      if (L.noise == UPF_NOISE_GRADIENT) {
        brick_layer_m<UPF_NOISE_GRADIENT, M>(c, L, rl, yz, x, y, z, raw);
      } else if (L.noise == UPF_NOISE_VALUE) {
        brick_layer_m<UPF_NOISE_VALUE, M>(c, L, rl, yz, x, y, z, raw);
      } else if (L.noise == UPF_NOISE_CELLULAR) {
#pragma unroll
        for (int m = 0; m < M; m++) {
          auto yz_m = [&](int r, float& yy, float& zz) { yz(kBrickRows * m + r, yy, zz); };
          raw[m] = brick_cell_layer(c, L, rl, yz_m, x, y[m], z[m]);
        }
      } else {
#pragma unroll
        for (int m = 0; m < M; m++) raw[m] = layer_value(L, x, y[m], z[m]);
      }
#pragma unroll
      for (int m = 0; m < M; m++) {
        const float v = L.weight * raw[m];
        if (L.blend == UPF_BLEND_ADD) acc[m] = acc[m] + v;
        else if (L.blend == UPF_BLEND_MASK) msk[m] = __builtin_amdgcn_fmed3f(v + 0.5f, 0.0f, 1.0f);
        else acc[m] = fmaf(msk[m], v - acc[m], acc[m]);
      }
    }
#pragma unroll
    for (int m = 0; m < M; m++) elev[m] = acc[m];
  }
#pragma unroll
  for (int m = 0; m < M; m++) {
    const float shape = s.compose == UPF_COMPOSE_ELEVATION_RAW ? 0.0f : shape_eval<false, COMPOSED>(s, x, y[m], z[m]);
    switch (s.compose) {
      case UPF_COMPOSE_LAND: out[m] = shape - s.escale * elev[m]; break;
      case UPF_COMPOSE_WATER: out[m] = shape; break;
      case UPF_COMPOSE_NAVIGATION: out[m] = shape - s.escale * fmaxf(elev[m], 0.0f); break;
      case UPF_COMPOSE_ELEVATION: out[m] = shape / s.escale; break;
      default: out[m] = elev[m]; break;
    }
  }
}

}  // namespace upf
