// upf_bake.hip -- UV atlas, UV rasterization and the per-texel material bake.
//
// Replaces meshGenerateTexture(mesh, {w, h, generator = Generator::pixel}) and the
// pixel callback (sources/textures.cpp:24-43, 65-71).  The per-texel material is the
// reduced, deterministic form of coloringDefault documented at upf_bake_spec
// (sources/coloringModes.cpp:905-930).
//
// The bake is HBM-bound per texel (triangle id 4 B + output 20 B, plus the mesh rows it
// touches, which stay in cache since neighbouring texels share triangles) once the
// material is light; the detail noise adds about 60 flops per octave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "upf_device.hpp"
#include "upf_internal.hpp"

namespace upf {

struct DBake {
  DSpec field;
  DLayer detail;
  float bands[3], band_width;
  float colors[5][3];
  float cliff[2];
  float detail_albedo;
  float roughness[4];
  float bump[2];
};

// Edge function without fused operations, so every rounding matches the CPU restatement.
__device__ __forceinline__ float edge_fn(float ax, float ay, float bx, float by, float px, float py) {
  const float u = (bx - ax) * (py - ay);
  const float v = (by - ay) * (px - ax);
  return u - v;
}

__global__ __launch_bounds__(256) void k_atlas(int64_t n, int side, int cols, float inv_w, float inv_h,
                                               float* __restrict__ uv) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const int64_t c = t >> 1;
  const float x0 = (float)((int)(c % cols) * side), y0 = (float)((int)(c / cols) * side);
  const float s = (float)side;
  float px[3], py[3];
  if ((t & 1) == 0) {  // lower-left half: (0,0) (s,0) (0,s)
    px[0] = x0; py[0] = y0; px[1] = x0 + s; py[1] = y0; px[2] = x0; py[2] = y0 + s;
  } else {             // upper-right half: (s,s) (0,s) (s,0)
    px[0] = x0 + s; py[0] = y0 + s; px[1] = x0; py[1] = y0 + s; px[2] = x0 + s; py[2] = y0;
  }
  for (int v = 0; v < 3; v++) {
    uv[6 * t + 2 * v] = px[v] * inv_w;
    uv[6 * t + 2 * v + 1] = py[v] * inv_h;
  }
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Texel-space set-up of triangle t: its corners, the sign that orients its edge
// functions, its box and the box's texels clipped to the image.  False when it covers
// nothing (zero or non-finite area, a box outside the image).  TIGHT: the box's texels
// are exactly those whose centres lie in the box (a texel range of at most 2^24 per
// side); otherwise the range is padded by up to one texel per side.  Either way
// rast_covers tests the box, so the covered set is the same.
struct RastTri {
  float ax, ay, bx, by, cx, cy, sgn, lx, hx, ly, hy;
  int x0, x1, y0, y1;
};

// Smallest x in [0, n] whose centre x + 0.5f is >= v (n when there is none), n <= 2^24.
// floorf(v - 0.5f) is within one texel of it (v - 0.5f rounds by at most half a unit
// where the centres are integers plus a half, by at most one above 2^23); two checked
// steps each way reach it without a data-dependent loop.
__device__ __forceinline__ int texel_first(float v, int n) {
  int x = min((int)fminf(fmaxf(floorf(v - 0.5f), 0.0f), 16777216.0f), n);
#pragma unroll
  for (int step = 0; step < 2; step++)
    if (x > 0 && (float)(x - 1) + 0.5f >= v) x--;
#pragma unroll
  for (int step = 0; step < 2; step++)
    if (x < n && (float)x + 0.5f < v) x++;
  return x;
}

// Largest x in [-1, n - 1] whose centre x + 0.5f is <= v (-1 when there is none).
__device__ __forceinline__ int texel_last(float v, int n) {
  int x = min((int)fminf(fmaxf(floorf(v - 0.5f), -1.0f), 16777216.0f), n - 1);
#pragma unroll
  for (int step = 0; step < 2; step++)
    if (x < n - 1 && (float)(x + 1) + 0.5f <= v) x++;
#pragma unroll
  for (int step = 0; step < 2; step++)
    if (x >= 0 && (float)x + 0.5f > v) x--;
  return x;
}

template <bool TIGHT>
__device__ __forceinline__ bool rast_setup(const float* __restrict__ uv, int64_t t, int w, int h, RastTri& r) {
  r.ax = uv[6 * t] * (float)w; r.ay = uv[6 * t + 1] * (float)h;
  r.bx = uv[6 * t + 2] * (float)w; r.by = uv[6 * t + 3] * (float)h;
  r.cx = uv[6 * t + 4] * (float)w; r.cy = uv[6 * t + 5] * (float)h;
  const float area = edge_fn(r.ax, r.ay, r.bx, r.by, r.cx, r.cy);
  if (!(area != 0.0f) || !isfinite(area)) return false;
  r.sgn = area > 0.0f ? 1.0f : -1.0f;
  r.lx = fminf(r.ax, fminf(r.bx, r.cx)); r.hx = fmaxf(r.ax, fmaxf(r.bx, r.cx));
  r.ly = fminf(r.ay, fminf(r.by, r.cy)); r.hy = fmaxf(r.ay, fmaxf(r.by, r.cy));
  if (TIGHT) {
    r.x0 = texel_first(r.lx, w); r.x1 = texel_last(r.hx, w);
    r.y0 = texel_first(r.ly, h); r.y1 = texel_last(r.hy, h);
  } else {
    r.x0 = max(0, (int)floorf(r.lx - 0.5f)); r.x1 = min(w - 1, (int)ceilf(r.hx - 0.5f));
    r.y0 = max(0, (int)floorf(r.ly - 0.5f)); r.y1 = min(h - 1, (int)ceilf(r.hy - 0.5f));
  }
  return r.x0 <= r.x1 && r.y0 <= r.y1;
}

// Whether the centre of texel (x, y) is inside (or on) the triangle; texel centres
// outside the triangle's box are outside, whatever the edge functions round to.
__device__ __forceinline__ bool rast_covers(const RastTri& r, int x, int y) {
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  if (px < r.lx || px > r.hx || py < r.ly || py > r.hy) return false;
  const float w0 = r.sgn * edge_fn(r.bx, r.by, r.cx, r.cy, px, py);
  const float w1 = r.sgn * edge_fn(r.cx, r.cy, r.ax, r.ay, px, py);
  const float w2 = r.sgn * edge_fn(r.ax, r.ay, r.bx, r.by, px, py);
  return w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f;
}

// One thread per triangle: every texel centre of its bounding box inside (or on) the UV
// triangle gets max(current, t).
__global__ __launch_bounds__(256) void k_rasterize(const float* __restrict__ uv, int64_t n, int w, int h,
                                                   int32_t* __restrict__ tri) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  RastTri r;
  if (!rast_setup<false>(uv, t, w, h, r)) return;
  for (int y = r.y0; y <= r.y1; y++)
    for (int x = r.x0; x <= r.x1; x++)
      if (rast_covers(r, x, y)) atomicMax(tri + (int64_t)y * w + x, (int32_t)t);
}

// ------------------------------------------------------------------ tile-binned rasterization
// The image is cut into tiles of 64 x 16 texels (a wave stores one 256-byte line).  The
// bin pass gives each tile that a triangle's texel box meets (at most kRastSlots of them)
// a slot in that tile's bin of kRastBin entries: consecutive triangles of an atlas share
// tiles, so the lanes of a wave with equal tiles in a row take their slots with one
// atomic, all runs of the wave at once.  The tile pass then resolves the largest
// covering id of each of a tile's 1024 texels in LDS, binned triangles spread over the
// workgroup's threads, and stores every texel once, -1 where nothing covers it: the
// image is written in whole lines, with no global atomics and no separate fill.  What
// does not fit, a (triangle, tile) pair beyond a full bin or a triangle over more than
// kRastSlots tiles, is listed as t * 4 + slot and added after the tile pass with global
// atomicMax over that tile (over the whole box for a triangle over more tiles), as
// k_rasterize does.
constexpr int kRastTW = 64, kRastTH = 16;
constexpr int kRastBin = 1024;
constexpr int kRastSlots = 4;
constexpr int kRastMaxSide = 1 << 24;

struct RastTiles {
  int tx0, ty0, cols, n;
};

__device__ __forceinline__ RastTiles rast_tiles(const RastTri& r) {
  RastTiles q;
  q.tx0 = r.x0 / kRastTW; q.ty0 = r.y0 / kRastTH;
  q.cols = r.x1 / kRastTW - q.tx0 + 1;
  q.n = q.cols * (r.y1 / kRastTH - q.ty0 + 1);
  return q;
}

__device__ __forceinline__ int rast_tile_key(const RastTiles& q, int tiles_x, int j) {
  return (q.ty0 + j / q.cols) * tiles_x + q.tx0 + j % q.cols;
}

__global__ __launch_bounds__(256) void k_rast_bin(const float* __restrict__ uv, int64_t n, int w, int h, int tiles_x,
                                                  uint32_t* __restrict__ count, uint32_t* __restrict__ bins,
                                                  uint32_t* __restrict__ over, uint32_t* __restrict__ nover) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63;
  RastTri r;
  const bool ok = t < n && rast_setup<true>(uv, t, w, h, r);
  RastTiles q = {0, 0, 1, 0};
  if (ok) q = rast_tiles(r);
  const bool small = q.n <= kRastSlots;
  if (ok && !small) over[atomicAdd(nover, 1u)] = (uint32_t)t * kRastSlots;
  const uint64_t below = (1ull << lane) - 1ull;  // lanes under this one
#pragma unroll
  for (int j = 0; j < kRastSlots; j++) {
    const int key = (small && j < q.n) ? rast_tile_key(q, tiles_x, j) : -1;
    if (__ballot(key >= 0) == 0) break;  // no lane has a j-th tile (uniform)
    // runs of equal keys among consecutive lanes: the first lane of a run takes the
    // run's slots, the others their offsets in it
    const int prev = __shfl_up(key, 1);
    const bool head = lane == 0 || key != prev;
    const uint64_t heads = __ballot(head);
    const uint64_t after = heads & ~(below | (1ull << lane));  // heads above this lane
    const int next = after ? __ffsll((unsigned long long)after) - 1 : 64;
    uint32_t base = 0;
    if (head && key >= 0) base = atomicAdd(count + key, (uint32_t)(next - lane));
    const int my_head = 63 - __clzll((unsigned long long)(heads & (below | (1ull << lane))));
    const uint32_t pos = (uint32_t)__shfl((int)base, my_head) + (uint32_t)(lane - my_head);
    if (key >= 0) {
      if (pos < (uint32_t)kRastBin)
        bins[(int64_t)key * kRastBin + pos] = (uint32_t)t;
      else
        over[atomicAdd(nover, 1u)] = (uint32_t)t * kRastSlots + (uint32_t)j;
    }
  }
}

__global__ __launch_bounds__(256) void k_rast_tiles(const float* __restrict__ uv, int w, int h, int tiles_x,
                                                    const uint32_t* __restrict__ count,
                                                    const uint32_t* __restrict__ bins, int32_t* __restrict__ tri) {
  constexpr int kTexels = kRastTW * kRastTH;
  __shared__ int s_best[kTexels];
  const int tile = blockIdx.x;
  const int bx0 = (tile % tiles_x) * kRastTW, by0 = (tile / tiles_x) * kRastTH;
  const uint32_t c = min(count[tile], (uint32_t)kRastBin);
  uint32_t ids[kRastBin / 256];
#pragma unroll
  for (int e = 0; e < kRastBin / 256; e++) {  // the bin's loads in flight together
    const uint32_t slot = threadIdx.x + 256 * e;
    ids[e] = slot < c ? bins[(int64_t)tile * kRastBin + slot] : 0xFFFFFFFFu;
  }
#pragma unroll
  for (int e = 0; e < kTexels / 256; e++) s_best[threadIdx.x + 256 * e] = -1;
  __syncthreads();
#pragma unroll
  for (int e = 0; e < kRastBin / 256; e++) {
    if (ids[e] == 0xFFFFFFFFu) continue;
    const int64_t t = ids[e];
    RastTri r;
    if (!rast_setup<true>(uv, t, w, h, r)) continue;  // true for every binned triangle
    const int x0 = max(r.x0, bx0), x1 = min(r.x1, bx0 + kRastTW - 1);
    const int y0 = max(r.y0, by0), y1 = min(r.y1, by0 + kRastTH - 1);
    for (int y = y0; y <= y1; y++)
      for (int x = x0; x <= x1; x++)
        if (rast_covers(r, x, y)) atomicMax(&s_best[(y - by0) * kRastTW + (x - bx0)], (int)t);
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < kTexels / 256; e++) {
    const int u = threadIdx.x + 256 * e;
    const int x = bx0 + (u & (kRastTW - 1)), y = by0 + u / kRastTW;
    if (x < w && y < h) tri[(int64_t)y * w + x] = s_best[u];
  }
}

// The listed (triangle, tile) pairs and triangles over more tiles, after the tile pass.
__global__ __launch_bounds__(256) void k_rast_over(const float* __restrict__ uv, int w, int h, int tiles_x,
                                                   const uint32_t* __restrict__ over,
                                                   const uint32_t* __restrict__ nover, int32_t* __restrict__ tri) {
  const uint32_t no = *nover;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < no; i += gridDim.x * 256) {
    const int64_t t = over[i] / kRastSlots;
    const int j = (int)(over[i] % kRastSlots);
    RastTri r;
    if (!rast_setup<true>(uv, t, w, h, r)) continue;
    const RastTiles q = rast_tiles(r);
    int x0 = r.x0, x1 = r.x1, y0 = r.y0, y1 = r.y1;
    if (q.n <= kRastSlots) {  // one tile of a binned triangle
      const int key = rast_tile_key(q, tiles_x, j);
      const int bx0 = (key % tiles_x) * kRastTW, by0 = (key / tiles_x) * kRastTH;
      x0 = max(x0, bx0); x1 = min(x1, bx0 + kRastTW - 1);
      y0 = max(y0, by0); y1 = min(y1, by0 + kRastTH - 1);
    }
    for (int y = y0; y <= y1; y++)
      for (int x = x0; x <= x1; x++)
        if (rast_covers(r, x, y)) atomicMax(tri + (int64_t)y * w + x, (int32_t)t);
  }
}

struct RastLayout {
  int tiles_x, tiles_y;
  int64_t tiles;
  size_t off_count, off_bins, off_over, bytes;
};

// count: tiles + 1 words (the last counts the list), bins: kRastBin words per tile,
// list: kRastSlots words per triangle (its upper bound).
RastLayout rast_layout(int64_t n, int w, int h) {
  RastLayout L;
  L.tiles_x = (w + kRastTW - 1) / kRastTW;
  L.tiles_y = (h + kRastTH - 1) / kRastTH;
  L.tiles = (int64_t)L.tiles_x * L.tiles_y;
  const auto a256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t o = 0;
  L.off_count = o; o = a256(o + 4 * (size_t)(L.tiles + 1));
  L.off_bins = o; o = a256(o + 4 * (size_t)kRastBin * (size_t)L.tiles);
  L.off_over = o; o = a256(o + 4 * (size_t)kRastSlots * (size_t)n);
  L.bytes = o;
  return L;
}

__device__ __forceinline__ float sat(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float smoothstep_(float e0, float e1, float x) {
  const float t = sat((x - e0) / (e1 - e0));
  return t * t * (3.0f - 2.0f * t);
}
__device__ __forceinline__ float mix_(float a, float b, float t) { return a + t * (b - a); }

template <bool EXT>
__global__ __launch_bounds__(256) void k_bake(const DBake b, const float* __restrict__ P, const float* __restrict__ N,
                                              int64_t nv, const uint32_t* __restrict__ I, int64_t nt,
                                              const float* __restrict__ uv, const int32_t* __restrict__ tri, int w,
                                              int h, float* __restrict__ albedo, float* __restrict__ bump,
                                              float* __restrict__ rough) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= w || y >= h) return;
  const int64_t o = (int64_t)y * w + x;
  const int32_t t = tri[o];
  uint32_t i0 = 0, i1 = 0, i2 = 0;
  const bool ok = t >= 0 && t < nt;
  if (ok) {
    i0 = I[3 * (int64_t)t];
    i1 = I[3 * (int64_t)t + 1];
    i2 = I[3 * (int64_t)t + 2];
  }
  if (!ok || i0 >= nv || i1 >= nv || i2 >= nv) {
    const float nan = __builtin_nanf("");
    albedo[3 * o] = albedo[3 * o + 1] = albedo[3 * o + 2] = nan;
    bump[o] = nan;
    rough[o] = nan;
    return;
  }
  // barycentrics of the texel centre in the UV triangle
  const float ax = uv[6 * (int64_t)t] * (float)w, ay = uv[6 * (int64_t)t + 1] * (float)h;
  const float bx = uv[6 * (int64_t)t + 2] * (float)w, by = uv[6 * (int64_t)t + 3] * (float)h;
  const float cx = uv[6 * (int64_t)t + 4] * (float)w, cy = uv[6 * (int64_t)t + 5] * (float)h;
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  const float area = edge_fn(ax, ay, bx, by, cx, cy);
  const float w0 = edge_fn(bx, by, cx, cy, px, py) / area;
  const float w1 = edge_fn(cx, cy, ax, ay, px, py) / area;
  const float w2 = edge_fn(ax, ay, bx, by, px, py) / area;
  float p[3], n[3];
  for (int a = 0; a < 3; a++) {
    p[a] = w0 * P[3 * (int64_t)i0 + a] + w1 * P[3 * (int64_t)i1 + a] + w2 * P[3 * (int64_t)i2 + a];
    n[a] = w0 * N[3 * (int64_t)i0 + a] + w1 * N[3 * (int64_t)i1 + a] + w2 * N[3 * (int64_t)i2 + a];
  }
  const float nl = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  const float pl = sqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  float cosang = 1.0f;
  if (nl > 0.0f && pl > 0.0f) cosang = (n[0] * p[0] + n[1] * p[1] + n[2] * p[2]) / (nl * pl);
  const float slope = 1.0f - cosang;
  // the shape and the detail noise round like the CPU restatement (X = true): the bands'
  // smoothsteps would amplify a few ulps of elevation by up to 150x
  const float elev = shape_eval<EXT, true, true>(b.field, p[0], p[1], p[2]) / b.field.escale;
  const float d = b.detail.weight * layer_value<true>(b.detail, p[0], p[1], p[2]);
  const float s1 = smoothstep_(b.bands[0] - b.band_width, b.bands[0] + b.band_width, elev);
  const float s2 = smoothstep_(b.bands[1] - b.band_width, b.bands[1] + b.band_width, elev);
  const float s3 = smoothstep_(b.bands[2] - b.band_width, b.bands[2] + b.band_width, elev);
  const float cl = smoothstep_(b.cliff[0], b.cliff[1], slope);
  const float m = 1.0f + b.detail_albedo * d;
  for (int a = 0; a < 3; a++) {
    float c = mix_(b.colors[0][a], b.colors[1][a], s1);
    c = mix_(c, b.colors[2][a], s2);
    c = mix_(c, b.colors[3][a], s3);
    c = mix_(c, b.colors[4][a], cl);
    albedo[3 * o + a] = sat(c * m);
  }
  rough[o] = sat(b.roughness[0] + b.roughness[1] * slope + b.roughness[2] * d + b.roughness[3] * s3);
  bump[o] = sat(0.5f + b.bump[0] * d + b.bump[1] * elev);
}

// Up to kDilR dilation rounds in one launch (layout [h][w][C], validity by channel 0).
// A workgroup loads its kDilT x kDilT tile with a kDilR-texel halo into LDS and runs the
// rounds there: in round k a texel whose first channel is still NaN takes, per channel,
// the mean of its 8 neighbours valid before round k (summed in the order dy = -1..1,
// dx = -1..1), as one pass over the whole image would.  Valid texels never change and a
// texel is filled at most once, so one copy of the values and a tag per texel suffice:
// tag 0 = valid from the start, k = filled in round k (so not yet visible to round k's
// other texels), kTagNone = still NaN, kTagOut = outside the image.  Round k computes
// the texels at least k from the loaded border, whose neighbours are right after round
// k - 1; the tile itself lies kDilR in.  One read of the tile and its halo and one write
// of the tile, instead of a read and a write of the whole image per round.
constexpr int kDilT = 32;                   // output tile side
constexpr int kDilR = 7;                    // rounds per launch (the halo)
constexpr int kDilS = kDilT + 2 * kDilR;    // loaded side
constexpr uint8_t kTagOut = 254, kTagNone = 255;
constexpr int kDilThreads = 512;  // 8 waves share a tile's rounds

// only_nan: write only the texels that were NaN when loaded (the last launch of an
// in-place dilation; k_dilate_merge then copies them back).
template <int C>
__global__ __launch_bounds__(kDilThreads) void k_dilate(const float* __restrict__ in, float* __restrict__ out, int w, int h,
                                                int rounds, bool only_nan) {
  __shared__ float s_v[C][kDilS * kDilS];
  __shared__ uint8_t s_tag[kDilS * kDilS];
  __shared__ uint16_t s_nan[kDilS * kDilS];  // the loaded texels that are NaN (only they can change)
  __shared__ int s_nnan;
  const int x0 = (int)blockIdx.x * kDilT - kDilR, y0 = (int)blockIdx.y * kDilT - kDilR;
  if (threadIdx.x == 0) s_nnan = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (int t0 = 0; t0 < kDilS * kDilS; t0 += kDilThreads) {  // uniform trip count: whole waves ballot
    const int t = t0 + (int)threadIdx.x;
    bool nan_here = false;
    if (t < kDilS * kDilS) {
      const int ly = t / kDilS, lx = t - ly * kDilS;
      const int x = x0 + lx, y = y0 + ly;
      uint8_t tag = kTagOut;
      if (x >= 0 && x < w && y >= 0 && y < h) {
        const int64_t o = ((int64_t)y * w + x) * C;
#pragma unroll
        for (int ch = 0; ch < C; ch++) s_v[ch][t] = in[o + ch];
        nan_here = isnan(s_v[0][t]);
        tag = nan_here ? kTagNone : 0;
      }
      s_tag[t] = tag;
    }
    const uint64_t m = __ballot(nan_here);
    int base = 0;
    if (lane == 0 && m) base = atomicAdd(&s_nnan, __popcll(m));
    base = __shfl(base, 0, 64);
    if (nan_here) s_nan[base + __popcll(m & ((1ull << lane) - 1))] = (uint16_t)t;
  }
  __syncthreads();
  const int nnan = s_nnan;
  for (int k = 1; k <= rounds; k++) {
    for (int e = threadIdx.x; e < nnan; e += kDilThreads) {
      const int idx = s_nan[e];
      const int ly = idx / kDilS, lx = idx - ly * kDilS;  // constant divisor
      if (lx < k || lx >= kDilS - k || ly < k || ly >= kDilS - k || s_tag[idx] != kTagNone) continue;
      float sum[C];
#pragma unroll
      for (int ch = 0; ch < C; ch++) sum[ch] = 0.0f;
      int cnt = 0;
#pragma unroll
      for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
          if (dx == 0 && dy == 0) continue;
          const int q = idx + dy * kDilS + dx;
          if (s_tag[q] >= k) continue;  // NaN, outside, or filled in this round
#pragma unroll
          for (int ch = 0; ch < C; ch++) sum[ch] = sum[ch] + s_v[ch][q];
          cnt++;
        }
      if (cnt) {
#pragma unroll
        for (int ch = 0; ch < C; ch++) s_v[ch][idx] = sum[ch] / (float)cnt;
        s_tag[idx] = (uint8_t)k;
      }
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < kDilT * kDilT; t += kDilThreads) {
    const int ly = kDilR + t / kDilT, lx = kDilR + t % kDilT;
    const int x = x0 + lx, y = y0 + ly;
    if (x < w && y < h && !(only_nan && s_tag[ly * kDilS + lx] == 0)) {
      const int64_t o = ((int64_t)y * w + x) * C;
#pragma unroll
      for (int ch = 0; ch < C; ch++) out[o + ch] = s_v[ch][ly * kDilS + lx];
    }
  }
}

// img takes the values k_dilate wrote into filled (only_nan) where its first channel is NaN.
// A thread takes kMergeTexels texels, 256 apart, and reads their first channels before
// using any, for more bytes in flight: 4096^2 images with 30% NaN merge in 39 / 47 / 62
// us for 1 / 2 / 3 channels, against 42 / 52 / 68 with one texel per thread.
constexpr int kMergeTexels = 4;

template <int C>
__global__ __launch_bounds__(256) void k_dilate_merge(float* __restrict__ img, const float* __restrict__ filled,
                                                      int64_t texels) {
  const int64_t t0 = (int64_t)blockIdx.x * (256 * kMergeTexels) + threadIdx.x;
  bool nan[kMergeTexels];
#pragma unroll
  for (int j = 0; j < kMergeTexels; j++) {
    const int64_t t = t0 + 256 * j;
    nan[j] = t < texels && isnan(img[t * C]);
  }
#pragma unroll
  for (int j = 0; j < kMergeTexels; j++) {
    const int64_t t = t0 + 256 * j;
    if (nan[j]) {
#pragma unroll
      for (int ch = 0; ch < C; ch++) img[t * C + ch] = filled[t * C + ch];
    }
  }
}

__global__ __launch_bounds__(256) void k_copy_f32(const float* __restrict__ in, float* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = in[i];
}

__device__ __forceinline__ uint8_t unit_to_u8(float n) { return (uint8_t)((n * 0.5f + 0.5f) * 255.0f + 0.5f); }

// Sobel normal of one texel from its 3 x 3 heights (rules at upf_image_height_to_normal),
// shared by both kernels below so that they round alike.
__device__ __forceinline__ void sobel_normal(uint32_t a8, uint32_t b8, uint32_t c8, uint32_t d8, uint32_t f8,
                                             uint32_t g8, uint32_t h8, uint32_t i8, float strength, uint8_t* o) {
  const float k = 1.0f / 255.0f;
  const float a = (float)a8 * k, b = (float)b8 * k, c = (float)c8 * k;
  const float d = (float)d8 * k, f = (float)f8 * k;
  const float g = (float)g8 * k, hh = (float)h8 * k, i = (float)i8 * k;
  const float gx = (c + 2.0f * f + i) - (a + 2.0f * d + g);
  const float gy = (g + 2.0f * hh + i) - (a + 2.0f * b + c);
  const float vx = -strength * gx, vy = -strength * gy;
  const float inv = 1.0f / sqrtf(vx * vx + vy * vy + 1.0f);
  o[0] = unit_to_u8(vx * inv);
  o[1] = unit_to_u8(vy * inv);
  o[2] = unit_to_u8(inv);
}

// One texel per lane over a 64 x 4 tile, texels [x_first, w) of each row.  Neighbour
// bytes come through L1/L2: 1 B read and 3 B written per texel from HBM.
__global__ __launch_bounds__(256) void k_height_to_normal(const uint8_t* __restrict__ in, int w, int h, float strength,
                                                          uint8_t* __restrict__ out, int x_first) {
  const int x = x_first + blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= w || y >= h) return;
  const int xm = x > 0 ? x - 1 : 0, xp = x + 1 < w ? x + 1 : w - 1;
  const int ym = y > 0 ? y - 1 : 0, yp = y + 1 < h ? y + 1 : h - 1;
  const uint8_t* rm = in + (int64_t)ym * w;
  const uint8_t* r0 = in + (int64_t)y * w;
  const uint8_t* rp = in + (int64_t)yp * w;
  uint8_t o[3];
  sobel_normal(rm[xm], rm[x], rm[xp], r0[xm], r0[xp], rp[xm], rp[x], rp[xp], strength, o);
  uint8_t* q = out + ((int64_t)y * w + x) * 3;
  q[0] = o[0];
  q[1] = o[1];
  q[2] = o[2];
}

// Four texels per lane (x = 4X .. 4X + 3, all inside the row): a 4-byte read per row plus
// the two bytes beside it, and the 12 output bytes as three 4-byte stores (the rows and
// the output 4-byte aligned).  With one byte store per channel the 4096^2 map took 45 us.
__global__ __launch_bounds__(256) void k_height_to_normal_x4(const uint8_t* __restrict__ in, int w, int h, float strength,
                                                             uint32_t* __restrict__ out) {
  const int X = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (X >= (w >> 2) || y >= h) return;
  const int x0 = 4 * X;
  const int ym = y > 0 ? y - 1 : 0, yp = y + 1 < h ? y + 1 : h - 1;
  const int xl = x0 > 0 ? x0 - 1 : 0, xr = x0 + 4 < w ? x0 + 4 : w - 1;
  uint32_t row[3][6];  // bytes x0 - 1 .. x0 + 4 of rows ym, y, yp (clamped)
  const int rows[3] = {ym, y, yp};
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const uint8_t* p = in + (int64_t)rows[r] * w;
    const uint32_t v = *reinterpret_cast<const uint32_t*>(p + x0);
    row[r][0] = p[xl];
#pragma unroll
    for (int j = 0; j < 4; j++) row[r][1 + j] = (v >> (8 * j)) & 0xFFu;
    row[r][5] = p[xr];
  }
  uint8_t o[12];
#pragma unroll
  for (int j = 0; j < 4; j++)
    sobel_normal(row[0][j], row[0][j + 1], row[0][j + 2], row[1][j], row[1][j + 2], row[2][j], row[2][j + 1],
                 row[2][j + 2], strength, o + 3 * j);
  uint32_t* q = out + ((int64_t)y * w + x0) * 3 / 4;
#pragma unroll
  for (int k = 0; k < 3; k++)
    q[k] = (uint32_t)o[4 * k] | ((uint32_t)o[4 * k + 1] << 8) | ((uint32_t)o[4 * k + 2] << 16) |
           ((uint32_t)o[4 * k + 3] << 24);
}

// (roughness, metallic) -> (255, roughness, metallic), texels [first, n), one per lane.
__global__ __launch_bounds__(256) void k_special_to_pbr(const uint8_t* __restrict__ in, int64_t first, int64_t n,
                                                        uint8_t* __restrict__ out) {
  const int64_t i = first + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  out[3 * i] = 255;
  out[3 * i + 1] = in[2 * i];
  out[3 * i + 2] = in[2 * i + 1];
}

// Four texels per lane: one 8-byte read and three 4-byte stores (in 8-byte and out
// 4-byte aligned; n4 groups of four).
__global__ __launch_bounds__(256) void k_special_to_pbr_x4(const uint2* __restrict__ in, int64_t n4,
                                                           uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n4) return;
  const uint2 v = in[i];  // r0 m0 r1 m1 | r2 m2 r3 m3
  const uint32_t r0 = v.x & 0xFFu, m0 = (v.x >> 8) & 0xFFu, r1 = (v.x >> 16) & 0xFFu, m1 = v.x >> 24;
  const uint32_t r2 = v.y & 0xFFu, m2 = (v.y >> 8) & 0xFFu, r3 = (v.y >> 16) & 0xFFu, m3 = v.y >> 24;
  // bytes: 255 r0 m0 255 | r1 m1 255 r2 | m2 255 r3 m3
  out[3 * i] = 0xFFu | (r0 << 8) | (m0 << 16) | (0xFFu << 24);
  out[3 * i + 1] = r1 | (m1 << 8) | (0xFFu << 16) | (r2 << 24);
  out[3 * i + 2] = m2 | (0xFFu << 8) | (r3 << 16) | (m3 << 24);
}

__device__ __forceinline__ uint32_t to_u8(float v) {
  return isnan(v) ? 0u : (uint32_t)(fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f + 0.5f);
}

__global__ __launch_bounds__(256) void k_to_u8(const float* __restrict__ in, int64_t n, uint8_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  out[i] = (uint8_t)to_u8(in[i]);
}

// Four values per thread: one 16-byte read and one 4-byte write (in 16-byte and out
// 4-byte aligned; n4 groups of four).
__global__ __launch_bounds__(256) void k_to_u8_x4(const float4* __restrict__ in, int64_t n4,
                                                  uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n4) return;
  const float4 v = in[i];
  out[i] = to_u8(v.x) | (to_u8(v.y) << 8) | (to_u8(v.z) << 16) | (to_u8(v.w) << 24);
}

int atlas_side(int64_t n, int w, int h) {
  const int64_t cells = (n + 1) / 2;
  int best = 0;
  for (int s = 2; s <= (w < h ? w : h); s++)
    if ((int64_t)(w / s) * (h / s) >= cells) best = s;
  return best;
}

}  // namespace upf

using namespace upf;

extern "C" {

int upf_atlas_per_triangle(int64_t n, int32_t w, int32_t h, float* d_uv, void* stream) {
  if (n < 0 || w < 2 || h < 2) return UPF_ERR_BAD_GRID;
  if (n == 0) return UPF_OK;
  if (!d_uv) return UPF_ERR_NULL;
  const int side = atlas_side(n, w, h);
  if (side < 2) return UPF_ERR_CAPACITY;
  hipLaunchKernelGGL(k_atlas, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), n, side, w / side,
                     1.0f / (float)w, 1.0f / (float)h, d_uv);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

int upf_bake_rasterize(const float* d_uv, int64_t n, int32_t w, int32_t h, int32_t* d_tri, void* stream) {
  if (n < 0 || w < 1 || h < 1) return UPF_ERR_BAD_GRID;
  if (!d_tri || (n > 0 && !d_uv)) return UPF_ERR_NULL;
  if (n >= ((int64_t)1 << 31)) return UPF_ERR_BAD_GRID;
  const int64_t texels = (int64_t)w * h;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((texels + 255) / 256)), dim3(256), 0, s, d_tri, texels, -1);
  if (n > 0) hipLaunchKernelGGL(k_rasterize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_uv, n, w, h, d_tri);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

size_t upf_bake_rasterize_workspace_bytes(int64_t n, int32_t w, int32_t h) {
  if (n < 0 || n >= ((int64_t)1 << 30) || w < 1 || h < 1 || w > kRastMaxSide || h > kRastMaxSide) return 0;
  const RastLayout L = rast_layout(n, w, h);
  return L.tiles < ((int64_t)1 << 31) ? L.bytes : 0;  // one workgroup per tile
}

int upf_bake_rasterize_ws(const float* d_uv, int64_t n, int32_t w, int32_t h, int32_t* d_tri, void* d_ws,
                          size_t ws_bytes, void* stream) {
  if (n < 0 || w < 1 || h < 1) return UPF_ERR_BAD_GRID;
  if (!d_tri || (n > 0 && !d_uv)) return UPF_ERR_NULL;
  // list entries are t * 4 + slot in 32 bits; tight boxes need sides of at most 2^24
  if (n >= ((int64_t)1 << 30) || w > kRastMaxSide || h > kRastMaxSide) return UPF_ERR_BAD_GRID;
  hipStream_t s = as_stream(stream);
  const RastLayout L = rast_layout(n, w, h);
  if (!d_ws || ws_bytes < L.bytes) return UPF_ERR_WORKSPACE;
  if (L.tiles >= ((int64_t)1 << 31)) return UPF_ERR_BAD_GRID;
  char* ws = static_cast<char*>(d_ws);
  uint32_t* count = reinterpret_cast<uint32_t*>(ws + L.off_count);
  uint32_t* nover = count + L.tiles;
  uint32_t* bins = reinterpret_cast<uint32_t*>(ws + L.off_bins);
  uint32_t* over = reinterpret_cast<uint32_t*>(ws + L.off_over);
  // the counts and the list length start at zero (a fill kernel: no memset nodes)
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((L.tiles + 1 + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<int32_t*>(count), L.tiles + 1, 0);
  if (n > 0)
    hipLaunchKernelGGL(k_rast_bin, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_uv, n, w, h, L.tiles_x,
                       count, bins, over, nover);
  hipLaunchKernelGGL(k_rast_tiles, dim3((unsigned)L.tiles), dim3(256), 0, s, d_uv, w, h, L.tiles_x, count, bins,
                     d_tri);
  if (n > 0)
    hipLaunchKernelGGL(k_rast_over, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1024)), dim3(256), 0, s, d_uv,
                       w, h, L.tiles_x, over, nover, d_tri);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

int upf_bake(const upf_bake_spec* spec, const float* d_P, const float* d_N, int64_t nv, const uint32_t* d_I,
             int64_t n, const float* d_uv, const int32_t* d_tri, int32_t w, int32_t h, float* d_albedo,
             float* d_bump, float* d_rough, void* stream) {
  if (!spec || !d_tri || !d_albedo || !d_bump || !d_rough) return UPF_ERR_NULL;
  if (n > 0 && (!d_P || !d_N || !d_I || !d_uv)) return UPF_ERR_NULL;
  if (w < 1 || h < 1 || n < 0 || nv < 0) return UPF_ERR_BAD_GRID;
  DBake b;
  int st = to_device(&spec->field, &b.field);
  if (st) return st;
  if (spec->field.elevation_scale == 0.0f) return UPF_ERR_INVALID_VALUE;
  upf_field_spec tmp = spec->field;  // validate the detail layer as a one-layer spec
  tmp.n_layers = 1;
  tmp.layers[0] = spec->detail;
  DSpec dd;
  if ((st = to_device(&tmp, &dd))) return st;
  b.detail = dd.L[0];
  const float* fl[] = {spec->bands, &spec->band_width, &spec->colors[0][0], spec->cliff, &spec->detail_albedo,
                       spec->roughness, spec->bump};
  const int nf[] = {3, 1, 15, 2, 1, 4, 2};
  for (int i = 0; i < 7; i++)
    for (int j = 0; j < nf[i]; j++)
      if (!std::isfinite(fl[i][j])) return UPF_ERR_INVALID_VALUE;
  for (int i = 0; i < 3; i++) b.bands[i] = spec->bands[i];
  b.band_width = spec->band_width;
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 3; j++) b.colors[i][j] = spec->colors[i][j];
  b.cliff[0] = spec->cliff[0];
  b.cliff[1] = spec->cliff[1];
  b.detail_albedo = spec->detail_albedo;
  for (int i = 0; i < 4; i++) b.roughness[i] = spec->roughness[i];
  b.bump[0] = spec->bump[0];
  b.bump[1] = spec->bump[1];
  const dim3 grid((w + 63) / 64, (h + 3) / 4);
  if (spec_ext(b.field))
    hipLaunchKernelGGL(k_bake<true>, grid, dim3(256), 0, as_stream(stream), b, d_P, d_N, nv, d_I, n, d_uv, d_tri, w, h,
                       d_albedo, d_bump, d_rough);
  else
    hipLaunchKernelGGL(k_bake<false>, grid, dim3(256), 0, as_stream(stream), b, d_P, d_N, nv, d_I, n, d_uv, d_tri, w,
                       h, d_albedo, d_bump, d_rough);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

int upf_image_dilate(float* d_img, float* d_scratch, int32_t w, int32_t h, int32_t channels, int32_t rounds,
                     void* stream) {
  if (w < 1 || h < 1 || rounds < 0) return UPF_ERR_BAD_GRID;
  if (channels < 1 || channels > 4) return UPF_ERR_INVALID_VALUE;
  if (rounds == 0) return UPF_OK;
  if (!d_img || !d_scratch) return UPF_ERR_NULL;
  hipStream_t s = as_stream(stream);
  const int64_t n = (int64_t)w * h * channels;
  // launches of up to kDilR rounds each, alternating between d_img and d_scratch; the
  // last one reads d_img and writes only the texels that were NaN into d_scratch, which
  // k_dilate_merge copies back.  An odd number of launches before it starts from a copy
  // in d_scratch, so that they end in d_img.
  const int launches = (rounds + kDilR - 1) / kDilR;
  float* a = d_img;
  float* b = d_scratch;
  if ((launches - 1) & 1) {
    hipLaunchKernelGGL(k_copy_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_img, d_scratch, n);
    a = d_scratch;
    b = d_img;
  }
  const dim3 grid((unsigned)((w + kDilT - 1) / kDilT), (unsigned)((h + kDilT - 1) / kDilT));
  const int64_t texels = (int64_t)w * h;
  const dim3 mgrid((unsigned)((texels + 256 * kMergeTexels - 1) / (256 * kMergeTexels)));
  for (int left = rounds; left > 0; left -= kDilR) {
    const int r = left < kDilR ? left : kDilR;
    const bool last = left <= kDilR;  // then a == d_img
    switch (channels) {
      case 1: hipLaunchKernelGGL(k_dilate<1>, grid, dim3(kDilThreads), 0, s, a, b, w, h, r, last); break;
      case 2: hipLaunchKernelGGL(k_dilate<2>, grid, dim3(kDilThreads), 0, s, a, b, w, h, r, last); break;
      case 3: hipLaunchKernelGGL(k_dilate<3>, grid, dim3(kDilThreads), 0, s, a, b, w, h, r, last); break;
      default: hipLaunchKernelGGL(k_dilate<4>, grid, dim3(kDilThreads), 0, s, a, b, w, h, r, last); break;
    }
    if (last) {
      switch (channels) {
        case 1: hipLaunchKernelGGL(k_dilate_merge<1>, mgrid, dim3(256), 0, s, a, b, texels); break;
        case 2: hipLaunchKernelGGL(k_dilate_merge<2>, mgrid, dim3(256), 0, s, a, b, texels); break;
        case 3: hipLaunchKernelGGL(k_dilate_merge<3>, mgrid, dim3(256), 0, s, a, b, texels); break;
        default: hipLaunchKernelGGL(k_dilate_merge<4>, mgrid, dim3(256), 0, s, a, b, texels); break;
      }
    }
    float* t = a;
    a = b;
    b = t;
  }
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

int upf_image_to_u8(const float* d_img, int64_t n, uint8_t* d_out, void* stream) {
  if (n < 0) return UPF_ERR_BAD_GRID;
  if (n == 0) return UPF_OK;
  if (!d_img || !d_out) return UPF_ERR_NULL;
  hipStream_t s = as_stream(stream);
  int64_t head = 0;  // values done four at a time
  if (reinterpret_cast<uintptr_t>(d_img) % 16 == 0 && reinterpret_cast<uintptr_t>(d_out) % 4 == 0 && n >= 4) {
    const int64_t n4 = n / 4;
    hipLaunchKernelGGL(k_to_u8_x4, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(d_img), n4, reinterpret_cast<uint32_t*>(d_out));
    head = 4 * n4;
  }
  if (head < n)
    hipLaunchKernelGGL(k_to_u8, dim3((unsigned)((n - head + 255) / 256)), dim3(256), 0, s, d_img + head, n - head,
                       d_out + head);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

int upf_image_height_to_normal(const uint8_t* d_height, int32_t w, int32_t h, float strength, uint8_t* d_out,
                               void* stream) {
  if (w < 1 || h < 1) return UPF_ERR_BAD_GRID;
  if (!(strength == strength) || isinf(strength)) return UPF_ERR_INVALID_VALUE;
  if (!d_height || !d_out) return UPF_ERR_NULL;
  hipStream_t s = as_stream(stream);
  // four texels per lane where the rows and the output allow 4-byte accesses (w a
  // multiple of 4, both buffers 4-byte aligned), one per lane otherwise
  int x_first = 0;
  if (w >= 4 && w % 4 == 0 && reinterpret_cast<uintptr_t>(d_height) % 4 == 0 &&
      reinterpret_cast<uintptr_t>(d_out) % 4 == 0) {
    hipLaunchKernelGGL(k_height_to_normal_x4, dim3((w / 4 + 63) / 64, (h + 3) / 4), dim3(256), 0, s, d_height, w, h,
                       strength, reinterpret_cast<uint32_t*>(d_out));
    x_first = w;
  }
  if (x_first < w)
    hipLaunchKernelGGL(k_height_to_normal, dim3((w - x_first + 63) / 64, (h + 3) / 4), dim3(256), 0, s, d_height, w,
                       h, strength, d_out, x_first);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

int upf_image_special_to_gltf_pbr(const uint8_t* d_special, int64_t n, uint8_t* d_out, void* stream) {
  if (n < 0) return UPF_ERR_BAD_GRID;
  if (n == 0) return UPF_OK;
  if (!d_special || !d_out) return UPF_ERR_NULL;
  hipStream_t s = as_stream(stream);
  int64_t first = 0;
  if (reinterpret_cast<uintptr_t>(d_special) % 8 == 0 && reinterpret_cast<uintptr_t>(d_out) % 4 == 0) {
    const int64_t n4 = n / 4;
    if (n4)
      hipLaunchKernelGGL(k_special_to_pbr_x4, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, s,
                         reinterpret_cast<const uint2*>(d_special), n4, reinterpret_cast<uint32_t*>(d_out));
    first = 4 * n4;
  }
  if (first < n)
    hipLaunchKernelGGL(k_special_to_pbr, dim3((unsigned)((n - first + 255) / 256)), dim3(256), 0, s, d_special, first,
                       n, d_out);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

}  // extern "C"
