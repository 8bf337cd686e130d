// upf_mc_device.hpp -- wave-level marching-cubes bodies shared by the single-grid
// kernels (upf_mc.hip) and the batched kernels (upf_batch.hip).  Included inside
// namespace upf.
#pragma once

#define UPF_MC_QUAL __constant__
#include "mc_table.inc"
#undef UPF_MC_QUAL

// edge -> (owner corner offset x, y, z, axis)
__constant__ int8_t c_edge_owner[12][4] = {
    {0, 0, 0, 0}, {1, 0, 0, 1}, {0, 1, 0, 0}, {0, 0, 0, 1}, {0, 0, 1, 0}, {1, 0, 1, 1},
    {0, 1, 1, 0}, {0, 0, 1, 1}, {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}};

// The same tables as compile-time data for the emit: the triangulation with rows padded
// to 16 bytes (one 16-byte load per cell), and the edge owners packed 5 bits per edge
// (dx | dy << 1 | dz << 2 | axis << 3) into one 64-bit immediate.
namespace mc_ct {
#define UPF_MC_QUAL static constexpr
#include "mc_table.inc"
#undef UPF_MC_QUAL
constexpr int8_t owner[12][4] = {{0, 0, 0, 0}, {1, 0, 0, 1}, {0, 1, 0, 0}, {0, 0, 0, 1}, {0, 0, 1, 0}, {1, 0, 1, 1},
                                 {0, 1, 1, 0}, {0, 0, 1, 1}, {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}};
struct Tris16 {
  int8_t v[256][16];
};
constexpr Tris16 pad_tris() {
  Tris16 t{};
  for (int c = 0; c < 256; c++) {
    for (int q = 0; q < 15; q++) t.v[c][q] = upf_mc_tris[c][q];
    t.v[c][15] = -1;
  }
  return t;
}
constexpr uint64_t pack_owner() {
  uint64_t r = 0;
  for (int e = 0; e < 12; e++)
    r |= (uint64_t)(owner[e][0] | (owner[e][1] << 1) | (owner[e][2] << 2) | (owner[e][3] << 3)) << (5 * e);
  return r;
}
}  // namespace mc_ct
__constant__ __align__(16) mc_ct::Tris16 c_tris16 = mc_ct::pad_tris();
constexpr uint64_t kEdgeOwner = mc_ct::pack_owner();

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// packed per-group counts: vertices (bits 0-7, <= 192), triangles (8-16, <= 320), active (17)
__device__ __forceinline__ uint32_t pack_counts(int nv, int nt) {
  return (uint32_t)nv | ((uint32_t)nt << 8) | ((nv | nt) ? (1u << 17) : 0u);
}

// ------------------------------------------------------------------ classify
// One wave: group column xb of row j, planes k0 <= k < k1.  The current z plane stays in
// registers; each step uses plane k+1 (loaded the step before) and loads plane k+2's rows
// j and j+1 at x and x+1 (clamped, so every load is unconditional and the loads of a step
// issue back to back).
__device__ __forceinline__ void mc_classify_column(const float* __restrict__ f, const McDims& d, float iso, int xb,
                                                   int j, int k0, int k1, const uint8_t* s_ntris,
                                                   uint64_t* __restrict__ S, uint32_t* __restrict__ counts) {
  const int lane = threadIdx.x & 63;
  const int x = xb * 64 + lane;
  const bool valid = x < d.nx;
  const int xc = min(x, d.nx - 1), x1 = min(x + 1, d.nx - 1);
  const int j1 = min(j + 1, d.ny - 1);
  const bool hx = valid && x + 1 < d.nx, hy = valid && j + 1 < d.ny;
  const int64_t r0 = (int64_t)j * d.nx, r1 = (int64_t)j1 * d.nx;
  // current plane: a0 = (x, j), a1 = (x+1, j), a2 = (x+1, j+1), a3 = (x, j+1)
  const float* p = f + (int64_t)k0 * d.nxy;
  float a0 = p[r0 + xc], a1 = p[r0 + x1], a2 = p[r1 + x1], a3 = p[r1 + xc];
  // plane k+1 is loaded one step ahead, so two planes of loads are in flight per step
  const float* q1 = f + (int64_t)min(k0 + 1, d.nz - 1) * d.nxy;
  float c0 = q1[r0 + xc], c1 = q1[r0 + x1], c2 = q1[r1 + x1], c3 = q1[r1 + xc];
  for (int k = k0; k < k1; k++) {
    const bool hz = valid && k + 1 < d.nz;
    const float b0 = c0, b1 = c1, b2 = c2, b3 = c3;
    const float* q = f + (int64_t)min(k + 2, d.nz - 1) * d.nxy;
    c0 = q[r0 + xc]; c1 = q[r0 + x1]; c2 = q[r1 + x1]; c3 = q[r1 + xc];
    const bool s0 = valid && a0 < iso;
    const bool cx = hx && ((a1 < iso) != s0);
    const bool cy = hy && ((a3 < iso) != s0);
    const bool cz = hz && ((b0 < iso) != s0);
    int nt = 0;
    if (hx && hy && hz) {
      const int code = (a0 < iso) | ((a1 < iso) << 1) | ((a2 < iso) << 2) | ((a3 < iso) << 3) | ((b0 < iso) << 4) |
                       ((b1 < iso) << 5) | ((b2 < iso) << 6) | ((b3 < iso) << 7);
      nt = s_ntris[code];
    }
    const uint64_t sm = ballot64(s0);
    const int nv = __popcll(ballot64(cx)) + __popcll(ballot64(cy)) + __popcll(ballot64(cz));
    // triangle total from bit-sliced ballots (nt <= 5)
    const int ntot = __popcll(ballot64(nt & 1)) + 2 * __popcll(ballot64(nt & 2)) + 4 * __popcll(ballot64(nt & 4));
    if (lane == 0) {
      const int64_t g = ((int64_t)k * d.ny + j) * d.gpr + xb;
      S[g] = sm;
      counts[g] = pack_counts(nv, ntot);
    }
    a0 = b0; a1 = b1; a2 = b2; a3 = b3;
  }
}

// ------------------------------------------------------------------ emit
struct Cross {
  uint64_t bx, by, bz;
  uint32_t voff, pad;
};

// Crossing masks of group (xb, j, k) (row-aligned, so its +y and +z neighbour groups
// share its lane alignment).  Lanes past nx hold 0 sign bits and are masked.
__device__ __forceinline__ Cross group_crossings(const uint64_t* __restrict__ S, const uint32_t* __restrict__ voff,
                                                 const McDims& d, int xb, int j, int k) {
  const int64_t gg = ((int64_t)k * d.ny + j) * d.gpr + xb;
  const int first = xb * 64;
  const int nval = min(64, d.nx - first);     // lanes inside the row
  const int nvx = min(64, d.nx - 1 - first);  // lanes with an x neighbour
  const uint64_t vmask = nval >= 64 ? ~0ull : ((1ull << nval) - 1);
  const uint64_t xmask = nvx >= 64 ? ~0ull : (nvx > 0 ? ((1ull << nvx) - 1) : 0ull);
  const uint64_t s = S[gg];
  const uint64_t snext = (xb + 1 < d.gpr) ? S[gg + 1] : 0ull;
  Cross c;
  c.bx = (s ^ ((s >> 1) | (snext << 63))) & xmask;
  c.by = (j + 1 < d.ny) ? ((s ^ S[gg + d.gpr]) & vmask) : 0ull;
  c.bz = (k + 1 < d.nz) ? ((s ^ S[gg + (int64_t)d.gpr * d.ny]) & vmask) : 0ull;
  c.voff = voff[gg];
  c.pad = 0;
  return c;
}

// One axis of the central-difference gradient at a node (one-sided on the grid faces),
// given f one step up (fp), at the node (f0) and one step down (fm).
__device__ __forceinline__ float grad_axis(int idx, int nmax, float fp, float f0, float fm, float sp) {
  if (idx == 0) return (fp - f0) / sp;
  if (idx == nmax - 1) return (f0 - fm) / sp;
  return (fp - fm) / (2.0f * sp);
}

// One wave emits active group gid (local to its grid): vertices at voff (global ids,
// stored at pos/nrm[id]) and triangles from toff (global), with index values made local
// by subtracting vbase.  Writes at or past cap_v / cap_t are dropped (the caller sized
// the buffers from the counts, or checks the counts afterwards).  s_nb: 8 LDS entries
// private to this wave.  Every field value a lane needs -- its cell's 8 corners and the
// central-difference neighbours of its node and of its three edges' far ends -- is
// loaded up front with clamped (in-bounds) indices, in the same round trip as the
// counts and sign masks.
__device__ __forceinline__ void mc_emit_group(const float* __restrict__ f, const McDims& d, const DGrid& g, float iso,
                                              const uint64_t* __restrict__ S, const uint32_t* __restrict__ counts,
                                              const uint32_t* __restrict__ voff, uint32_t gid, uint32_t toff,
                                              uint32_t vbase, Cross* s_nb, float* __restrict__ pos,
                                              float* __restrict__ nrm, uint32_t* __restrict__ idx, int64_t cap_v,
                                              int64_t cap_t) {
  const int lane = threadIdx.x & 63;
  const uint32_t row = gid / (uint32_t)d.gpr;
  const int xb = (int)(gid - row * (uint32_t)d.gpr);
  const int k = (int)(row / (uint32_t)d.ny);
  const int j = (int)(row - (uint32_t)k * (uint32_t)d.ny);
  const int i = xb * 64 + lane;
  const bool valid = i < d.nx;

  // field values (clamped coordinates; a clamped value is never used), loaded only by
  // the lanes that own a vertex
  const int ic = min(i, d.nx - 1);
  const int64_t sy = d.nx, sz = d.nxy;
  const int xs[4] = {max(ic - 1, 0), ic, min(ic + 1, d.nx - 1), min(ic + 2, d.nx - 1)};  // i-1 .. i+2
  const int ys[4] = {max(j - 1, 0), j, min(j + 1, d.ny - 1), min(j + 2, d.ny - 1)};
  const int zs[4] = {max(k - 1, 0), k, min(k + 1, d.nz - 1), min(k + 2, d.nz - 1)};
  auto at = [&](int xi, int yi, int zi) { return f[(int64_t)zs[zi] * sz + (int64_t)ys[yi] * sy + xs[xi]]; };
  const uint32_t c = counts[gid];
  const int cnt_v = c & 0xFF, cnt_t = (c >> 8) & 0x1FF;

  // crossing masks of the (up to) 8 groups a cell of this group takes edges from:
  // slot = gx | dy << 1 | dz << 2, gx = 1 for the next group along x
  if (lane < 8) {
    const int gx = lane & 1, dy = (lane >> 1) & 1, dz = lane >> 2;
    if (xb + gx < d.gpr && j + dy < d.ny && k + dz < d.nz) s_nb[lane] = group_crossings(S, voff, d, xb + gx, j + dy, k + dz);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  const uint64_t lt = (1ull << lane) - 1;

  // ---- vertices
  if (cnt_v) {
    const Cross cc = s_nb[0];
    uint32_t vid = cc.voff + (uint32_t)(__popcll(cc.bx & lt) + __popcll(cc.by & lt) + __popcll(cc.bz & lt));
    const uint64_t bits[3] = {cc.bx, cc.by, cc.bz};
    const bool any = ((cc.bx | cc.by | cc.bz) >> lane) & 1ull;
    if (any) {
      const float c0 = at(1, 1, 1), c1 = at(2, 1, 1), c3 = at(1, 2, 1), c4 = at(1, 1, 2);
      float c2 = 0.f, c5 = 0.f, c7 = 0.f;
      float xm = 0.f, ym = 0.f, zm = 0.f, x2 = 0.f, xym = 0.f, xzm = 0.f, y2 = 0.f, yxm = 0.f, yzm = 0.f;
      float z2 = 0.f, zxm = 0.f, zym = 0.f;
      if (nrm) {
        c2 = at(2, 2, 1); c5 = at(2, 1, 2); c7 = at(1, 2, 2);
        xm = at(0, 1, 1); ym = at(1, 0, 1); zm = at(1, 1, 0);
        x2 = at(3, 1, 1); xym = at(2, 0, 1); xzm = at(2, 1, 0);
        y2 = at(1, 3, 1); yxm = at(0, 2, 1); yzm = at(1, 2, 0);
        z2 = at(1, 1, 3); zxm = at(0, 1, 2); zym = at(1, 0, 2);
      }
      const float f0 = c0;
      const float p0[3] = {node_coord(g.o[0], g.s[0], i), node_coord(g.o[1], g.s[1], j),
                           node_coord(g.o[2], g.s[2], k)};
      const float f1s[3] = {c1, c3, c4};
      const int ijk[3] = {i, j, k};
      float g0[3] = {0.0f, 0.0f, 0.0f};
      if (nrm) {
        g0[0] = grad_axis(i, d.nx, c1, c0, xm, g.s[0]);
        g0[1] = grad_axis(j, d.ny, c3, c0, ym, g.s[1]);
        g0[2] = grad_axis(k, d.nz, c4, c0, zm, g.s[2]);
      }
#pragma unroll
      for (int ax = 0; ax < 3; ax++) {
        if (!((bits[ax] >> lane) & 1ull)) continue;
        const float f1 = f1s[ax];
        const float t = (iso - f0) / (f1 - f0);
        const float p1a = node_coord(g.o[ax], g.s[ax], ijk[ax] + 1);
        float p[3] = {p0[0], p0[1], p0[2]};
        p[ax] = p0[ax] + t * (p1a - p0[ax]);
        if ((int64_t)vid < cap_v) {
          pos[3 * (size_t)vid + 0] = p[0];
          pos[3 * (size_t)vid + 1] = p[1];
          pos[3 * (size_t)vid + 2] = p[2];
          if (nrm) {
            // gradient at the edge's far end
            float g1[3];
            if (ax == 0) {
              g1[0] = grad_axis(i + 1, d.nx, x2, c1, c0, g.s[0]);
              g1[1] = grad_axis(j, d.ny, c2, c1, xym, g.s[1]);
              g1[2] = grad_axis(k, d.nz, c5, c1, xzm, g.s[2]);
            } else if (ax == 1) {
              g1[0] = grad_axis(i, d.nx, c2, c3, yxm, g.s[0]);
              g1[1] = grad_axis(j + 1, d.ny, y2, c3, c0, g.s[1]);
              g1[2] = grad_axis(k, d.nz, c7, c3, yzm, g.s[2]);
            } else {
              g1[0] = grad_axis(i, d.nx, c5, c4, zxm, g.s[0]);
              g1[1] = grad_axis(j, d.ny, c7, c4, zym, g.s[1]);
              g1[2] = grad_axis(k + 1, d.nz, z2, c4, c0, g.s[2]);
            }
            const float gx = g0[0] + t * (g1[0] - g0[0]);
            const float gy = g0[1] + t * (g1[1] - g0[1]);
            const float gz = g0[2] + t * (g1[2] - g0[2]);
            const float len = __builtin_sqrtf(gx * gx + gy * gy + gz * gz);
            const float inv = len > 0.0f ? 1.0f / len : 0.0f;
            nrm[3 * (size_t)vid + 0] = gx * inv;
            nrm[3 * (size_t)vid + 1] = gy * inv;
            nrm[3 * (size_t)vid + 2] = gz * inv;
          }
        }
        vid++;
      }
    }
  }

  // ---- triangles
  if (cnt_t) {
    const bool cell = valid && i + 1 < d.nx && j + 1 < d.ny && k + 1 < d.nz;
    // the case code from the sign words (bit = f < iso) of the cell's 4 corner rows and
    // their +x neighbour groups: wave-uniform 64-bit loads instead of 8 values per lane
    const bool nx1 = xb + 1 < d.gpr, hy = j + 1 < d.ny, hz = k + 1 < d.nz;
    const int64_t gr = (int64_t)gid, plane = (int64_t)d.ny * d.gpr;
    const uint64_t wa = S[gr], wax = nx1 ? S[gr + 1] : 0ull;
    const uint64_t way = hy ? S[gr + d.gpr] : 0ull, waxy = (hy && nx1) ? S[gr + d.gpr + 1] : 0ull;
    const uint64_t wb = hz ? S[gr + plane] : 0ull, wbx = (hz && nx1) ? S[gr + plane + 1] : 0ull;
    const uint64_t wby = (hz && hy) ? S[gr + plane + d.gpr] : 0ull;
    const uint64_t wbxy = (hz && hy && nx1) ? S[gr + plane + d.gpr + 1] : 0ull;
    const uint64_t wa1 = (wa >> 1) | (wax << 63), way1 = (way >> 1) | (waxy << 63);
    const uint64_t wb1 = (wb >> 1) | (wbx << 63), wby1 = (wby >> 1) | (wbxy << 63);
    int code = 0;
    if (cell)
      code = (int)((wa >> lane) & 1ull) | ((int)((wa1 >> lane) & 1ull) << 1) | ((int)((way1 >> lane) & 1ull) << 2) |
             ((int)((way >> lane) & 1ull) << 3) | ((int)((wb >> lane) & 1ull) << 4) |
             ((int)((wb1 >> lane) & 1ull) << 5) | ((int)((wby1 >> lane) & 1ull) << 6) |
             ((int)((wby >> lane) & 1ull) << 7);
    const int nt = upf_mc_ntris[code];
    int inc = nt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(inc, o, 64);
      if (lane >= o) inc += v;
    }
    const uint32_t tid0 = toff + (uint32_t)(inc - nt);
    // the cell's triangulation row in one 16-byte load; edge owners from the immediate
    const int4 row = *reinterpret_cast<const int4*>(&c_tris16.v[code][0]);
    const uint32_t rw[4] = {(uint32_t)row.x, (uint32_t)row.y, (uint32_t)row.z, (uint32_t)row.w};
#pragma unroll
    for (int t = 0; t < UPF_MC_MAX_TRIS; t++) {
      if (t >= nt) break;
      const uint32_t tid = tid0 + (uint32_t)t;
#pragma unroll
      for (int v = 0; v < 3; v++) {
        const int q = 3 * t + v;
        const int e = (int)(int8_t)(uint8_t)(rw[q >> 2] >> (8 * (q & 3)));
        const int own = (int)((kEdgeOwner >> (5 * e)) & 31u);
        const int ol = lane + (own & 1);
        const int slot = (ol >> 6) | (((own >> 1) & 1) << 1) | (((own >> 2) & 1) << 2);
        const int axis = own >> 3;
        const Cross cc = s_nb[slot];
        const int l = ol & 63;
        const uint64_t m = (1ull << l) - 1;
        uint32_t id = cc.voff + (uint32_t)(__popcll(cc.bx & m) + __popcll(cc.by & m) + __popcll(cc.bz & m));
        if (axis >= 1) id += (uint32_t)((cc.bx >> l) & 1ull);
        if (axis == 2) id += (uint32_t)((cc.by >> l) & 1ull);
        if ((int64_t)tid < cap_t) idx[3 * (size_t)tid + v] = id - vbase;
      }
    }
  }
}
