// upf_mc_device.hpp -- wave-level marching-cubes bodies shared by the single-grid
// kernels (upf_mc.hip) and the batched kernels (upf_batch.hip).  Included inside
// namespace upf.
#pragma once

#define UPF_MC_QUAL __constant__
#include "mc_table.inc"
#undef UPF_MC_QUAL

// The same tables as compile-time data for the emits: the triangulation with rows padded
// to 16 bytes (one 16-byte load per cell), and the edge owners (the id table's order
// check, and a 64-bit immediate for the narrow band's emit).
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
// true when edge e is owned by corner (x, y, z) along axis a as the emit's id table
// assumes: x edges e = 0, 2, 4, 6 at corners 000, 010, 001, 011; y edges 1, 3, 5, 7 at
// 100, 000, 101, 001; z edges 8..11 at 000, 100, 110, 010
constexpr bool owner_order_ok() {
  constexpr int8_t want[12][4] = {{0, 0, 0, 0}, {1, 0, 0, 1}, {0, 1, 0, 0}, {0, 0, 0, 1}, {0, 0, 1, 0}, {1, 0, 1, 1},
                                  {0, 1, 1, 0}, {0, 0, 1, 1}, {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}};
  for (int e = 0; e < 12; e++)
    for (int q = 0; q < 4; q++)
      if (owner[e][q] != want[e][q]) return false;
  return true;
}
// the edge owners packed 5 bits per edge (dx | dy << 1 | dz << 2 | axis << 3)
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
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int src) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, src);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

// v plus the value of the lane that DPP control CTRL names, in the rows ROWS enables
// (lanes of other rows, and lanes whose source lies outside their row, add 0).
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_add_i(int v) {
  return v + __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, false);
}

// Inclusive sum of v over lanes 0..lane for the first LANES lanes (8 or 64; the other
// lanes get partial sums): within each 16-lane row by row shifts of 1, 2, 4 and 8, then
// rows 1 and 3 add lane 15 of the row before (row_bcast:15) and rows 2 and 3 add lane 31
// (row_bcast:31).  DPP adds in place of the LDS permutes of a __shfl_up scan.  Every
// lane of the wave must be active.
template <int LANES = 64>
__device__ __forceinline__ int wave_inclusive_sum(int v) {
  static_assert(LANES == 8 || LANES == 64, "8 or 64 lanes");
  v = dpp_add_i<0x111, 0xF>(v);
  v = dpp_add_i<0x112, 0xF>(v);
  v = dpp_add_i<0x114, 0xF>(v);
  if (LANES == 64) {
    v = dpp_add_i<0x118, 0xF>(v);
    v = dpp_add_i<0x142, 0xA>(v);
    v = dpp_add_i<0x143, 0xC>(v);
  }
  return v;
}

// packed per-group counts: vertices (bits 0-7, <= 192), triangles (8-16, <= 320), active (17)
__device__ __forceinline__ uint32_t pack_counts(int nv, int nt) {
  return (uint32_t)nv | ((uint32_t)nt << 8) | ((nv | nt) ? (1u << 17) : 0u);
}

// ------------------------------------------------------------------ scan helpers
// Three running sums at once: x = vertices, y = triangles, z = active groups.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint4 add4(uint4 a, uint4 b) { return make_uint4(a.x + b.x, a.y + b.y, a.z + b.z, 0); }
// Count words: kScanMc = marching-cubes groups (vertices bits 0-7, triangles 8-16,
// active bit 17); kScanBlocks = narrow-band blocks (vertices 0-11, triangles 12-23,
// active bit 24).
template <int FMT>
__device__ __forceinline__ uint4 unpack_counts(uint32_t c) {
  if (FMT == kScanMc) return make_uint4(c & 0xFFu, (c >> 8) & 0x1FFu, (c >> 17) & 1u, 0);
  return make_uint4(c & 0xFFFu, (c >> 12) & 0xFFFu, (c >> 24) & 1u, 0);
}

// Block-wide exclusive scan of one uint4 per thread; the block total goes to *total.
__device__ __forceinline__ uint4 block_exclusive_scan(uint4 v, uint4* total) {
  __shared__ uint4 wsum[kScanThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint4 inc = make_uint4((unsigned)wave_inclusive_sum((int)v.x), (unsigned)wave_inclusive_sum((int)v.y),
                               (unsigned)wave_inclusive_sum((int)v.z), 0u);
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint4 before = make_uint4(0, 0, 0, 0), all = make_uint4(0, 0, 0, 0);
  for (int w = 0; w < kScanThreads / 64; w++) {
    if (w < wave) before = add4(before, wsum[w]);
    all = add4(all, wsum[w]);
  }
  __syncthreads();
  *total = all;
  return make_uint4(before.x + inc.x - v.x, before.y + inc.y - v.y, before.z + inc.z - v.z, 0);
}

// ------------------------------------------------------------------ classify
// Packed counts of a group from the sign words alone (written by the field kernel or
// k_mc_signs), one lane per group.  Crossing masks come from S of the group, of its +x
// neighbour group and of the +y / +z rows (8 words, load_group_words), so vertex counts
// are a few 64-bit operations per lane.  Triangle counts need the codes of the mixed
// cells (corners not all on one side; the other codes have no triangles).  The lanes
// list their groups' mixed cells in LDS; in passes of at most 64 cells (a group's cells
// never split), lane t takes listed cell t, builds its code from the group's 8 words (in
// LDS) and adds the code's triangle count to the group's LDS counter.  A wave with ~6 active groups of ~7 mixed cells each needs
// one pass, instead of one wave-wide step per active group.  Lanes may belong to
// different grids (batched chunks): everything but the pass bookkeeping is per lane.

// groups per thread of the classification kernels (thread t of a workgroup takes groups
// t + 256 u), a whole number of scan tiles per workgroup so that they write the tile sums
constexpr int kClassifyGroups = 8;
static_assert((256 * kClassifyGroups) % kScanTile == 0 && kScanTile % 256 == 0, "whole scan tiles per workgroup");

struct SignPass {  // one wave's LDS
  uint64_t w[8][64];  // word q of lane (group) t at w[q][t]: consecutive lanes, no bank conflicts
  uint32_t nt[64];
  uint16_t item[64];
};

// The 8 sign words a group's classification reads (zero past the grid's faces), with
// the row flags; loaded apart from the classification so that a lane can have the next
// group's words in flight while it classifies the current one.
struct GroupWords {
  uint64_t a, ax, ay, axy, b, bx, by, bxy;
  uint32_t seg;  // the group's segment in its row | (a +y row) << 30 | (a +z row) << 31
};

__device__ __forceinline__ GroupWords load_group_words(const uint64_t* __restrict__ S, const McDims& d, int64_t g,
                                                       bool in) {
  GroupWords w;
  w.a = w.ax = w.ay = w.axy = w.b = w.bx = w.by = w.bxy = 0ull;
  w.seg = 0u;
  if (in) {
    const uint32_t row = (uint32_t)(g / d.gpr);
    const int xb = (int)(g - (int64_t)row * d.gpr);
    const int k = (int)(row / (uint32_t)d.ny), j = (int)(row - (uint32_t)k * (uint32_t)d.ny);
    const bool nx1 = xb + 1 < d.gpr, hy = j + 1 < d.ny, hz = k + 1 < d.nz;
    const int64_t plane = (int64_t)d.ny * d.gpr;
    w.seg = (uint32_t)xb | ((uint32_t)hy << 30) | ((uint32_t)hz << 31);
    w.a = S[g];
    w.ax = nx1 ? S[g + 1] : 0ull;
    w.ay = hy ? S[g + d.gpr] : 0ull;
    w.axy = (hy && nx1) ? S[g + d.gpr + 1] : 0ull;
    w.b = hz ? S[g + plane] : 0ull;
    w.bx = (hz && nx1) ? S[g + plane + 1] : 0ull;
    w.by = (hz && hy) ? S[g + plane + d.gpr] : 0ull;
    w.bxy = (hz && hy && nx1) ? S[g + plane + d.gpr + 1] : 0ull;
  }
  return w;
}

// W: the group's words (load_group_words), nx: its grid's row length
__device__ __forceinline__ uint32_t classify_group_words(const GroupWords& W, int nx, bool in,
                                                         const uint8_t* s_ntris, SignPass& P) {
  const int lane = threadIdx.x & 63;
  uint64_t a = 0, a1 = 0, ay = 0, ay1 = 0, b = 0, b1 = 0, by = 0, by1 = 0, cells = 0;
  int nv = 0;
  if (in) {
    const bool hy = (W.seg >> 30) & 1u, hz = W.seg >> 31;
    a = W.a; ay = W.ay; b = W.b; by = W.by;
    const int first = (int)(W.seg & 0x3FFFFFFFu) * 64;
    const int nval = min(64, nx - first), nvx = min(64, nx - 1 - first);
    const uint64_t vmask = nval >= 64 ? ~0ull : ((1ull << nval) - 1);
    const uint64_t xmask = nvx >= 64 ? ~0ull : (nvx > 0 ? ((1ull << nvx) - 1) : 0ull);
    // bit i of *1 = node i + 1 of the same row
    a1 = (a >> 1) | (W.ax << 63); ay1 = (ay >> 1) | (W.axy << 63);
    b1 = (b >> 1) | (W.bx << 63); by1 = (by >> 1) | (W.bxy << 63);
    nv = __popcll((a ^ a1) & xmask) + (hy ? __popcll((a ^ ay) & vmask) : 0) + (hz ? __popcll((a ^ b) & vmask) : 0);
    cells = (hy && hz) ? xmask : 0ull;
  }
  const uint64_t any = (a | a1 | ay | ay1 | b | b1 | by | by1) & cells;
  const uint64_t all = (a & a1 & ay & ay1 & b & b1 & by & by1) & cells;
  const uint64_t mixed = any & ~all;
  const int nmix = __popcll(mixed);
  // the words in code-bit order: bit q of cell c's code is bit c of word q
  P.w[0][lane] = a; P.w[1][lane] = a1; P.w[2][lane] = ay1; P.w[3][lane] = ay;
  P.w[4][lane] = b; P.w[5][lane] = b1; P.w[6][lane] = by1; P.w[7][lane] = by;
  P.nt[lane] = 0u;
  const int incl = wave_inclusive_sum(nmix);
  const int total = __builtin_amdgcn_readlane(incl, 63);
  for (int base = 0; base < total;) {
    // the groups whose cells lie within [base, base + 64); the first group left starts at
    // base and has at most 64 cells, so every pass takes one
    const bool take = nmix > 0 && incl - nmix >= base && incl <= base + 64;
    const uint64_t tk = ballot64(take);
    const int end = __builtin_amdgcn_readlane(incl, 63 - __clzll((long long)tk));
    // list the taken groups' mixed cells: one taken group at a time, each lane placing
    // the cell of its own bit position
    for (uint64_t tm = tk; tm; tm &= tm - 1) {
      const int src = __ffsll((unsigned long long)tm) - 1;
      const uint64_t mm = readlane64(mixed, src);
      const int p0 = __builtin_amdgcn_readlane(incl - nmix, src) - base;
      if ((mm >> lane) & 1ull) P.item[p0 + __popcll(mm & ((1ull << lane) - 1))] = (uint16_t)((src << 6) | lane);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < end - base) {
      const int e = P.item[lane];
      const int grp = e >> 6, c = e & 63;
      int code = 0;
#pragma unroll
      for (int q = 0; q < 8; q++) code |= (int)((P.w[q][grp] >> c) & 1ull) << q;
      atomicAdd(&P.nt[grp], (uint32_t)s_ntris[code]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    base = end;
  }
  return pack_counts(nv, (int)P.nt[lane]);
}

// ------------------------------------------------------------------ emit
struct Cross {
  uint64_t bx, by, bz;  // crossing edges along x, y, z (bit = owning node of the row segment)
  uint64_t w;           // the sign word itself (bit = f < iso)
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
  c.w = s;
  c.voff = voff[gg];
  c.pad = 0;
  return c;
}

// One axis of the central-difference gradient at a node (one-sided on the grid faces),
// given f one step up (fp), at the node (f0) and one step down (fm).
// One division with selected operands: written as three returns, the three divisions
// were all computed and selected per lane (the branches diverge), three times the
// division sequence for the same result.
__device__ __forceinline__ float grad_axis(int idx, int nmax, float fp, float f0, float fm, float sp) {
  const bool lo = idx == 0, hi = idx == nmax - 1;
  const float num = lo ? fp - f0 : (hi ? f0 - fm : fp - fm);
  const float den = (lo || hi) ? sp : 2.0f * sp;
  return num / den;
}

// Emission, one pass of a wave: up to kEmitPassGroups consecutive entries of the
// active-group list (act[a..], all of one grid: global group ids gbase + local id below
// glimit) whose cells with work -- a vertex on one of the node's own edges, or a case
// code other than 0 and 255 (exactly the codes with triangles) -- number at most 64;
// each lane takes one such cell.  A group holds 64 cells of which about a tenth cross
// the surface, so filling the lanes with cells that have work instead of the 64 cells
// of one group cuts the instructions per group several-fold.
//   1. lane 8 b + s loads the crossing masks and sign word of neighbour slot s of group
//      b (slot = gx | dy << 1 | dz << 2: the groups a cell takes edges from);
//   2. lane b counts its group's cells with work; the pass takes the leading groups
//      whose counts sum to at most 64 (at least one: a group has 64 cells);
//   3. the cells are listed in LDS in group order and, within a group, in cell order;
//   4. lane t takes listed cell t.
// Vertex ids (voff + the crossings of the lower nodes of the row segment, x < y < z per
// node), triangle ids (toff + the triangles of the lower cells of the group, table
// order) and every float operation are those of one group per wave, so the mesh does
// not depend on how the groups are split into passes.  Writes at or past cap_v / cap_t
// are dropped; index values are made local by subtracting vbase.  Every field value a
// lane needs -- its node's central-difference neighbours and those of its three edges'
// far ends -- is loaded with clamped (in-bounds) indices by the lanes that own a vertex.
// Returns the number of list entries consumed.
constexpr int kEmitPassGroups = 8;

struct EmitPass {  // one wave's LDS
  Cross nb[kEmitPassGroups][8];
  uint32_t toff[kEmitPassGroups];
  int xb[kEmitPassGroups], j[kEmitPassGroups], k[kEmitPassGroups], pre[kEmitPassGroups];
  uint16_t item[64];  // (group << 6) | cell
  uint32_t eid[12][64];  // lane t: the vertex id (less vbase) on edge e of listed cell t
  uint16_t vitem[192];   // the pass's vertices in id order: (listed cell << 2) | axis
};

__device__ __forceinline__ void emit_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Step 1 for lane 8 b + s: the crossing record of neighbour slot s of group b of a pass,
// from the group's active-list entry e (global id e.x, first triangle e.y).  Lanes past
// the pass's groups return an empty record.
struct EmitRec {
  Cross c;
  uint32_t toff;
  int xb, j, k;
};

__device__ __forceinline__ EmitRec emit_record(const uint64_t* __restrict__ S, const uint32_t* __restrict__ voff,
                                               const McDims& d, uint2 e, bool in, uint32_t gbase) {
  const int lane = threadIdx.x & 63;
  const int sl = lane & 7;
  EmitRec r;
  r.c.bx = r.c.by = r.c.bz = r.c.w = 0ull;
  r.c.voff = r.c.pad = 0u;
  r.toff = e.y;
  r.xb = r.j = r.k = 0;
  if (in) {
    const uint32_t gid = e.x - gbase;
    const uint32_t row = gid / (uint32_t)d.gpr;
    r.xb = (int)(gid - row * (uint32_t)d.gpr);
    r.k = (int)(row / (uint32_t)d.ny);
    r.j = (int)(row - (uint32_t)r.k * (uint32_t)d.ny);
    const int gx = sl & 1, dy = (sl >> 1) & 1, dz = sl >> 2;
    if (r.xb + gx < d.gpr && r.j + dy < d.ny && r.k + dz < d.nz)
      r.c = group_crossings(S, voff, d, r.xb + gx, r.j + dy, r.k + dz);
  }
  return r;
}

__device__ __forceinline__ void emit_put(EmitPass& P, const EmitRec& r, int ng) {
  const int lane = threadIdx.x & 63;
  const int b = lane >> 3, sl = lane & 7;
  if (b < ng) {
    P.nb[b][sl] = r.c;
    if (sl == 0) {
      P.toff[b] = r.toff;
      P.xb[b] = r.xb;
      P.j[b] = r.j;
      P.k[b] = r.k;
    }
  }
}

// Steps 2 and 3: the cells with work of each of the pass's ng groups (lane b), how many
// groups the pass takes (returned) and the list of its cells (P.item, P.pre); *nitems =
// the number of listed cells.  The records of the ng groups are in P.
__device__ __forceinline__ int emit_plan(EmitPass& P, int ng, const McDims& d, int* nitems_out) {
  const int lane = threadIdx.x & 63;
  uint64_t items = 0ull;
  int nit = 0;
  if (lane < ng) {
    const Cross* n = P.nb[lane];
    const int xb = P.xb[lane], j = P.j[lane], k = P.k[lane];
    const uint64_t wa = n[0].w, wax = n[1].w, way = n[2].w, waxy = n[3].w;
    const uint64_t wb = n[4].w, wbx = n[5].w, wby = n[6].w, wbxy = n[7].w;
    const uint64_t wa1 = (wa >> 1) | (wax << 63), way1 = (way >> 1) | (waxy << 63);
    const uint64_t wb1 = (wb >> 1) | (wbx << 63), wby1 = (wby >> 1) | (wbxy << 63);
    const int nvx = min(64, d.nx - 1 - xb * 64);
    const uint64_t xmask = nvx >= 64 ? ~0ull : (nvx > 0 ? ((1ull << nvx) - 1) : 0ull);
    const uint64_t cells = (j + 1 < d.ny && k + 1 < d.nz) ? xmask : 0ull;
    const uint64_t any = (wa | wa1 | way | way1 | wb | wb1 | wby | wby1) & cells;
    const uint64_t all = (wa & wa1 & way & way1 & wb & wb1 & wby & wby1) & cells;
    items = (any & ~all) | n[0].bx | n[0].by | n[0].bz;
    nit = __popcll(items);
  }
  static_assert(kEmitPassGroups == 8, "the group scan covers 8 lanes");
  const int incl = wave_inclusive_sum<kEmitPassGroups>(nit);
  const int pre = incl - nit;
  const int npass = __popcll(ballot64(lane < ng && incl <= 64));  // >= 1
  *nitems_out = __builtin_amdgcn_readlane(incl, npass - 1);
  if (lane < npass) P.pre[lane] = pre;
  for (int b = 0; b < npass; b++) {
    const uint64_t it = readlane64(items, b);
    const int p = __builtin_amdgcn_readlane(pre, b);
    if ((it >> lane) & 1ull) P.item[p + __popcll(it & ((1ull << lane) - 1))] = (uint16_t)((b << 6) | lane);
  }
  emit_sync();
  return npass;
}

// One vertex per lane: vertex ve = (listed cell t << 2) | axis of the pass, on the edge
// from the cell's node along the axis (an edge with a crossing, so both ends are inside
// the grid); its id is the id table's entry for that edge.  Position and normal use the
// float operations of the reference's per-edge interpolation: t = (iso - f0) / (f1 - f0)
// along the axis, the central-difference gradients at both ends (one-sided on the grid
// faces) interpolated by t and normalised.  Field reads use offsets clamped to the grid.
// vertex_load issues the reads, vertex_store computes and writes, so that other work
// can run while the reads are in flight.
struct VertexIn {
  float c0, f1;
  float up[3], dn[3];  // the node's neighbours one step up / down each axis (clamped)
  float fp[3], fm[3];  // the far end's: beyond it along the axis, or up / down the others
  int ijk[3], ax;
  uint32_t vid;
};

__device__ __forceinline__ VertexIn vertex_load(const float* __restrict__ f, const McDims& d, const EmitPass& P,
                                                int ve, uint32_t vbase, bool normals) {
  VertexIn v;
  const int t = ve >> 2, a = ve & 3;
  const int e = P.item[t];
  const int b = e >> 6;
  v.ijk[0] = P.xb[b] * 64 + (e & 63);
  v.ijk[1] = P.j[b];
  v.ijk[2] = P.k[b];
  v.ax = a;
  v.vid = vbase + (a == 0 ? P.eid[0][t] : (a == 1 ? P.eid[3][t] : P.eid[8][t]));
  const int nn[3] = {d.nx, d.ny, d.nz};
  const int64_t st[3] = {1, d.nx, d.nxy};
  const float* fn = f + ((int64_t)v.ijk[2] * d.nxy + (int64_t)v.ijk[1] * d.nx + v.ijk[0]);
  const int ia = a == 0 ? v.ijk[0] : (a == 1 ? v.ijk[1] : v.ijk[2]);
  const int na = a == 0 ? nn[0] : (a == 1 ? nn[1] : nn[2]);
  const int64_t sa = a == 0 ? st[0] : (a == 1 ? st[1] : st[2]);
  v.c0 = fn[0];
  v.f1 = fn[sa];
#pragma unroll
  for (int q = 0; q < 3; q++) v.up[q] = v.dn[q] = v.fp[q] = v.fm[q] = 0.0f;
  if (normals) {
    const int64_t up2 = ia + 2 < na ? sa : 0;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const int64_t up = v.ijk[q] + 1 < nn[q] ? st[q] : 0, dn = v.ijk[q] > 0 ? st[q] : 0;
      v.up[q] = fn[up];
      v.dn[q] = fn[-dn];
      const bool along = q == a;
      v.fp[q] = fn[sa + (along ? up2 : up)];
      v.fm[q] = along ? v.c0 : fn[sa - dn];
    }
  }
  return v;
}

__device__ __forceinline__ void vertex_store(const VertexIn& v, const McDims& d, const DGrid& g, float iso,
                                             float* __restrict__ pos, float* __restrict__ nrm, int64_t cap_v) {
  const int a = v.ax;
  const int nn[3] = {d.nx, d.ny, d.nz};
  const float tt = (iso - v.c0) / (v.f1 - v.c0);
  float p[3];
#pragma unroll
  for (int q = 0; q < 3; q++) p[q] = node_coord(g.o[q], g.s[q], v.ijk[q]);
  const int ia = a == 0 ? v.ijk[0] : (a == 1 ? v.ijk[1] : v.ijk[2]);
  const float oa = a == 0 ? g.o[0] : (a == 1 ? g.o[1] : g.o[2]);
  const float spa = a == 0 ? g.s[0] : (a == 1 ? g.s[1] : g.s[2]);
  const float pa = a == 0 ? p[0] : (a == 1 ? p[1] : p[2]);
  const float pe = pa + tt * (node_coord(oa, spa, ia + 1) - pa);
  if ((int64_t)v.vid >= cap_v) return;
  pos[3 * (size_t)v.vid + 0] = a == 0 ? pe : p[0];
  pos[3 * (size_t)v.vid + 1] = a == 1 ? pe : p[1];
  pos[3 * (size_t)v.vid + 2] = a == 2 ? pe : p[2];
  if (!nrm) return;
  float gv[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const float g0 = grad_axis(v.ijk[q], nn[q], v.up[q], v.c0, v.dn[q], g.s[q]);
    const bool along = q == a;
    const float g1 = grad_axis(along ? v.ijk[q] + 1 : v.ijk[q], nn[q], v.fp[q], v.f1, v.fm[q], g.s[q]);
    gv[q] = g0 + tt * (g1 - g0);
  }
  const float len = __builtin_sqrtf(gv[0] * gv[0] + gv[1] * gv[1] + gv[2] * gv[2]);
  const float inv = len > 0.0f ? 1.0f / len : 0.0f;
  nrm[3 * (size_t)v.vid + 0] = gv[0] * inv;
  nrm[3 * (size_t)v.vid + 1] = gv[1] * inv;
  nrm[3 * (size_t)v.vid + 2] = gv[2] * inv;
}

// Step 4: one listed cell per lane for the edge id table and the cell's triangles; the
// vertices on the listed nodes' own edges listed in id order, one per lane.
__device__ __forceinline__ void emit_cells(const float* __restrict__ f, const McDims& d, const DGrid& g, float iso,
                                           EmitPass& P, int nitems, uint32_t vbase, float* __restrict__ pos,
                                           float* __restrict__ nrm, uint32_t* __restrict__ idx, int64_t cap_v,
                                           int64_t cap_t) {
  const int lane = threadIdx.x & 63;
  const bool act_lane = lane < nitems;
  const int e = act_lane ? (int)P.item[lane] : 0;
  const int b = e >> 6, l = e & 63;
  const Cross* nb = P.nb[b];
  const int xb = P.xb[b], j = P.j[b], k = P.k[b];
  const int i = xb * 64 + l;

  // triangles of the cell: the case code from the sign words of its 4 corner rows and
  // their +x neighbour groups; ids from the in-group prefix of the triangle counts
  int code = 0;
  if (act_lane && i + 1 < d.nx && j + 1 < d.ny && k + 1 < d.nz) {
    // corner at x = i: bit l of the slot's word; at x = i + 1: bit l + 1, or bit 0 of the
    // next group's word (slot + 1) for the last cell of the segment
    auto bit = [&](int s0) { return (int)((nb[s0].w >> l) & 1ull); };
    auto bit1 = [&](int s0) { return l < 63 ? (int)((nb[s0].w >> (l + 1)) & 1ull) : (int)(nb[s0 + 1].w & 1ull); };
    code = bit(0) | (bit1(0) << 1) | (bit1(2) << 2) | (bit(2) << 3) | (bit(4) << 4) | (bit1(4) << 5) |
           (bit1(6) << 6) | (bit(6) << 7);
  }
  const int nt = upf_mc_ntris[code];
  const int inc = wave_inclusive_sum(nt);
  const int exc = inc - nt;
  // the triangles of the pass's earlier groups: exc at the first cell of this lane's group
  // (b is the lane's own group, so the source lane differs between lanes)
  const int seg = __shfl(exc, P.pre[b], 64);
  // The vertex ids of the cell's 12 edges, once per lane, into the lane's column of
  // P.eid: an edge's vertex is owned by corner (dx, dy, dz) of the cell (its x, y and z
  // crossings, in that order, after those of the lower nodes of the corner's row
  // segment).  Edges without a crossing get an id that no triangle uses.
  {
    struct Node { uint32_t base, xb, yb; };
    auto node = [&](int dx, int dy, int dz) {
      const int ol = l + dx;
      const Cross& cc = nb[(ol >> 6) | (dy << 1) | (dz << 2)];
      const int lo = ol & 63;
      const uint64_t m = (1ull << lo) - 1;
      Node n;
      n.base = cc.voff - vbase + (uint32_t)(__popcll(cc.bx & m) + __popcll(cc.by & m) + __popcll(cc.bz & m));
      n.xb = (uint32_t)((cc.bx >> lo) & 1ull);
      n.yb = (uint32_t)((cc.by >> lo) & 1ull);
      return n;
    };
    static_assert(mc_ct::owner_order_ok(), "edge order of the id table");
    // corner by corner, each corner's edges stored as soon as it is known
    {
      const Node n = node(0, 0, 0);
      P.eid[0][lane] = n.base;               // x
      P.eid[3][lane] = n.base + n.xb;        // y
      P.eid[8][lane] = n.base + n.xb + n.yb; // z
    }
    {
      const Node n = node(1, 0, 0);
      P.eid[1][lane] = n.base + n.xb;        // y
      P.eid[9][lane] = n.base + n.xb + n.yb; // z
    }
    {
      const Node n = node(0, 1, 0);
      P.eid[2][lane] = n.base;                // x
      P.eid[11][lane] = n.base + n.xb + n.yb; // z
    }
    {
      const Node n = node(0, 0, 1);
      P.eid[4][lane] = n.base;         // x
      P.eid[7][lane] = n.base + n.xb;  // y
    }
    P.eid[5][lane] = [&] { const Node n = node(1, 0, 1); return n.base + n.xb; }();           // y
    P.eid[6][lane] = node(0, 1, 1).base;                                                     // x
    P.eid[10][lane] = [&] { const Node n = node(1, 1, 0); return n.base + n.xb + n.yb; }();  // z
  }
  // the node's own vertices (x, y, z crossings in that order), listed in id order: the
  // listed cells are in group order and, within a group, in node order
  int vinc_all;
  {
    uint32_t ob = 0u;
    if (act_lane)
      ob = (uint32_t)((nb[0].bx >> l) & 1ull) | ((uint32_t)((nb[0].by >> l) & 1ull) << 1) |
           ((uint32_t)((nb[0].bz >> l) & 1ull) << 2);
    const int nvo = __popc(ob);
    const int vinc = wave_inclusive_sum(nvo);
    vinc_all = vinc;
    int vp = vinc - nvo;
    if (ob & 1u) P.vitem[vp++] = (uint16_t)(lane << 2);
    if (ob & 2u) P.vitem[vp++] = (uint16_t)((lane << 2) | 1);
    if (ob & 4u) P.vitem[vp] = (uint16_t)((lane << 2) | 2);
    emit_sync();
  }
  // the first 64 vertices' reads in flight while the triangles are written
  const int nverts = uniform(__builtin_amdgcn_readlane(vinc_all, 63));
  const bool vfirst = lane < nverts;
  VertexIn vin;
  if (vfirst) vin = vertex_load(f, d, P, P.vitem[lane], vbase, nrm != nullptr);
  if (nt) {
    const uint32_t tid0 = P.toff[b] + (uint32_t)(exc - seg);
    // the cell's triangulation row in one 16-byte load
    const int4 row = *reinterpret_cast<const int4*>(&c_tris16.v[code][0]);
    const uint32_t rw[4] = {(uint32_t)row.x, (uint32_t)row.y, (uint32_t)row.z, (uint32_t)row.w};
#pragma unroll
    for (int t = 0; t < UPF_MC_MAX_TRIS; t++) {
      if (t >= nt) break;
      const uint32_t tid = tid0 + (uint32_t)t;
#pragma unroll
      for (int v = 0; v < 3; v++) {
        const int q = 3 * t + v;
        const int ed = (int)((rw[q >> 2] >> (8 * (q & 3))) & 0xFu);  // 0..11 on a triangle row
        if ((int64_t)tid < cap_t) idx[3 * (size_t)tid + v] = P.eid[ed][lane];
      }
    }
  }
  if (vfirst) vertex_store(vin, d, g, iso, pos, nrm, cap_v);
  for (int v0 = 64; v0 < nverts; v0 += 64)
    if (v0 + lane < nverts)
      vertex_store(vertex_load(f, d, P, P.vitem[v0 + lane], vbase, nrm != nullptr), d, g, iso, pos, nrm, cap_v);
}

__device__ __forceinline__ int mc_emit_pass(const float* __restrict__ f, const McDims& d, const DGrid& g, float iso,
                                            const uint64_t* __restrict__ S, const uint32_t* __restrict__ voff,
                                            const uint2* __restrict__ act, int64_t a, int ng, uint32_t gbase,
                                            uint32_t glimit, uint32_t vbase, EmitPass& P, float* __restrict__ pos,
                                            float* __restrict__ nrm, uint32_t* __restrict__ idx, int64_t cap_v,
                                            int64_t cap_t) {
  const int lane = threadIdx.x & 63;
  // entries of another grid end the pass (the list is sorted, so a grid's entries are
  // contiguous); entry a is the caller's grid, so every pass takes at least one group
  {
    const bool mine = lane < ng && act[a + lane].x < glimit;
    ng = max(__popcll(ballot64(mine) & 0xFFull), 1);
  }
  {
    const int b = lane >> 3;
    const uint2 e = b < ng ? act[a + b] : make_uint2(0u, 0u);
    emit_put(P, emit_record(S, voff, d, e, b < ng, gbase), ng);
  }
  emit_sync();
  int nitems;
  const int npass = emit_plan(P, ng, d, &nitems);
  emit_cells(f, d, g, iso, P, nitems, vbase, pos, nrm, idx, cap_v, cap_t);
  emit_sync();  // this pass's reads of P before the next pass's writes
  return npass;
}
