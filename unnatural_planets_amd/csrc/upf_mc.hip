// upf_mc.hip -- marching cubes over an FP32 grid: classify, scan, emit.
//
// Replaces cage's MarchingCubes::makeMesh() (sources/meshGeneration.cpp:39).
//
// Layout.  Nodes are cut into "groups" of 64 consecutive x nodes of one row (a row of
// nx nodes has gpr = ceil(nx/64) groups), one wavefront per group, lane = x offset.
//  classify  each lane reads its node and the +x/+y/+z neighbours (and the other
//            cell corners), the wave ballots the inside bits into a 64-bit sign mask
//            S[g] and counts the crossing edges the group's nodes own (+x, +y, +z) and
//            the triangles of the cells whose min corner it holds;
//  scan      exclusive prefix over the per-group (vertex, triangle) counts;
//  emit      groups with work recompute their crossings, write each owned vertex at
//            voff[g] + rank (rank = popcounts of the lower lanes' crossing masks, then
//            the axis order x<y<z), and write each triangle's three vertex ids.  The
//            id of an edge owned by another group is rebuilt from that group's sign
//            mask and its +x/+y/+z neighbour masks, so one pass deduplicates vertices
//            without a hash table or a dense edge-id array.
// HBM traffic: the field is read once by classify (neighbour reads hit L2) plus the
// crossing cells' corners again in emit; scratch is 16 B per group (S, counts,
// offsets), about 0.25 B per node.
#include <hip/hip_runtime.h>

#include <cmath>

#include "upf_internal.hpp"

namespace upf {

#define UPF_MC_QUAL __constant__
#include "mc_table.inc"
#undef UPF_MC_QUAL
#define c_ntris upf_mc_ntris
#define c_tris upf_mc_tris

// edge -> (owner corner offset x,y,z, axis)
__constant__ int8_t c_edge_owner[12][4] = {
    {0, 0, 0, 0}, {1, 0, 0, 1}, {0, 1, 0, 0}, {0, 0, 0, 1}, {0, 0, 1, 0}, {1, 0, 1, 1},
    {0, 1, 1, 0}, {0, 0, 1, 1}, {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}};

struct McDims {
  int nx, ny, nz, gpr;
  int64_t nxy;
  int64_t groups;
};

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

__device__ __forceinline__ int case_code(const float* __restrict__ f, int64_t n, int nx, int64_t nxy, float iso) {
  const float c0 = f[n], c1 = f[n + 1], c3 = f[n + nx], c2 = f[n + nx + 1];
  const float c4 = f[n + nxy], c5 = f[n + nxy + 1], c7 = f[n + nxy + nx], c6 = f[n + nxy + nx + 1];
  return (c0 < iso) | ((c1 < iso) << 1) | ((c2 < iso) << 2) | ((c3 < iso) << 3) | ((c4 < iso) << 4) |
         ((c5 < iso) << 5) | ((c6 < iso) << 6) | ((c7 < iso) << 7);
}

// ------------------------------------------------------------------ classify
__global__ __launch_bounds__(256) void k_mc_classify(const float* __restrict__ f, const McDims d, float iso,
                                                     uint64_t* __restrict__ S, uint2* __restrict__ counts) {
  const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= d.groups) return;
  const int lane = threadIdx.x & 63;
  const int xb = (int)(g % d.gpr);
  const int64_t row = g / d.gpr;
  const int j = (int)(row % d.ny), k = (int)(row / d.ny);
  const int i = xb * 64 + lane;
  const bool valid = i < d.nx;
  const int64_t n = row * d.nx + i;
  const float f0 = valid ? f[n] : 0.0f;
  const bool s0 = valid && f0 < iso;
  const bool hx = valid && i + 1 < d.nx, hy = valid && j + 1 < d.ny, hz = valid && k + 1 < d.nz;
  const bool cx = hx && ((f[n + 1] < iso) != s0);
  const bool cy = hy && ((f[n + d.nx] < iso) != s0);
  const bool cz = hz && ((f[n + d.nxy] < iso) != s0);
  int nt = 0;
  if (hx && hy && hz) nt = c_ntris[case_code(f, n, d.nx, d.nxy, iso)];
  const uint64_t sm = ballot64(s0);
  const int nv = __popcll(ballot64(cx)) + __popcll(ballot64(cy)) + __popcll(ballot64(cz));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) nt += __shfl_xor(nt, off, 64);
  if (lane == 0) {
    S[g] = sm;
    counts[g] = make_uint2((unsigned)nv, (unsigned)nt);
  }
}

// ------------------------------------------------------------------ scan
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint2 add2(uint2 a, uint2 b) { return make_uint2(a.x + b.x, a.y + b.y); }

// Block-wide exclusive scan of one uint2 per thread; returns the block total in *total.
__device__ __forceinline__ uint2 block_exclusive_scan(uint2 v, uint2* total) {
  __shared__ uint2 wsum[kScanThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint2 inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned ax = __shfl_up(inc.x, off, 64), ay = __shfl_up(inc.y, off, 64);
    if (lane >= off) { inc.x += ax; inc.y += ay; }
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint2 before = make_uint2(0, 0), all = make_uint2(0, 0);
  for (int w = 0; w < kScanThreads / 64; w++) {
    if (w < wave) before = add2(before, wsum[w]);
    all = add2(all, wsum[w]);
  }
  __syncthreads();
  *total = all;
  return make_uint2(before.x + inc.x - v.x, before.y + inc.y - v.y);
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const uint2* __restrict__ in, int64_t n,
                                                             uint2* __restrict__ tile_sums) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint2 s = make_uint2(0, 0);
#pragma unroll
  for (int t = 0; t < kScanItems; t++)
    if (base + t < n) s = add2(s, in[base + t]);
  uint2 total;
  (void)block_exclusive_scan(s, &total);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Single workgroup: exclusive scan of the tile sums in place, grand total to *total.
__global__ __launch_bounds__(kScanThreads) void k_scan_top(uint2* __restrict__ sums, int64_t n,
                                                           uint2* __restrict__ total) {
  uint2 carry = make_uint2(0, 0);
  for (int64_t base = 0; base < n; base += kScanTile) {
    const int64_t b = base + (int64_t)threadIdx.x * kScanItems;
    uint2 v[kScanItems];
    uint2 s = make_uint2(0, 0);
#pragma unroll
    for (int t = 0; t < kScanItems; t++) {
      v[t] = (b + t < n) ? sums[b + t] : make_uint2(0, 0);
      s = add2(s, v[t]);
    }
    uint2 chunk;
    uint2 ex = add2(block_exclusive_scan(s, &chunk), carry);
#pragma unroll
    for (int t = 0; t < kScanItems; t++) {
      if (b + t < n) sums[b + t] = ex;
      ex = add2(ex, v[t]);
    }
    carry = add2(carry, chunk);
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint2* __restrict__ in, int64_t n,
                                                             const uint2* __restrict__ tile_off,
                                                             uint2* __restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint2 v[kScanItems];
  uint2 s = make_uint2(0, 0);
#pragma unroll
  for (int t = 0; t < kScanItems; t++) {
    v[t] = (base + t < n) ? in[base + t] : make_uint2(0, 0);
    s = add2(s, v[t]);
  }
  uint2 total;
  uint2 ex = add2(block_exclusive_scan(s, &total), tile_off[blockIdx.x]);
#pragma unroll
  for (int t = 0; t < kScanItems; t++) {
    if (base + t < n) out[base + t] = ex;
    ex = add2(ex, v[t]);
  }
}

// ------------------------------------------------------------------ emit
// Crossing masks of group gg (row-aligned, so +y and +z neighbour groups share its
// lane alignment).  Lanes past nx hold 0 sign bits and are masked.
struct Cross {
  uint64_t bx, by, bz;
};

__device__ __forceinline__ Cross group_crossings(const uint64_t* __restrict__ S, const McDims& d, int64_t gg) {
  const int xb = (int)(gg % d.gpr);
  const int64_t row = gg / d.gpr;
  const int j = (int)(row % d.ny), k = (int)(row / d.ny);
  const int first = xb * 64;
  const int nval = min(64, d.nx - first);                 // lanes inside the row
  const int nvx = min(64, d.nx - 1 - first);              // lanes with an x neighbour
  const uint64_t vmask = nval >= 64 ? ~0ull : ((1ull << nval) - 1);
  const uint64_t xmask = nvx >= 64 ? ~0ull : (nvx > 0 ? ((1ull << nvx) - 1) : 0ull);
  const uint64_t s = S[gg];
  const uint64_t snext = (xb + 1 < d.gpr) ? S[gg + 1] : 0ull;
  Cross c;
  c.bx = (s ^ ((s >> 1) | (snext << 63))) & xmask;
  c.by = (j + 1 < d.ny) ? ((s ^ S[gg + d.gpr]) & vmask) : 0ull;
  c.bz = (k + 1 < d.nz) ? ((s ^ S[gg + (int64_t)d.gpr * d.ny]) & vmask) : 0ull;
  return c;
}

__device__ __forceinline__ float node_gradient_axis(const float* __restrict__ f, int64_t n, int idx, int nmax,
                                                    int64_t step, float sp) {
  if (idx == 0) return (f[n + step] - f[n]) / sp;
  if (idx == nmax - 1) return (f[n] - f[n - step]) / sp;
  return (f[n + step] - f[n - step]) / (2.0f * sp);
}

__device__ __forceinline__ void node_gradient(const float* __restrict__ f, const McDims& d, const DGrid& g, int i,
                                              int j, int k, float out[3]) {
  const int64_t n = (int64_t)k * d.nxy + (int64_t)j * d.nx + i;
  out[0] = node_gradient_axis(f, n, i, d.nx, 1, g.s[0]);
  out[1] = node_gradient_axis(f, n, j, d.ny, d.nx, g.s[1]);
  out[2] = node_gradient_axis(f, n, k, d.nz, d.nxy, g.s[2]);
}

__global__ __launch_bounds__(256) void k_mc_emit(const float* __restrict__ f, const McDims d, const DGrid g,
                                                 float iso, const uint64_t* __restrict__ S,
                                                 const uint2* __restrict__ counts, const uint2* __restrict__ offs,
                                                 float* __restrict__ pos, float* __restrict__ nrm,
                                                 uint32_t* __restrict__ idx) {
  const int64_t gid = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (gid >= d.groups) return;
  const uint2 cnt = counts[gid];
  if (cnt.x == 0 && cnt.y == 0) return;
  const uint2 off = offs[gid];
  const int lane = threadIdx.x & 63;
  const int xb = (int)(gid % d.gpr);
  const int64_t row = gid / d.gpr;
  const int j = (int)(row % d.ny), k = (int)(row / d.ny);
  const int i = xb * 64 + lane;
  const bool valid = i < d.nx;
  const int64_t n = row * d.nx + i;

  // ---- vertices
  if (cnt.x) {
    const Cross c = group_crossings(S, d, gid);
    const uint64_t lt = (1ull << lane) - 1;
    uint32_t vid = off.x + (uint32_t)(__popcll(c.bx & lt) + __popcll(c.by & lt) + __popcll(c.bz & lt));
    const float f0 = valid ? f[n] : 0.0f;
    const float p0[3] = {node_coord(g.o[0], g.s[0], i), node_coord(g.o[1], g.s[1], j),
                         node_coord(g.o[2], g.s[2], k)};
    float g0[3] = {0.0f, 0.0f, 0.0f};
    const bool any = ((c.bx | c.by | c.bz) >> lane) & 1ull;
    if (any && nrm) node_gradient(f, d, g, i, j, k, g0);
    const uint64_t bits[3] = {c.bx, c.by, c.bz};
    const int64_t step[3] = {1, d.nx, d.nxy};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      if (!((bits[a] >> lane) & 1ull)) continue;
      const float f1 = f[n + step[a]];
      const float t = (iso - f0) / (f1 - f0);
      const int ijk[3] = {i, j, k};
      const float p1a = node_coord(g.o[a], g.s[a], ijk[a] + 1);
      float p[3] = {p0[0], p0[1], p0[2]};
      p[a] = p0[a] + t * (p1a - p0[a]);
      pos[3 * (size_t)vid + 0] = p[0];
      pos[3 * (size_t)vid + 1] = p[1];
      pos[3 * (size_t)vid + 2] = p[2];
      if (nrm) {
        float g1[3];
        node_gradient(f, d, g, i + (a == 0), j + (a == 1), k + (a == 2), g1);
        const float gx = g0[0] + t * (g1[0] - g0[0]);
        const float gy = g0[1] + t * (g1[1] - g0[1]);
        const float gz = g0[2] + t * (g1[2] - g0[2]);
        const float len = __builtin_sqrtf(gx * gx + gy * gy + gz * gz);
        const float inv = len > 0.0f ? 1.0f / len : 0.0f;
        nrm[3 * (size_t)vid + 0] = gx * inv;
        nrm[3 * (size_t)vid + 1] = gy * inv;
        nrm[3 * (size_t)vid + 2] = gz * inv;
      }
      vid++;
    }
  }

  // ---- triangles
  if (cnt.y) {
    const bool cell = valid && i + 1 < d.nx && j + 1 < d.ny && k + 1 < d.nz;
    const int code = cell ? case_code(f, n, d.nx, d.nxy, iso) : 0;
    const int nt = c_ntris[code];
    int inc = nt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int a = __shfl_up(inc, o, 64);
      if (lane >= o) inc += a;
    }
    uint32_t tid = off.y + (uint32_t)(inc - nt);
    for (int t = 0; t < nt; t++, tid++) {
#pragma unroll
      for (int v = 0; v < 3; v++) {
        const int e = c_tris[code][3 * t + v];
        const int ox = i + c_edge_owner[e][0], oy = j + c_edge_owner[e][1], oz = k + c_edge_owner[e][2];
        const int axis = c_edge_owner[e][3];
        const int64_t og = ((int64_t)oz * d.ny + oy) * d.gpr + (ox >> 6);
        const int ol = ox & 63;
        const Cross c = group_crossings(S, d, og);
        const uint64_t lt = (1ull << ol) - 1;
        uint32_t id = offs[og].x + (uint32_t)(__popcll(c.bx & lt) + __popcll(c.by & lt) + __popcll(c.bz & lt));
        if (axis >= 1) id += (uint32_t)((c.bx >> ol) & 1ull);
        if (axis == 2) id += (uint32_t)((c.by >> ol) & 1ull);
        idx[3 * (size_t)tid + v] = id;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_mc_codes(const float* __restrict__ f, const McDims d, float iso,
                                                  uint8_t* __restrict__ codes) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int k = blockIdx.z;
  if (i >= d.nx - 1 || j >= d.ny - 1) return;
  const int64_t n = (int64_t)k * d.nxy + (int64_t)j * d.nx + i;
  codes[((int64_t)k * (d.ny - 1) + j) * (d.nx - 1) + i] = (uint8_t)case_code(f, n, d.nx, d.nxy, iso);
}

namespace {

struct McLayout {
  McDims d;
  int64_t tiles;
  size_t off_S, off_counts, off_offs, off_tiles, off_total, bytes;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

McLayout mc_layout(const DGrid& g) {
  McLayout L;
  L.d.nx = g.n[0];
  L.d.ny = g.n[1];
  L.d.nz = g.n[2];
  L.d.gpr = (g.n[0] + 63) / 64;
  L.d.nxy = (int64_t)g.n[0] * g.n[1];
  L.d.groups = (int64_t)L.d.gpr * g.n[1] * g.n[2];
  L.tiles = (L.d.groups + kScanTile - 1) / kScanTile;
  size_t o = 0;
  L.off_S = o; o = align256(o + sizeof(uint64_t) * L.d.groups);
  L.off_counts = o; o = align256(o + sizeof(uint2) * L.d.groups);
  L.off_offs = o; o = align256(o + sizeof(uint2) * L.d.groups);
  L.off_tiles = o; o = align256(o + sizeof(uint2) * L.tiles);
  L.off_total = o; o = align256(o + sizeof(uint2));
  L.bytes = o;
  return L;
}

}  // namespace

}  // namespace upf

using namespace upf;

extern "C" {

size_t upf_mc_workspace_bytes(const upf_grid* grid) {
  DGrid g;
  if (to_device(grid, &g)) return 0;
  return mc_layout(g).bytes;
}

int upf_mc_count(const float* d_field, const upf_grid* grid, float iso, void* d_ws, size_t ws_bytes,
                 int64_t* n_vertices, int64_t* n_triangles, void* stream) {
  if (!d_field || !n_vertices || !n_triangles) return UPF_ERR_NULL;
  if (!std::isfinite(iso)) return UPF_ERR_INVALID_VALUE;
  DGrid g;
  const int st = to_device(grid, &g);
  if (st) return st;
  const McLayout L = mc_layout(g);
  if (!d_ws || ws_bytes < L.bytes) return UPF_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  char* ws = static_cast<char*>(d_ws);
  uint64_t* S = reinterpret_cast<uint64_t*>(ws + L.off_S);
  uint2* counts = reinterpret_cast<uint2*>(ws + L.off_counts);
  uint2* offs = reinterpret_cast<uint2*>(ws + L.off_offs);
  uint2* tiles = reinterpret_cast<uint2*>(ws + L.off_tiles);
  uint2* total = reinterpret_cast<uint2*>(ws + L.off_total);
  const unsigned nblk = (unsigned)((L.d.groups + 3) / 4);
  hipLaunchKernelGGL(k_mc_classify, dim3(nblk), dim3(256), 0, s, d_field, L.d, iso, S, counts);
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)L.tiles), dim3(kScanThreads), 0, s, counts, L.d.groups, tiles);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, tiles, L.tiles, total);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)L.tiles), dim3(kScanThreads), 0, s, counts, L.d.groups, tiles,
                     offs);
  UPF_LAUNCH_CHECK();
  uint2 h_total;
  UPF_HIP_CHECK(hipMemcpyAsync(&h_total, total, sizeof(uint2), hipMemcpyDeviceToHost, s));
  UPF_HIP_CHECK(hipStreamSynchronize(s));
  *n_vertices = h_total.x;
  *n_triangles = h_total.y;
  return UPF_OK;
}

int upf_mc_emit(const float* d_field, const upf_grid* grid, float iso, upf_mesh_buffers* out, void* d_ws,
                size_t ws_bytes, void* stream) {
  if (!d_field || !out) return UPF_ERR_NULL;
  if (!std::isfinite(iso)) return UPF_ERR_INVALID_VALUE;
  DGrid g;
  const int st = to_device(grid, &g);
  if (st) return st;
  const McLayout L = mc_layout(g);
  if (!d_ws || ws_bytes < L.bytes) return UPF_ERR_WORKSPACE;
  if (out->n_vertices > out->cap_vertices || out->n_triangles > out->cap_triangles) return UPF_ERR_CAPACITY;
  if (out->n_vertices == 0 && out->n_triangles == 0) return UPF_OK;
  if (!out->positions || !out->indices) return UPF_ERR_NULL;
  char* ws = static_cast<char*>(d_ws);
  const uint64_t* S = reinterpret_cast<const uint64_t*>(ws + L.off_S);
  const uint2* counts = reinterpret_cast<const uint2*>(ws + L.off_counts);
  const uint2* offs = reinterpret_cast<const uint2*>(ws + L.off_offs);
  const unsigned nblk = (unsigned)((L.d.groups + 3) / 4);
  hipLaunchKernelGGL(k_mc_emit, dim3(nblk), dim3(256), 0, as_stream(stream), d_field, L.d, g, iso, S, counts, offs,
                     out->positions, out->normals, out->indices);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

int upf_mc_extract(const float* d_field, const upf_grid* grid, float iso, upf_mesh_buffers* out, void* d_ws,
                   size_t ws_bytes, void* stream) {
  if (!out) return UPF_ERR_NULL;
  int64_t nv = 0, nt = 0;
  const int st = upf_mc_count(d_field, grid, iso, d_ws, ws_bytes, &nv, &nt, stream);
  if (st) return st;
  out->n_vertices = nv;
  out->n_triangles = nt;
  if (nv > out->cap_vertices || nt > out->cap_triangles) return UPF_ERR_CAPACITY;
  return upf_mc_emit(d_field, grid, iso, out, d_ws, ws_bytes, stream);
}

int upf_mc_case_codes(const float* d_field, const upf_grid* grid, float iso, uint8_t* d_codes, void* stream) {
  if (!d_field || !d_codes) return UPF_ERR_NULL;
  DGrid g;
  const int st = to_device(grid, &g);
  if (st) return st;
  const McLayout L = mc_layout(g);
  const dim3 blocks((g.n[0] - 1 + 63) / 64, (g.n[1] - 1 + 3) / 4, g.n[2] - 1);
  hipLaunchKernelGGL(k_mc_codes, blocks, dim3(256), 0, as_stream(stream), d_field, L.d, iso, d_codes);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

}  // extern "C"
