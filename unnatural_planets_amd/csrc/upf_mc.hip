// upf_mc.hip -- marching cubes over an FP32 grid: classify, scan, emit.
//
// Replaces cage's MarchingCubes::makeMesh() (sources/meshGeneration.cpp:39).
//
// Layout.  Nodes are cut into "groups" of 64 consecutive x nodes of one row (a row of
// nx nodes has gpr = ceil(nx/64) groups); a wavefront handles a group, lane = x offset.
//  classify  a workgroup owns one 64-wide column of 4 rows and sweeps a z range,
//            keeping the current z plane in registers (each step loads only the next
//            plane: rows j and j+1 at x and x+1).  Per group it ballots the inside bits
//            into a 64-bit sign mask S[g] and packs the counts of the crossing edges
//            the group's nodes own (+x, +y, +z) and of the triangles of the cells whose
//            min corner it holds into one word;
//  scan      exclusive prefix of (vertices, triangles, active groups); writes the dense
//            per-group vertex offset and the compacted list of active groups;
//  emit      one wave per active group writes each owned vertex at voff[g] + rank
//            (popcounts of the lower lanes' crossing masks, then axis order x<y<z) and
//            each triangle's three vertex ids.  An edge owned by a neighbouring group is
//            resolved from that group's crossing masks, rebuilt from the sign masks, so
//            one pass deduplicates vertices with neither a hash table nor a dense
//            per-edge id array.
// Scratch is 24 B per group (S, packed counts, voff, active entry), ~0.4 B per node.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>

#include "upf_internal.hpp"
#include "upf_mc.hpp"

namespace upf {

#include "upf_mc_device.hpp"

// ------------------------------------------------------------------ sign words / emit kernels
// Sign words of a stored field: bit x of S[g] = (node x of the group's row segment
// exists and f < iso), the words the field kernel stores when it evaluates a grid, so a
// stored field is classified by the same sign-word kernel as a fused one.  One wave per
// grid row and run of kSignGroups consecutive groups (blockIdx.y): kSignGroups loads in
// flight per lane, one ballot per group, and lane u stores the word of group u.  The
// field is read once, 4 B per node.
constexpr int kSignGroups = 8;

__global__ __launch_bounds__(256) void k_mc_signs(const float* __restrict__ f, const McDims d, float iso,
                                                  int64_t rows, uint64_t* __restrict__ S) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;  // whole waves
  const int lane = threadIdx.x & 63;
  const int g0 = blockIdx.y * kSignGroups;
  const float* fr = f + row * d.nx;
  float v[kSignGroups];
#pragma unroll
  for (int u = 0; u < kSignGroups; u++) {
    const int x = (g0 + u) * 64 + lane;
    v[u] = x < d.nx ? fr[x] : 0.0f;
  }
  uint64_t mine = 0ull;
#pragma unroll
  for (int u = 0; u < kSignGroups; u++) {
    const int x = (g0 + u) * 64 + lane;
    const uint64_t m = ballot64(x < d.nx && v[u] < iso);
    if (lane == u) mine = m;
  }
  if (lane < kSignGroups && g0 + lane < d.gpr) S[row * d.gpr + g0 + lane] = mine;
}

// At least five waves per SIMD (at most 102 VGPRs): the edge-id table's corners took the
// kernel to 112 VGPRs and four waves, 0.224 against 0.211 ms stand-alone at 512^3.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_mc_emit(
                                                 const float* __restrict__ f, const McDims d, const DGrid g,
                                                 float iso, const uint64_t* __restrict__ S,
                                                 const uint32_t* __restrict__ voff, const uint2* __restrict__ act,
                                                 const uint4* __restrict__ total, float* __restrict__ pos,
                                                 float* __restrict__ nrm, uint32_t* __restrict__ idx, int64_t cap_v,
                                                 int64_t cap_t, int64_t* __restrict__ counts_out) {
  __shared__ EmitPass s_pass[4];
  const int wave = uniform(threadIdx.x >> 6);
  if (counts_out && blockIdx.x == 0 && threadIdx.x < 2)
    counts_out[threadIdx.x] = (int64_t)(threadIdx.x ? total->y : total->x);
  // each wave takes a contiguous run of the active-group list (neighbouring groups share
  // sign words and field values in cache) in passes of up to kEmitPassGroups groups; the
  // grid is sized to keep the CUs full, the list's length is known only on the device.
  // At 512^3 (185,747 active groups) the passes issue ~32 M VALU wave instructions
  // (~52 us of issue) in ~69 us: issue and the pass's dependent reads (records, then the
  // vertices' field values) share the time.  A software-pipelined variant that issued
  // the next pass's record loads before the current pass's field reads ran no faster.
  const int64_t nact = (int64_t)total->z, nw = (int64_t)gridDim.x * 4;
  const int64_t per = (nact + nw - 1) / nw;
  int64_t a = ((int64_t)blockIdx.x * 4 + wave) * per;
  const int64_t a1 = min(a + per, nact);
  while (a < a1)
    a += mc_emit_pass(f, d, g, iso, S, voff, act, a, (int)min<int64_t>(kEmitPassGroups, a1 - a), 0u, 0xFFFFFFFFu, 0u,
                      s_pass[wave], pos, nrm, idx, cap_v, cap_t);
}

// Packed counts of every group, and the sums of the scan tiles the workgroup covers
// (the scan's first pass): thread t takes groups t + 256 u, so tile q of the workgroup
// is u in [q * kScanTile / 256, (q + 1) * kScanTile / 256).
__global__ __launch_bounds__(256) void k_mc_classify_signs(const uint64_t* __restrict__ S, const McDims d,
                                                           uint32_t* __restrict__ counts,
                                                           uint4* __restrict__ tile_sums) {
  constexpr int kTiles = 256 * kClassifyGroups / kScanTile, kPerTile = kScanTile / 256;
  __shared__ uint8_t s_ntris[256];
  __shared__ SignPass s_pass[4];
  s_ntris[threadIdx.x] = upf_mc_ntris[threadIdx.x];
  __syncthreads();
  // the next group's words are loaded before the current group is classified
  const int64_t g0 = (int64_t)blockIdx.x * (256 * kClassifyGroups) + threadIdx.x;
  uint4 sum[kTiles];
#pragma unroll
  for (int q = 0; q < kTiles; q++) sum[q] = make_uint4(0, 0, 0, 0);
  GroupWords w = load_group_words(S, d, g0, g0 < d.groups);
#pragma unroll
  for (int u = 0; u < kClassifyGroups; u++) {
    const int64_t g = g0 + 256 * u;
    const bool in = g < d.groups;
    GroupWords next = w;
    if (u + 1 < kClassifyGroups) next = load_group_words(S, d, g + 256, g + 256 < d.groups);
    const uint32_t c = classify_group_words(w, d.nx, in, s_ntris, s_pass[threadIdx.x >> 6]);
    if (in) counts[g] = c;
    sum[u / kPerTile] = add4(sum[u / kPerTile], unpack_counts<kScanMc>(in ? c : 0u));
    w = next;
  }
  const int64_t ntiles = (d.groups + kScanTile - 1) / kScanTile;
#pragma unroll
  for (int q = 0; q < kTiles; q++) {
    uint4 total;
    (void)block_exclusive_scan(sum[q], &total);
    const int64_t tile = (int64_t)blockIdx.x * kTiles + q;
    if (threadIdx.x == 0 && tile < ntiles) tile_sums[tile] = total;
  }
}

// ------------------------------------------------------------------ scan
template <int FMT>
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const uint32_t* __restrict__ counts, int64_t n,
                                                             uint4* __restrict__ tile_sums) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint4 s = make_uint4(0, 0, 0, 0);
#pragma unroll
  for (int t = 0; t < kScanItems; t++)
    if (base + t < n) s = add4(s, unpack_counts<FMT>(counts[base + t]));
  uint4 total;
  (void)block_exclusive_scan(s, &total);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the tile sums in place, grand total to *total.
__global__ __launch_bounds__(kScanThreads) void k_scan_top(uint4* __restrict__ sums, int64_t n,
                                                           uint4* __restrict__ total) {
  uint4 carry = make_uint4(0, 0, 0, 0);
  for (int64_t base = 0; base < n; base += kScanTile) {
    const int64_t b = base + (int64_t)threadIdx.x * kScanItems;
    uint4 v[kScanItems];
    uint4 s = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int t = 0; t < kScanItems; t++) {
      v[t] = (b + t < n) ? sums[b + t] : make_uint4(0, 0, 0, 0);
      s = add4(s, v[t]);
    }
    uint4 chunk;
    uint4 ex = add4(block_exclusive_scan(s, &chunk), carry);
#pragma unroll
    for (int t = 0; t < kScanItems; t++) {
      if (b + t < n) sums[b + t] = ex;
      ex = add4(ex, v[t]);
    }
    carry = add4(carry, chunk);
  }
  if (threadIdx.x == 0) *total = carry;
}

// OWN_PREFIX: tiles holds the tiles' sums (not yet scanned) and each workgroup adds up
// those before its own tile itself (at most kScanOwnPrefixTiles of them, L2-resident),
// the last one writing the grand total -- no separate top-level scan launch.  Otherwise
// tiles holds the tiles' exclusive offsets from k_scan_top.
constexpr int64_t kScanOwnPrefixTiles = 4096;

template <int FMT, bool OWN_PREFIX>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t* __restrict__ counts, int64_t n,
                                                             const uint4* __restrict__ tiles,
                                                             uint32_t* __restrict__ voff, uint2* __restrict__ act,
                                                             uint32_t* __restrict__ toff, uint4* __restrict__ grand) {
  uint4 before;
  if (OWN_PREFIX) {
    uint4 p = make_uint4(0, 0, 0, 0);
    for (int t = threadIdx.x; t < (int)blockIdx.x; t += kScanThreads) p = add4(p, tiles[t]);
    (void)block_exclusive_scan(p, &before);
  } else {
    before = tiles[blockIdx.x];
  }
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint4 v[kScanItems];
  uint4 s = make_uint4(0, 0, 0, 0);
#pragma unroll
  for (int t = 0; t < kScanItems; t++) {
    v[t] = (base + t < n) ? unpack_counts<FMT>(counts[base + t]) : make_uint4(0, 0, 0, 0);
    s = add4(s, v[t]);
  }
  uint4 total;
  uint4 ex = add4(block_exclusive_scan(s, &total), before);
  if (OWN_PREFIX && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) *grand = add4(before, total);
#pragma unroll
  for (int t = 0; t < kScanItems; t++) {
    if (base + t < n) {
      if (voff) voff[base + t] = ex.x;
      if (toff) toff[base + t] = ex.y;
      if (act && v[t].z) act[ex.z] = make_uint2((uint32_t)(base + t), ex.y);
    }
    ex = add4(ex, v[t]);
  }
}

__global__ __launch_bounds__(256) void k_mc_codes(const float* __restrict__ f, const McDims d, float iso,
                                                  uint8_t* __restrict__ codes) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int k = blockIdx.z;
  if (i >= d.nx - 1 || j >= d.ny - 1) return;
  const int64_t n = (int64_t)k * d.nxy + (int64_t)j * d.nx + i;
  const float c0 = f[n], c1 = f[n + 1], c3 = f[n + d.nx], c2 = f[n + d.nx + 1];
  const float c4 = f[n + d.nxy], c5 = f[n + d.nxy + 1], c7 = f[n + d.nxy + d.nx], c6 = f[n + d.nxy + d.nx + 1];
  codes[((int64_t)k * (d.ny - 1) + j) * (d.nx - 1) + i] =
      (uint8_t)((c0 < iso) | ((c1 < iso) << 1) | ((c2 < iso) << 2) | ((c3 < iso) << 3) | ((c4 < iso) << 4) |
                ((c5 < iso) << 5) | ((c6 < iso) << 6) | ((c7 < iso) << 7));
}

// ------------------------------------------------------------------ host side
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
  L.off_counts = o; o = align256(o + sizeof(uint32_t) * L.d.groups);
  L.off_voff = o; o = align256(o + sizeof(uint32_t) * L.d.groups);
  L.off_act = o; o = align256(o + sizeof(uint2) * L.d.groups);
  L.off_tiles = o; o = align256(o + sizeof(uint4) * L.tiles);
  L.off_total = o; o = align256(o + sizeof(uint4));
  L.bytes = o;
  return L;
}

McPtrs mc_ptrs(const McLayout& L, void* ws) {
  char* w = static_cast<char*>(ws);
  return {reinterpret_cast<uint64_t*>(w + L.off_S), reinterpret_cast<uint32_t*>(w + L.off_counts),
          reinterpret_cast<uint32_t*>(w + L.off_voff), reinterpret_cast<uint2*>(w + L.off_act),
          reinterpret_cast<uint4*>(w + L.off_tiles), reinterpret_cast<uint4*>(w + L.off_total)};
}

int mc_grid_ok(const DGrid& g) {
  // group ids and vertex ids are 32-bit
  const int64_t groups = (int64_t)((g.n[0] + 63) / 64) * g.n[1] * g.n[2];
  return groups < ((int64_t)1 << 31) ? UPF_OK : UPF_ERR_BAD_GRID;
}

int mc_classify_scan(const float* d_field, const DGrid& g, float iso, void* d_ws, size_t ws_bytes, hipStream_t s) {
  const McLayout L = mc_layout(g);
  if (!d_ws || ws_bytes < L.bytes) return UPF_ERR_WORKSPACE;
  const McPtrs P = mc_ptrs(L, d_ws);
  const int64_t rows = (int64_t)L.d.ny * L.d.nz;
  hipLaunchKernelGGL(k_mc_signs, dim3((unsigned)((rows + 3) / 4), (unsigned)((L.d.gpr + kSignGroups - 1) / kSignGroups)),
                     dim3(256), 0, s, d_field, L.d, iso, rows, P.S);
  UPF_LAUNCH_CHECK();
  return mc_classify_signs_scan(g, d_ws, ws_bytes, s);
}

int mc_scan_launch(const uint32_t* counts, int64_t groups, uint4* tiles, uint4* total, uint32_t* voff, uint2* act,
                   uint32_t* toff, hipStream_t s, int fmt, bool tiles_ready) {
  const int64_t ntiles = (groups + kScanTile - 1) / kScanTile;
  if (ntiles == 0) return UPF_ERR_BAD_GRID;
  if (tiles_ready) {
    // the tile sums were written by the kernel that produced the counts
  } else if (fmt == kScanMc)
    hipLaunchKernelGGL(k_scan_tiles<kScanMc>, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, counts, groups, tiles);
  else
    hipLaunchKernelGGL(k_scan_tiles<kScanBlocks>, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, counts, groups,
                       tiles);
  const dim3 grid((unsigned)ntiles);
  if (ntiles <= kScanOwnPrefixTiles) {
    if (fmt == kScanMc)
      hipLaunchKernelGGL((k_scan_apply<kScanMc, true>), grid, dim3(kScanThreads), 0, s, counts, groups, tiles, voff,
                         act, toff, total);
    else
      hipLaunchKernelGGL((k_scan_apply<kScanBlocks, true>), grid, dim3(kScanThreads), 0, s, counts, groups, tiles,
                         voff, act, toff, total);
  } else {
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, tiles, ntiles, total);
    if (fmt == kScanMc)
      hipLaunchKernelGGL((k_scan_apply<kScanMc, false>), grid, dim3(kScanThreads), 0, s, counts, groups, tiles, voff,
                         act, toff, total);
    else
      hipLaunchKernelGGL((k_scan_apply<kScanBlocks, false>), grid, dim3(kScanThreads), 0, s, counts, groups, tiles,
                         voff, act, toff, total);
  }
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

int mc_classify_signs_scan(const DGrid& g, void* d_ws, size_t ws_bytes, hipStream_t s) {
  const McLayout L = mc_layout(g);
  if (!d_ws || ws_bytes < L.bytes) return UPF_ERR_WORKSPACE;
  const McPtrs P = mc_ptrs(L, d_ws);
  const int64_t per_block = 256 * kClassifyGroups;
  hipLaunchKernelGGL(k_mc_classify_signs, dim3((unsigned)((L.d.groups + per_block - 1) / per_block)), dim3(256), 0, s,
                     P.S, L.d, P.counts, P.tiles);
  UPF_LAUNCH_CHECK();
  return mc_scan_launch(P.counts, L.d.groups, P.tiles, P.total, P.voff, P.act, nullptr, s, kScanMc, true);
}

int64_t mc_scan_tiles(int64_t groups) { return (groups + kScanTile - 1) / kScanTile; }


// ------------------------------------------------------------------ allocation-free scratch
// The workspace of upf_field_mesh_async_static: a file-scope device array, one copy per
// GPU, held until the process ends.  64 MiB holds the marching-cubes workspace of grids
// up to 512^3 nodes (50.4 MB there: 24 B per 64-node group and the scan tiles).  In host
// code `scratch` is only a stand-in for the symbol; each device's address comes from
// hipGetSymbolAddress, looked up once per device on the entry point's first call there.
constexpr size_t kStaticScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// This device's address of `scratch` (the current device's), or an error.
int static_scratch(void** out) {
  int dev = -1;
  UPF_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return UPF_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return UPF_ERR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return UPF_OK;
}

// About 32 workgroups of k_mc_emit per CU, rounded down to whole rounds of the resident
// ones (occupancy API, cached per device): the runs are equal, so the workgroups finish
// in rounds, and a last round only partly filled costs as long as a full one.
constexpr int kEmitBlocksPerCu = 32;
constexpr int kEmitMaxDevices = 64;

int64_t emit_grid_waves() {
  static std::atomic<int64_t> cache[kEmitMaxDevices];
  int dev = 0;
  const bool cached = hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < kEmitMaxDevices;
  if (cached) {
    const int64_t w = cache[dev].load(std::memory_order_relaxed);
    if (w > 0) return w;
  }
  const int cus = device_cus();
  const int per_cu = std::max(1, resident_blocks(reinterpret_cast<const void*>(&k_mc_emit), 256) / cus);
  const int64_t w = (int64_t)cus * per_cu * std::max(1, kEmitBlocksPerCu / per_cu) * 4;
  if (cached) cache[dev].store(w, std::memory_order_relaxed);
  return w;
}

int mc_emit_launch(const float* d_field, const DGrid& g, float iso, void* d_ws, int64_t max_active, float* pos,
                   float* nrm, uint32_t* idx, int64_t cap_v, int64_t cap_t, hipStream_t s, int64_t* counts_out) {
  const McLayout L = mc_layout(g);
  const McPtrs P = mc_ptrs(L, d_ws);
  if (max_active <= 0) return UPF_OK;
  // whole rounds of resident workgroups of 4 waves, each wave looping over its run
  max_active = std::min<int64_t>(max_active, emit_grid_waves());
  hipLaunchKernelGGL(k_mc_emit, dim3((unsigned)((max_active + 3) / 4)), dim3(256), 0, s, d_field, L.d, g, iso, P.S,
                     P.voff, P.act, P.total, pos, nrm, idx, cap_v, cap_t, counts_out);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

}  // namespace upf

using namespace upf;

extern "C" {

size_t upf_mc_workspace_bytes(const upf_grid* grid) {
  DGrid g;
  if (to_device(grid, &g) || mc_grid_ok(g)) return 0;
  return mc_layout(g).bytes;
}

int upf_mc_count(const float* d_field, const upf_grid* grid, float iso, void* d_ws, size_t ws_bytes,
                 int64_t* n_vertices, int64_t* n_triangles, void* stream) {
  if (!d_field || !n_vertices || !n_triangles) return UPF_ERR_NULL;
  if (!std::isfinite(iso)) return UPF_ERR_INVALID_VALUE;
  DGrid g;
  int st = to_device(grid, &g);
  if (st || (st = mc_grid_ok(g))) return st;
  hipStream_t s = as_stream(stream);
  if ((st = mc_classify_scan(d_field, g, iso, d_ws, ws_bytes, s))) return st;
  const McPtrs P = mc_ptrs(mc_layout(g), d_ws);
  uint4 h_total;
  UPF_HIP_CHECK(hipMemcpyAsync(&h_total, P.total, sizeof(uint4), hipMemcpyDeviceToHost, s));
  UPF_HIP_CHECK(hipStreamSynchronize(s));
  *n_vertices = h_total.x;
  *n_triangles = h_total.y;
  return UPF_OK;
}

int upf_field_mesh_count(const upf_field_spec* spec, const upf_grid* grid, float iso, float* d_field,
                         const upf_eval_opts* opts, void* d_ws, size_t ws_bytes, int64_t* n_vertices,
                         int64_t* n_triangles, void* stream) {
  if (!d_field || !n_vertices || !n_triangles) return UPF_ERR_NULL;
  if (!std::isfinite(iso)) return UPF_ERR_INVALID_VALUE;
  DSpec sp;
  DGrid g;
  int st = to_device(spec, &sp);
  if (st || (st = to_device(grid, &g)) || (st = mc_grid_ok(g))) return st;
  const McLayout L = mc_layout(g);
  if (!d_ws || ws_bytes < L.bytes) return UPF_ERR_WORKSPACE;
  const McPtrs P = mc_ptrs(L, d_ws);
  hipStream_t s = as_stream(stream);
  if ((st = field_grid_launch(sp, g, d_field, opts ? opts->variant : UPF_EVAL_AUTO, P.S, iso, s))) return st;
  if ((st = mc_classify_signs_scan(g, d_ws, ws_bytes, s))) return st;
  uint4 h_total;
  UPF_HIP_CHECK(hipMemcpyAsync(&h_total, P.total, sizeof(uint4), hipMemcpyDeviceToHost, s));
  UPF_HIP_CHECK(hipStreamSynchronize(s));
  *n_vertices = h_total.x;
  *n_triangles = h_total.y;
  return UPF_OK;
}

int upf_field_mesh_async(const upf_field_spec* spec, const upf_grid* grid, float iso, float* d_field,
                         const upf_eval_opts* opts, upf_mesh_buffers* out, int64_t* d_counts, void* d_ws,
                         size_t ws_bytes, void* stream) {
  if (!d_field || !out || !d_counts) return UPF_ERR_NULL;
  if (!std::isfinite(iso)) return UPF_ERR_INVALID_VALUE;
  if (out->cap_vertices < 0 || out->cap_triangles < 0) return UPF_ERR_CAPACITY;
  if ((out->cap_vertices && !out->positions) || (out->cap_triangles && !out->indices)) return UPF_ERR_NULL;
  DSpec sp;
  DGrid g;
  int st = to_device(spec, &sp);
  if (st || (st = to_device(grid, &g)) || (st = mc_grid_ok(g))) return st;
  const McLayout L = mc_layout(g);
  if (!d_ws || ws_bytes < L.bytes) return UPF_ERR_WORKSPACE;
  const McPtrs P = mc_ptrs(L, d_ws);
  hipStream_t s = as_stream(stream);
  if ((st = field_grid_launch(sp, g, d_field, opts ? opts->variant : UPF_EVAL_AUTO, P.S, iso, s))) return st;
  if ((st = mc_classify_signs_scan(g, d_ws, ws_bytes, s))) return st;
  // the number of active groups is known only on the device: launch the persistent grid
  // for the largest possible list; the kernel also stores the totals in d_counts
  return mc_emit_launch(d_field, g, iso, d_ws, L.d.groups, out->positions, out->normals, out->indices,
                        out->cap_vertices, out->cap_triangles, s, d_counts);
}

size_t upf_static_scratch_bytes(void) { return kStaticScratchBytes; }

int upf_field_mesh_async_static(const upf_field_spec* spec, const upf_grid* grid, float iso, float* d_field,
                                const upf_eval_opts* opts, upf_mesh_buffers* out, int64_t* d_counts, void* stream) {
  DGrid g;
  int st = to_device(grid, &g);
  if (st || (st = mc_grid_ok(g))) return st;
  if (mc_layout(g).bytes > kStaticScratchBytes) return UPF_ERR_WORKSPACE;
  void* ws = nullptr;
  if ((st = static_scratch(&ws))) return st;
  // the same launches as upf_field_mesh_async; every word of the workspace is written
  // before it is read within a call, so nothing is cleared
  return upf_field_mesh_async(spec, grid, iso, d_field, opts, out, d_counts, ws, kStaticScratchBytes, stream);
}

int upf_mc_emit(const float* d_field, const upf_grid* grid, float iso, upf_mesh_buffers* out, void* d_ws,
                size_t ws_bytes, void* stream) {
  if (!d_field || !out) return UPF_ERR_NULL;
  if (!std::isfinite(iso)) return UPF_ERR_INVALID_VALUE;
  DGrid g;
  int st = to_device(grid, &g);
  if (st || (st = mc_grid_ok(g))) return st;
  const McLayout L = mc_layout(g);
  if (!d_ws || ws_bytes < L.bytes) return UPF_ERR_WORKSPACE;
  if (out->n_vertices > out->cap_vertices || out->n_triangles > out->cap_triangles) return UPF_ERR_CAPACITY;
  if (out->n_vertices == 0 && out->n_triangles == 0) return UPF_OK;
  if (!out->positions || !out->indices) return UPF_ERR_NULL;
  // The kernel reads the number of active groups from the workspace; launch enough
  // waves for every group that can own a vertex or a triangle (at most one per vertex
  // or triangle, at most one per group) and let the surplus exit.
  const int64_t bound = std::min<int64_t>(L.d.groups, out->n_vertices + out->n_triangles);
  return mc_emit_launch(d_field, g, iso, d_ws, bound, out->positions, out->normals, out->indices, out->cap_vertices,
                        out->cap_triangles, as_stream(stream));
}

int upf_mc_extract_async(const float* d_field, const upf_grid* grid, float iso, upf_mesh_buffers* out,
                         int64_t* d_counts, void* d_ws, size_t ws_bytes, void* stream) {
  if (!d_field || !out || !d_counts) return UPF_ERR_NULL;
  if (!std::isfinite(iso)) return UPF_ERR_INVALID_VALUE;
  if (out->cap_vertices < 0 || out->cap_triangles < 0) return UPF_ERR_CAPACITY;
  if ((out->cap_vertices && !out->positions) || (out->cap_triangles && !out->indices)) return UPF_ERR_NULL;
  DGrid g;
  int st = to_device(grid, &g);
  if (st || (st = mc_grid_ok(g))) return st;
  const McLayout L = mc_layout(g);
  if (!d_ws || ws_bytes < L.bytes) return UPF_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  if ((st = mc_classify_scan(d_field, g, iso, d_ws, ws_bytes, s))) return st;
  // as upf_field_mesh_async: the persistent emit grid for the largest possible list, the
  // totals stored in d_counts by the kernel
  return mc_emit_launch(d_field, g, iso, d_ws, L.d.groups, out->positions, out->normals, out->indices,
                        out->cap_vertices, out->cap_triangles, s, d_counts);
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
