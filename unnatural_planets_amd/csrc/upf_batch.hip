// upf_batch.hip -- many chunks of uneven sizes evaluated and meshed by one set of
// launches (no counterpart in the reference, which meshes one chunk per task,
// sources/generator.cpp:159).
//
// Work items, enumerated chunk after chunk:
//   field     64 x nodes of a pair of grid rows            (chunk, xb, j / 2, k), which
//             also stores the rows' sign words: one item per wave over all chunks' items
//   classify  one marching-cubes group per lane, from the sign words alone
//   emit      packed passes over the active groups of the concatenated scan, on a
//             fixed grid
// Small chunks share the machine with large ones and nothing waits on the host.  The
// marching-cubes groups of all chunks are concatenated, so a single scan numbers
// vertices and triangles batch-wide; per-chunk ranges come out of the same scan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "upf_device_row.hpp"
#include "upf_internal.hpp"
#include "upf_mc.hpp"

namespace upf {

#include "upf_mc_device.hpp"

struct BChunk {
  DSpec spec;
  DGrid grid;
  McDims md;
  float iso;
  int32_t gx;          // 64-node row segments per row
  int64_t node_off;    // first field value of the chunk
  int64_t group_off;   // first marching-cubes group of the chunk
  int64_t row_off;     // first field item (row pairs)
};

// chunk holding item r of a list whose per-chunk starts are at member `off`
template <int64_t BChunk::*OFF>
__device__ __forceinline__ int find_chunk(const BChunk* __restrict__ ch, int n, int64_t r) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ch[mid].*OFF <= r) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One field item: nodes xb * 64 + lane of rows 2 jp and 2 jp + 1 (when it exists) of
// layer k of chunk C, their values into the chunk's field and their sign words into S.
template <int VARIANT, bool EXT, bool COMPOSED>
__device__ __forceinline__ void batch_field_item(const RowCtx& ctx, const BChunk& C, int xb, int jp, int k,
                                                 float* __restrict__ field, uint64_t* __restrict__ S) {
  const int lane = threadIdx.x & 63;
  const int j = 2 * jp;
  const bool two = j + 1 < C.grid.n[1];
  const int ic = min(xb * 64 + lane, C.grid.n[0] - 1);
  const float x = node_coord(C.grid.o[0], C.grid.s[0], ic);
  const float y[2] = {node_coord(C.grid.o[1], C.grid.s[1], j),
                      node_coord(C.grid.o[1], C.grid.s[1], two ? j + 1 : j)};
  const float z = node_coord(C.grid.o[2], C.grid.s[2], k);
  float v[2];
  if (VARIANT == UPF_EVAL_PER_LANE) {
    // a chunk rounds as its own single-grid evaluation would: the exact form for an
    // elevation mode or a further shape, the fused form otherwise (the choice is
    // uniform across the wave: one chunk per row pair)
    if (EXT && spec_ext(C.spec)) {
      v[0] = field_eval<true, true>(C.spec, x, y[0], z);
      v[1] = two ? field_eval<true, true>(C.spec, x, y[1], z) : v[0];
    } else {
      v[0] = field_eval<EXT, false>(C.spec, x, y[0], z);
      v[1] = two ? field_eval<EXT, false>(C.spec, x, y[1], z) : v[0];
    }
  } else {
    row_field<2, COMPOSED>(ctx, C.spec, x, y, z, v);
  }
  // the node index recomputed from the context's lane: not live across the layers
  const int in = xb * 64 + (VARIANT == UPF_EVAL_PER_LANE ? lane : lane_of(ctx));
  const bool valid = in < C.grid.n[0];
  if (valid) {
    const int64_t o = C.node_off + ((int64_t)k * C.grid.n[1] + j) * C.grid.n[0] + in;
    field[o] = v[0];
    if (two) field[o + C.grid.n[0]] = v[1];
  }
  // the rows' sign words (bit = node exists and f < iso), as the single-grid kernels
  const uint64_t m0 = ballot64(valid && v[0] < C.iso), m1 = ballot64(valid && v[1] < C.iso);
  if (lane == 0) {
    const int64_t w = C.group_off + ((int64_t)k * C.grid.n[1] + j) * C.gx + xb;
    S[w] = m0;
    if (two) S[w + C.gx] = m1;
  }
}

// The field pass: one field item per wave over all chunks' items, as the single-grid
// kernels run.  A workgroup takes 4 consecutive items; blk_start[workgroup] holds its
// first item's chunk, segment, row pair and plane (16 bits each), from which a wave
// steps to its own item with carries.  An item in a later chunk than the first (chunks
// of tiny grids hold fewer than 4 items, empty ones none) walks on to its chunk and is
// decoded by division (two 32-bit divisions per item cost ~50 VALU instructions, ~9% of
// an item of the 64-chunk batch).  A
// persistent loop of blocks_per_cu workgroups per CU striding over spans of items ran
// the row form at 1.06-1.08 ms on the 17 192^3 chunks of the 64-chunk batch against
// 0.83 ms for the single-grid kernel over the same chunks one by one (20 B per lane
// spilled in the loop); this launch takes the 64-chunk batch's row-form pass from 1.33
// to 1.09 ms and its per-lane form from 2.52 to 2.23 ms.  Seven waves per SIMD as the
// single-grid row kernel, six for the per-lane elevation modes and further shapes.
template <int VARIANT, bool EXT, bool COMPOSED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EXT ? 6 : 7))) void k_batch_field_items(
    const BChunk* __restrict__ ch, int n_chunks, const uint64_t* __restrict__ blk_start, int64_t n_items,
    float* __restrict__ field, uint64_t* __restrict__ S) {
  __shared__ float4 s_slice[4][kRowSlice];
  __shared__ float4 s_gtab[16];
  __shared__ float4 s_oct[4][kRowOctTab];
  if (threadIdx.x < 16) s_gtab[threadIdx.x] = gradient_entry(threadIdx.x);
  __syncthreads();
  const int wave = uniform(threadIdx.x >> 6);
  const int64_t r = (int64_t)blockIdx.x * 4 + wave;
  if (r >= n_items) return;  // whole waves
  const uint64_t st = blk_start[blockIdx.x];
  int c = (int)(st & 0xFFFFu);
  int xb = (int)((st >> 16) & 0xFFFFu), jp = (int)((st >> 32) & 0xFFFFu), k = (int)(st >> 48);
  if (c + 1 < n_chunks && ch[c + 1].row_off <= r) {
    do c++; while (c + 1 < n_chunks && ch[c + 1].row_off <= r);
    c = uniform(c);
    const BChunk& C = ch[c];
    const uint32_t local = (uint32_t)(r - C.row_off);  // < 2^31: a chunk's groups fit 31 bits
    const uint32_t pairs = (uint32_t)(C.grid.n[1] + 1) / 2, rest = local / (uint32_t)C.gx;
    xb = (int)(local % (uint32_t)C.gx);
    jp = (int)(rest % pairs);
    k = (int)(rest / pairs);
  } else {
    const BChunk& C = ch[c];
    const int pairs = (C.grid.n[1] + 1) / 2;
    for (int w = 0; w < wave; w++)
      if (++xb == C.gx) {
        xb = 0;
        if (++jp == pairs) {
          jp = 0;
          k++;
        }
      }
  }
  c = uniform(c);
  const RowCtx ctx = make_row_ctx(s_slice[wave], s_gtab, s_oct[wave], threadIdx.x & 63);
  batch_field_item<VARIANT, EXT, COMPOSED>(ctx, ch[c], uniform(xb), uniform(jp), uniform(k), field, S);
}

// Counts of every group of every chunk (global id g: chunk, then its local group) from
// the sign words; lanes of one wave may belong to two chunks.  One group per thread: the
// batch's small chunks have several times the 512^3 grid's mixed cells per group, and
// their classification passes need the occupancy (four or eight groups per thread with
// the next group's words loaded ahead and the tile sums written, as the single-grid
// kernel does, took 67 and 82 us against 50 + 5.6 us for this kernel and the separate
// tile pass).
__global__ __launch_bounds__(256) void k_batch_classify(const BChunk* __restrict__ ch, int n_chunks, int64_t groups,
                                                        const uint64_t* __restrict__ S,
                                                        uint32_t* __restrict__ counts) {
  __shared__ uint8_t s_ntris[256];
  __shared__ SignPass s_pass[4];
  s_ntris[threadIdx.x] = upf_mc_ntris[threadIdx.x];
  __syncthreads();
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool in = g < groups;
  // the chunk of the workgroup's first group (one search per workgroup), then a short
  // walk per lane: a workgroup's 256 groups rarely cross a chunk boundary
  int c = uniform(find_chunk<&BChunk::group_off>(ch, n_chunks, (int64_t)blockIdx.x * 256));
  while (c + 1 < n_chunks && ch[c + 1].group_off <= (in ? g : groups - 1)) c++;
  const BChunk& C = ch[c];
  const GroupWords w = load_group_words(S + C.group_off, C.md, g - C.group_off, in);
  const uint32_t cnt = classify_group_words(w, C.md.nx, in, s_ntris, s_pass[threadIdx.x >> 6]);
  if (in) counts[g] = cnt;
}

// Per-chunk vertex / triangle ranges from the dense exclusive offsets.
__global__ void k_batch_meshes(const BChunk* __restrict__ ch, int n_chunks, const uint32_t* __restrict__ voff,
                               const uint32_t* __restrict__ toff, const uint4* __restrict__ total,
                               upf_chunk_mesh* __restrict__ meshes) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_chunks) return;
  const uint32_t v0 = voff[ch[c].group_off], t0 = toff[ch[c].group_off];
  const uint32_t v1 = c + 1 < n_chunks ? voff[ch[c + 1].group_off] : total->x;
  const uint32_t t1 = c + 1 < n_chunks ? toff[ch[c + 1].group_off] : total->y;
  meshes[c].vertex_offset = v0;
  meshes[c].n_vertices = (int64_t)v1 - v0;
  meshes[c].triangle_offset = t0;
  meshes[c].n_triangles = (int64_t)t1 - t0;
}

// At least five waves per SIMD, as k_mc_emit (the same pass).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_batch_emit(
                                                    const BChunk* __restrict__ ch, int n_chunks,
                                                    const float* __restrict__ field, const uint64_t* __restrict__ S,
                                                    const uint32_t* __restrict__ voff, const uint2* __restrict__ act,
                                                    const uint4* __restrict__ total, float* __restrict__ pos,
                                                    float* __restrict__ nrm, uint32_t* __restrict__ idx, int64_t cap_v,
                                                    int64_t cap_t) {
  __shared__ EmitPass s_pass[4];
  const int wave = uniform(threadIdx.x >> 6);
  // contiguous runs of the active-group list per wave, in passes of one chunk each
  const int64_t n_active = (int64_t)total->z, nw = (int64_t)gridDim.x * 4;
  const int64_t per = (n_active + nw - 1) / nw;
  int64_t a = ((int64_t)blockIdx.x * 4 + wave) * per;
  const int64_t a1 = min(a + per, n_active);
  int c = a < a1 ? uniform(find_chunk<&BChunk::group_off>(ch, n_chunks, (int64_t)(uint32_t)uniform((int)act[a].x))) : 0;
  while (a < a1) {
    // the list is sorted, so the chunk only advances
    const int64_t g = (int64_t)(uint32_t)uniform((int)act[a].x);
    while (c + 1 < n_chunks && ch[c + 1].group_off <= g) c++;
    c = uniform(c);
    const BChunk& C = ch[c];
    const uint32_t vbase = voff[C.group_off];
    a += mc_emit_pass(field + C.node_off, C.md, C.grid, C.iso, S + C.group_off, voff + C.group_off, act, a,
                      (int)min<int64_t>(kEmitPassGroups, a1 - a), (uint32_t)C.group_off,
                      (uint32_t)(C.group_off + C.md.groups), vbase, s_pass[wave], pos, nrm, idx, cap_v, cap_t);
  }
}

// ------------------------------------------------------------------ host plan
namespace {

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

}  // namespace upf

struct upf_batch {
  std::vector<upf::BChunk> host;
  upf::BChunk* dev = nullptr;
  uint64_t* blk_start = nullptr;  // first field item of each workgroup of 4: chunk | xb << 16 | jp << 32 | k << 48
  int64_t item_blocks = 0;
  int device = 0;
  int n = 0;
  int cus = 256;
  int emit_blocks = 256 * 32;  // k_batch_emit's grid
  int64_t nodes = 0, groups = 0, rows = 0;
  bool any_ext = false;  // a chunk with an elevation mode or a further shape: the field runs per lane
  bool any_composed = false;  // a chunk whose shape is a composition (spec_composed)
  upf_batch_opts opts{};
  size_t off_S = 0, off_counts = 0, off_voff = 0, off_toff = 0, off_act = 0, off_tiles = 0, off_total = 0,
         ws_bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace upf {
namespace {

int normalize_opts(const upf_batch_opts* in, upf_batch_opts* out) {
  upf_batch_opts o{};
  if (in) o = *in;
  if (o.field_variant < 0 || o.field_variant > UPF_EVAL_PER_LANE) return UPF_ERR_BAD_SPEC;
  if (o.blocks_per_cu == 0) o.blocks_per_cu = 4;
  if (o.blocks_per_cu < 1 || o.blocks_per_cu > 32) return UPF_ERR_BAD_SPEC;
  if (o.classify_z == 0) o.classify_z = 32;
  if (o.classify_z != 16 && o.classify_z != 32 && o.classify_z != 64) return UPF_ERR_BAD_SPEC;
  o.reserved = 0;
  *out = o;
  return UPF_OK;
}

int upload(upf_batch* b) {
  UPF_HIP_CHECK(hipMemcpy(b->dev, b->host.data(), sizeof(BChunk) * b->n, hipMemcpyHostToDevice));
  return UPF_OK;
}

int run(upf_batch* b, float* field, void* ws, size_t ws_bytes, float* pos, float* nrm, uint32_t* idx, int64_t cap_v,
        int64_t cap_t, upf_chunk_mesh* meshes, hipStream_t s) {
  if (!field || !ws || !meshes || (cap_v > 0 && !pos) || (cap_t > 0 && !idx)) return UPF_ERR_NULL;
  if (ws_bytes < b->ws_bytes) return UPF_ERR_WORKSPACE;
  char* w = static_cast<char*>(ws);
  uint64_t* S = reinterpret_cast<uint64_t*>(w + b->off_S);
  uint32_t* counts = reinterpret_cast<uint32_t*>(w + b->off_counts);
  uint32_t* voff = reinterpret_cast<uint32_t*>(w + b->off_voff);
  uint32_t* toff = reinterpret_cast<uint32_t*>(w + b->off_toff);
  uint2* act = reinterpret_cast<uint2*>(w + b->off_act);
  uint4* tiles = reinterpret_cast<uint4*>(w + b->off_tiles);
  uint4* total = reinterpret_cast<uint4*>(w + b->off_total);
  const dim3 items((unsigned)b->item_blocks);
  if (b->any_ext)
    hipLaunchKernelGGL((k_batch_field_items<UPF_EVAL_PER_LANE, true, true>), items, dim3(256), 0, s, b->dev, b->n,
                       b->blk_start, b->rows, field, S);
  else if (b->opts.field_variant == UPF_EVAL_PER_LANE)
    hipLaunchKernelGGL((k_batch_field_items<UPF_EVAL_PER_LANE, false, true>), items, dim3(256), 0, s, b->dev, b->n,
                       b->blk_start, b->rows, field, S);
  else if (b->any_composed)
    hipLaunchKernelGGL((k_batch_field_items<UPF_EVAL_ROWS, false, true>), items, dim3(256), 0, s, b->dev, b->n,
                       b->blk_start, b->rows, field, S);
  else
    hipLaunchKernelGGL((k_batch_field_items<UPF_EVAL_ROWS, false, false>), items, dim3(256), 0, s, b->dev, b->n,
                       b->blk_start, b->rows, field, S);
  hipLaunchKernelGGL(k_batch_classify, dim3((unsigned)((b->groups + 255) / 256)), dim3(256), 0, s, b->dev, b->n,
                     b->groups, S, counts);
  UPF_LAUNCH_CHECK();
  int st = mc_scan_launch(counts, b->groups, tiles, total, voff, act, toff, s);
  if (st) return st;
  hipLaunchKernelGGL(k_batch_meshes, dim3((b->n + 63) / 64), dim3(64), 0, s, b->dev, b->n, voff, toff, total, meshes);
  // the emission runs at fewer waves per SIMD than the field pass (its registers), so it
  // gets its own grid: about 32 small workgroups per CU in whole rounds of its resident
  // ones, several rounds of short runs, instead of one and a fraction of the field pass's
  hipLaunchKernelGGL(k_batch_emit, dim3((unsigned)b->emit_blocks), dim3(256), 0, s, b->dev, b->n, field, S, voff,
                     act, total, pos, nrm, idx, cap_v, cap_t);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

}  // namespace
}  // namespace upf

using namespace upf;

extern "C" {

int upf_batch_create(const upf_chunk* chunks, int32_t n_chunks, const upf_batch_opts* opts, upf_batch** out) {
  if (!out || (!chunks && n_chunks > 0)) return UPF_ERR_NULL;
  *out = nullptr;
  if (n_chunks < 1 || n_chunks > 65536) return UPF_ERR_BAD_SPEC;
  upf_batch* b = new (std::nothrow) upf_batch();
  if (!b) return UPF_ERR_HIP;
  int st = normalize_opts(opts, &b->opts);
  if (st) { delete b; return st; }
  b->n = n_chunks;
  b->host.resize(n_chunks);
  for (int c = 0; c < n_chunks; c++) {
    BChunk& C = b->host[c];
    std::memset(&C, 0, sizeof(C));
    if ((st = to_device(&chunks[c].spec, &C.spec)) || (st = to_device(&chunks[c].grid, &C.grid)) ||
        (st = mc_grid_ok(C.grid))) {
      delete b;
      return st;
    }
    if (!std::isfinite(chunks[c].iso)) { delete b; return UPF_ERR_INVALID_VALUE; }
    C.md = mc_layout(C.grid).d;
    C.iso = chunks[c].iso;
    if (spec_ext(C.spec)) b->any_ext = true;
    if (spec_composed(C.spec)) b->any_composed = true;
    C.gx = (C.grid.n[0] + 63) / 64;
    C.node_off = b->nodes;
    C.group_off = b->groups;
    C.row_off = b->rows;
    b->nodes += (int64_t)C.grid.n[0] * C.grid.n[1] * C.grid.n[2];
    b->groups += C.md.groups;
    b->rows += (int64_t)C.gx * ((C.grid.n[1] + 1) / 2) * C.grid.n[2];
  }
  if (b->groups >= ((int64_t)1 << 31)) { delete b; return UPF_ERR_BAD_GRID; }
  size_t o = 0;
  b->off_S = o; o = al256(o + sizeof(uint64_t) * b->groups);
  b->off_counts = o; o = al256(o + sizeof(uint32_t) * b->groups);
  b->off_voff = o; o = al256(o + sizeof(uint32_t) * b->groups);
  b->off_toff = o; o = al256(o + sizeof(uint32_t) * b->groups);
  b->off_act = o; o = al256(o + sizeof(uint2) * b->groups);
  b->off_tiles = o; o = al256(o + sizeof(uint4) * mc_scan_tiles(b->groups));
  b->off_total = o; o = al256(o + sizeof(uint4));
  b->ws_bytes = o;
  if (hipGetDevice(&b->device) != hipSuccess ||
      hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess ||
      hipMalloc(&b->dev, sizeof(BChunk) * n_chunks) != hipSuccess) {
    delete b;
    return UPF_ERR_HIP;
  }
  {
    // each workgroup's first field item (chunk, segment, row pair, plane; all < 2^16: at
    // most 65536 chunks of at most 65535 nodes per axis), for the field kernel
    b->item_blocks = (b->rows + 3) / 4;
    if (b->item_blocks >= ((int64_t)1 << 31)) { upf_batch_destroy(b); return UPF_ERR_BAD_GRID; }
    std::vector<uint64_t> tab((size_t)b->item_blocks);
    int c = 0;
    for (int64_t q = 0; q < b->item_blocks; q++) {
      while (c + 1 < n_chunks && b->host[c + 1].row_off <= 4 * q) c++;
      const BChunk& C = b->host[c];
      const int64_t local = 4 * q - C.row_off, pairs = (C.grid.n[1] + 1) / 2;
      const uint64_t xb = (uint64_t)(local % C.gx), rest = (uint64_t)(local / C.gx);
      tab[(size_t)q] = (uint64_t)c | xb << 16 | (rest % pairs) << 32 | (rest / pairs) << 48;
    }
    if (hipMalloc(&b->blk_start, sizeof(uint64_t) * tab.size()) != hipSuccess ||
        hipMemcpy(b->blk_start, tab.data(), sizeof(uint64_t) * tab.size(), hipMemcpyHostToDevice) != hipSuccess) {
      upf_batch_destroy(b);
      return UPF_ERR_HIP;
    }
  }
  {
    const int per_cu = std::max(1, resident_blocks(reinterpret_cast<const void*>(&k_batch_emit), 256) / b->cus);
    b->emit_blocks = b->cus * per_cu * std::max(1, 32 / per_cu);
  }
  if ((st = upload(b))) { upf_batch_destroy(b); return st; }
  *out = b;
  return UPF_OK;
}

void upf_batch_destroy(upf_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->dev) (void)hipFree(b->dev);
  if (b->blk_start) (void)hipFree(b->blk_start);
  delete b;
}

int64_t upf_batch_field_floats(const upf_batch* b) { return b ? b->nodes : 0; }
int64_t upf_batch_field_offset(const upf_batch* b, int32_t chunk) {
  return (b && chunk >= 0 && chunk < b->n) ? b->host[chunk].node_off : -1;
}
size_t upf_batch_workspace_bytes(const upf_batch* b) { return b ? b->ws_bytes : 0; }

int upf_batch_set_opts(upf_batch* b, const upf_batch_opts* opts) {
  if (!b) return UPF_ERR_NULL;
  upf_batch_opts o;
  const int st = normalize_opts(opts, &o);
  if (st) return st;
  // a graph captured with other options would replay the old launch shapes: drop it, so
  // that upf_batch_replay reports UPF_ERR_NULL until upf_batch_capture runs again
  if (o.field_variant != b->opts.field_variant || o.blocks_per_cu != b->opts.blocks_per_cu ||
      o.classify_z != b->opts.classify_z) {
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  }
  b->opts = o;  // the options are host-side launch parameters: nothing to upload
  return UPF_OK;
}

int upf_batch_get_opts(const upf_batch* b, upf_batch_opts* opts) {
  if (!b || !opts) return UPF_ERR_NULL;
  *opts = b->opts;
  return UPF_OK;
}

int upf_batch_run(upf_batch* b, float* d_field, void* d_ws, size_t ws_bytes, float* d_pos, float* d_nrm,
                  uint32_t* d_idx, int64_t cap_v, int64_t cap_t, upf_chunk_mesh* d_meshes, void* stream) {
  if (!b) return UPF_ERR_NULL;
  return run(b, d_field, d_ws, ws_bytes, d_pos, d_nrm, d_idx, cap_v, cap_t, d_meshes, as_stream(stream));
}

int upf_batch_totals(const upf_batch* b, const void* d_ws, int64_t* n_vertices, int64_t* n_triangles, void* stream) {
  if (!b || !d_ws || !n_vertices || !n_triangles) return UPF_ERR_NULL;
  uint4 t;
  hipStream_t s = as_stream(stream);
  UPF_HIP_CHECK(hipMemcpyAsync(&t, static_cast<const char*>(d_ws) + b->off_total, sizeof(t), hipMemcpyDeviceToHost, s));
  UPF_HIP_CHECK(hipStreamSynchronize(s));
  *n_vertices = t.x;
  *n_triangles = t.y;
  return UPF_OK;
}

int upf_batch_capture(upf_batch* b, float* d_field, void* d_ws, size_t ws_bytes, float* d_pos, float* d_nrm,
                      uint32_t* d_idx, int64_t cap_v, int64_t cap_t, upf_chunk_mesh* d_meshes, void* stream) {
  if (!b) return UPF_ERR_NULL;
  hipStream_t s = as_stream(stream);
  if (!s) return UPF_ERR_NULL;  // capture needs a non-default stream
  if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  UPF_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int st = run(b, d_field, d_ws, ws_bytes, d_pos, d_nrm, d_idx, cap_v, cap_t, d_meshes, s);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (st) { if (g) (void)hipGraphDestroy(g); return st; }
  if (e != hipSuccess) return UPF_ERR_HIP;
  b->graph = g;
  UPF_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return UPF_OK;
}

int upf_batch_replay(upf_batch* b, void* stream) {
  if (!b || !b->exec) return UPF_ERR_NULL;
  UPF_HIP_CHECK(hipGraphLaunch(b->exec, as_stream(stream)));
  return UPF_OK;
}

int upf_batch_autotune(upf_batch* b, float* d_field, void* d_ws, size_t ws_bytes, float* d_pos, float* d_nrm,
                       uint32_t* d_idx, int64_t cap_v, int64_t cap_t, upf_chunk_mesh* d_meshes, void* stream,
                       upf_batch_opts* best) {
  if (!b) return UPF_ERR_NULL;
  hipStream_t s = as_stream(stream);
  hipEvent_t e0, e1;
  UPF_HIP_CHECK(hipEventCreate(&e0));
  UPF_HIP_CHECK(hipEventCreate(&e1));
  const upf_batch_opts start = b->opts;
  upf_batch_opts best_o = start;
  float best_ms = 1e30f;
  int st = UPF_OK;
  // the row form and the per-lane form (blocks_per_cu has no effect on either)
  for (int cand = 0; cand < 2 && !st; cand++) {
    const upf_batch_opts o{cand == 0 ? UPF_EVAL_ROWS : UPF_EVAL_PER_LANE, start.blocks_per_cu, start.classify_z, 0};
    if ((st = upf_batch_set_opts(b, &o))) break;
    if ((st = run(b, d_field, d_ws, ws_bytes, d_pos, d_nrm, d_idx, cap_v, cap_t, d_meshes, s))) break;  // warm-up
    if (hipEventRecord(e0, s) != hipSuccess) { st = UPF_ERR_HIP; break; }
    for (int r = 0; r < 2 && !st; r++) st = run(b, d_field, d_ws, ws_bytes, d_pos, d_nrm, d_idx, cap_v, cap_t, d_meshes, s);
    if (st) break;
    float ms = 0.0f;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      st = UPF_ERR_HIP;
      break;
    }
    if (ms < best_ms) { best_ms = ms; best_o = o; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const int st2 = upf_batch_set_opts(b, st ? &start : &best_o);
  if (st) return st;
  if (best) *best = best_o;
  return st2;
}

}  // extern "C"
