// upf_meshpost.hip -- mesh post-processing after marching cubes: keep the largest
// connected component (meshRemoveDisconnected, sources/meshGeneration.cpp:41) and the
// water-mesh vertex mask with its two-ring expansion (sources/meshGeneration.cpp:62-95).
//
// Connected components by label propagation over the triangles' edges: every vertex
// starts with its own id as label; each pass hooks the larger of two differing labels
// of an edge to the smaller (atomicMin) and then halves every label chain (pointer
// jumping), until a pass changes nothing.  The final label of a vertex is the smallest
// vertex id of its component, whatever the order the passes ran in.  The kept component
// is the one with the most triangles (ties: the smaller label).  Both operations end in
// one order-preserving compaction driven by the shared scan: vertex i keeps its order
// among the kept vertices, triangle t among the kept triangles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "upf_internal.hpp"
#include "upf_mc.hpp"

namespace upf {

struct PostLayout {
  int64_t nv, nt, n;  // n = max(nv, nt): scan length
  size_t off_label, off_cnt, off_a, off_b, off_counts, off_voff, off_toff, off_tiles, off_total, off_misc, bytes;
};

inline size_t al256p(size_t x) { return (x + 255) & ~(size_t)255; }

PostLayout post_layout(int64_t nv, int64_t nt) {
  PostLayout L;
  L.nv = nv;
  L.nt = nt;
  L.n = std::max<int64_t>(std::max(nv, nt), 1);
  size_t o = 0;
  L.off_label = o; o = al256p(o + 4 * (size_t)nv);
  L.off_cnt = o; o = al256p(o + 4 * (size_t)nv);
  L.off_a = o; o = al256p(o + (size_t)nv);
  L.off_b = o; o = al256p(o + (size_t)nv);
  L.off_counts = o; o = al256p(o + 4 * (size_t)L.n);
  L.off_voff = o; o = al256p(o + 4 * (size_t)L.n);
  L.off_toff = o; o = al256p(o + 4 * (size_t)L.n);
  L.off_tiles = o; o = al256p(o + sizeof(uint4) * (size_t)mc_scan_tiles(L.n));
  L.off_total = o; o = al256p(o + sizeof(uint4));
  L.off_misc = o; o = al256p(o + 16);  // best component key (u64), changed flag (u32)
  L.bytes = o;
  return L;
}

__global__ void k_pp_iota(int32_t* __restrict__ label, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) label[i] = (int32_t)i;
}

__global__ void k_pp_zero_u32(uint32_t* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__device__ __forceinline__ void hook(int32_t* label, int32_t a, int32_t b, uint32_t* changed) {
  const int32_t la = label[a], lb = label[b];
  if (la == lb) return;
  const int32_t hi = max(la, lb), lo = min(la, lb);
  if (atomicMin(label + hi, lo) > lo) *changed = 1u;
}

// Triangles naming a vertex outside [0, nv) are skipped by every kernel (and never kept).
__device__ __forceinline__ bool tri_ok(const uint32_t* idx, int64_t t, int64_t nv) {
  return idx[3 * t] < nv && idx[3 * t + 1] < nv && idx[3 * t + 2] < nv;
}

__global__ void k_pp_hook(const uint32_t* __restrict__ idx, int64_t nt, int64_t nv, int32_t* label,
                          uint32_t* changed) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nt || !tri_ok(idx, t, nv)) return;
  const int32_t a = (int32_t)idx[3 * t], b = (int32_t)idx[3 * t + 1], c = (int32_t)idx[3 * t + 2];
  hook(label, a, b, changed);
  hook(label, b, c, changed);
}

__global__ void k_pp_jump(int32_t* label, int64_t nv, uint32_t* changed) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  const int32_t l = label[v], ll = label[l];
  if (ll != l) {
    label[v] = ll;
    *changed = 1u;
  }
}

// Triangles per component label.  A mesh is mostly one component, so nearly every
// triangle adds to the same counter; one atomic per triangle would serialise on that
// address.  Each wave instead adds once per distinct label among its lanes: the lowest
// pending lane's label is broadcast, the lanes holding it are counted by ballot, and
// its owner adds the count (one round for a wave inside one component).
__global__ void k_pp_count(const uint32_t* __restrict__ idx, int64_t nt, int64_t nv,
                           const int32_t* __restrict__ label, uint32_t* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int32_t l = -1;
  if (t < nt && tri_ok(idx, t, nv)) l = label[idx[3 * t]];
  const int lane = (int)(threadIdx.x & 63);
  uint64_t todo = __ballot(l >= 0);
  while (todo) {
    const int lead = __builtin_ctzll(todo);
    const int32_t key = __shfl(l, lead);
    const uint64_t same = __ballot(l == key) & todo;
    if (lane == lead) atomicAdd(cnt + key, (uint32_t)__popcll(same));
    todo &= ~same;
  }
}

__global__ void k_pp_best(const uint32_t* __restrict__ cnt, int64_t nv, unsigned long long* best) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv || cnt[v] == 0u) return;
  atomicMax(best, ((unsigned long long)cnt[v] << 32) | (unsigned long long)(0xFFFFFFFFu - (uint32_t)v));
}

// keep flags in the kScanMc count format: vertex bit 0, triangle bit 8
__global__ void k_pp_flags_component(const uint32_t* __restrict__ idx, int64_t nv, int64_t nt,
                                     const int32_t* __restrict__ label, const unsigned long long* best,
                                     uint32_t* __restrict__ counts, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int32_t keep = (int32_t)(0xFFFFFFFFu - (uint32_t)(*best & 0xFFFFFFFFull));
  const uint32_t kv = (i < nv && label[i] == keep) ? 1u : 0u;
  const uint32_t kt = (i < nt && tri_ok(idx, i, nv) && label[idx[3 * i]] == keep) ? 1u : 0u;
  counts[i] = kv | (kt << 8);
}

__global__ void k_pp_water_valid(const DSpec s, const float* __restrict__ pos, int64_t nv, float threshold,
                                 uint8_t* __restrict__ valid) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v < nv) valid[v] = field_eval(s, pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]) < threshold ? 1 : 0;
}

__global__ void k_pp_copy_u8(const uint8_t* __restrict__ a, uint8_t* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// one expansion pass: every triangle touching a valid vertex (in `a`) makes its three
// vertices valid (in `b`, a copy of `a`); concurrent writers store the same value
__global__ void k_pp_expand(const uint32_t* __restrict__ idx, int64_t nt, int64_t nv, const uint8_t* __restrict__ a,
                            uint8_t* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nt || !tri_ok(idx, t, nv)) return;
  const uint32_t i0 = idx[3 * t], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
  if (a[i0] | a[i1] | a[i2]) {
    b[i0] = 1;
    b[i1] = 1;
    b[i2] = 1;
  }
}

__global__ void k_pp_flags_valid(const uint32_t* __restrict__ idx, int64_t nv, int64_t nt,
                                 const uint8_t* __restrict__ valid, uint32_t* __restrict__ counts, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t kv = (i < nv && valid[i]) ? 1u : 0u;
  const uint32_t kt =
      (i < nt && tri_ok(idx, i, nv) && valid[idx[3 * i]] && valid[idx[3 * i + 1]] && valid[idx[3 * i + 2]]) ? 1u : 0u;
  counts[i] = kv | (kt << 8);
}

__global__ void k_pp_compact(const float* __restrict__ pos, const float* __restrict__ nrm,
                             const uint32_t* __restrict__ idx, int64_t nv, int64_t nt,
                             const uint32_t* __restrict__ counts, const uint32_t* __restrict__ voff,
                             const uint32_t* __restrict__ toff, int64_t n, float* __restrict__ opos,
                             float* __restrict__ onrm, uint32_t* __restrict__ oidx) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = counts[i];
  if (c & 1u) {
    const uint32_t o = voff[i];
    opos[3 * (size_t)o] = pos[3 * i];
    opos[3 * (size_t)o + 1] = pos[3 * i + 1];
    opos[3 * (size_t)o + 2] = pos[3 * i + 2];
    if (onrm && nrm) {
      onrm[3 * (size_t)o] = nrm[3 * i];
      onrm[3 * (size_t)o + 1] = nrm[3 * i + 1];
      onrm[3 * (size_t)o + 2] = nrm[3 * i + 2];
    }
  }
  if ((c >> 8) & 1u) {
    const uint32_t o = toff[i];
    oidx[3 * (size_t)o] = voff[idx[3 * i]];
    oidx[3 * (size_t)o + 1] = voff[idx[3 * i + 1]];
    oidx[3 * (size_t)o + 2] = voff[idx[3 * i + 2]];
  }
}

inline dim3 blocks_for(int64_t n) { return dim3((unsigned)((std::max<int64_t>(n, 1) + 255) / 256)); }

int check_mesh(const upf_mesh_buffers* in, const upf_mesh_buffers* out) {
  if (!in || !out) return UPF_ERR_NULL;
  if (in->n_vertices < 0 || in->n_triangles < 0 || in->n_vertices >= ((int64_t)1 << 31) ||
      in->n_triangles >= ((int64_t)1 << 31))
    return UPF_ERR_BAD_GRID;
  if ((in->n_vertices && !in->positions) || (in->n_triangles && !in->indices)) return UPF_ERR_NULL;
  return UPF_OK;
}

// scan the keep flags, read the totals back, check capacity and compact
int finish(const PostLayout& L, char* w, const upf_mesh_buffers* in, upf_mesh_buffers* out, hipStream_t s) {
  uint32_t* counts = reinterpret_cast<uint32_t*>(w + L.off_counts);
  uint32_t* voff = reinterpret_cast<uint32_t*>(w + L.off_voff);
  uint32_t* toff = reinterpret_cast<uint32_t*>(w + L.off_toff);
  uint4* total = reinterpret_cast<uint4*>(w + L.off_total);
  int st = mc_scan_launch(counts, L.n, reinterpret_cast<uint4*>(w + L.off_tiles), total, voff, nullptr, toff, s,
                          kScanMc);
  if (st) return st;
  uint4 h;
  UPF_HIP_CHECK(hipMemcpyAsync(&h, total, sizeof(uint4), hipMemcpyDeviceToHost, s));
  UPF_HIP_CHECK(hipStreamSynchronize(s));
  out->n_vertices = h.x;
  out->n_triangles = h.y;
  if (out->n_vertices > out->cap_vertices || out->n_triangles > out->cap_triangles) return UPF_ERR_CAPACITY;
  if ((out->n_vertices && !out->positions) || (out->n_triangles && !out->indices)) return UPF_ERR_NULL;
  if (L.n > 0 && (h.x || h.y))
    hipLaunchKernelGGL(k_pp_compact, blocks_for(L.n), dim3(256), 0, s, in->positions, in->normals, in->indices,
                       L.nv, L.nt, counts, voff, toff, L.n, out->positions, out->normals, out->indices);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

}  // namespace upf

using namespace upf;

extern "C" {

size_t upf_mesh_post_workspace_bytes(int64_t n_vertices, int64_t n_triangles) {
  if (n_vertices < 0 || n_triangles < 0 || n_vertices >= ((int64_t)1 << 31) || n_triangles >= ((int64_t)1 << 31))
    return 0;
  return post_layout(n_vertices, n_triangles).bytes;
}

int upf_mesh_remove_disconnected(const upf_mesh_buffers* in, upf_mesh_buffers* out, void* d_ws, size_t ws_bytes,
                                 void* stream) {
  int st = check_mesh(in, out);
  if (st) return st;
  const PostLayout L = post_layout(in->n_vertices, in->n_triangles);
  if (!d_ws || ws_bytes < L.bytes) return UPF_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  char* w = static_cast<char*>(d_ws);
  int32_t* label = reinterpret_cast<int32_t*>(w + L.off_label);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(w + L.off_cnt);
  uint32_t* misc = reinterpret_cast<uint32_t*>(w + L.off_misc);
  unsigned long long* best = reinterpret_cast<unsigned long long*>(misc);
  uint32_t* changed = misc + 2;
  const int64_t nv = L.nv, nt = L.nt;
  if (nt == 0) {  // no triangles: nothing is kept
    out->n_vertices = out->n_triangles = 0;
    return UPF_OK;
  }
  hipLaunchKernelGGL(k_pp_iota, blocks_for(nv), dim3(256), 0, s, label, nv);
  // hook + jump until a pass changes nothing; every pass lowers a label or stops, so the
  // loop ends within nv passes (in practice a few tens)
  for (int64_t pass = 0;; pass++) {
    if (pass > nv + 1) return UPF_ERR_HIP;
    hipLaunchKernelGGL(k_pp_zero_u32, dim3(1), dim3(256), 0, s, changed, (int64_t)1);
    hipLaunchKernelGGL(k_pp_hook, blocks_for(nt), dim3(256), 0, s, in->indices, nt, nv, label, changed);
    hipLaunchKernelGGL(k_pp_jump, blocks_for(nv), dim3(256), 0, s, label, nv, changed);
    hipLaunchKernelGGL(k_pp_jump, blocks_for(nv), dim3(256), 0, s, label, nv, changed);
    UPF_LAUNCH_CHECK();
    uint32_t h = 0;
    UPF_HIP_CHECK(hipMemcpyAsync(&h, changed, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    UPF_HIP_CHECK(hipStreamSynchronize(s));
    if (!h) break;
  }
  hipLaunchKernelGGL(k_pp_zero_u32, blocks_for(nv), dim3(256), 0, s, cnt, nv);
  hipLaunchKernelGGL(k_pp_zero_u32, dim3(1), dim3(256), 0, s, misc, (int64_t)2);
  hipLaunchKernelGGL(k_pp_count, blocks_for(nt), dim3(256), 0, s, in->indices, nt, nv, label, cnt);
  hipLaunchKernelGGL(k_pp_best, blocks_for(nv), dim3(256), 0, s, cnt, nv, best);
  hipLaunchKernelGGL(k_pp_flags_component, blocks_for(L.n), dim3(256), 0, s, in->indices, nv, nt, label, best,
                     reinterpret_cast<uint32_t*>(w + L.off_counts), L.n);
  UPF_LAUNCH_CHECK();
  return finish(L, w, in, out, s);
}

int upf_mesh_water_mask(const upf_field_spec* spec, float threshold, const upf_mesh_buffers* in,
                        upf_mesh_buffers* out, void* d_ws, size_t ws_bytes, void* stream) {
  int st = check_mesh(in, out);
  if (st) return st;
  if (!spec) return UPF_ERR_NULL;
  if (!std::isfinite(threshold)) return UPF_ERR_INVALID_VALUE;
  upf_field_spec raw = *spec;
  raw.compose = UPF_COMPOSE_ELEVATION_RAW;
  DSpec ds;
  if ((st = to_device(&raw, &ds))) return st;
  const PostLayout L = post_layout(in->n_vertices, in->n_triangles);
  if (!d_ws || ws_bytes < L.bytes) return UPF_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  char* w = static_cast<char*>(d_ws);
  uint8_t* a = reinterpret_cast<uint8_t*>(w + L.off_a);
  uint8_t* b = reinterpret_cast<uint8_t*>(w + L.off_b);
  const int64_t nv = L.nv, nt = L.nt;
  if (nv == 0) {
    out->n_vertices = out->n_triangles = 0;
    return UPF_OK;
  }
  hipLaunchKernelGGL(k_pp_water_valid, blocks_for(nv), dim3(256), 0, s, ds, in->positions, nv, threshold, a);
  for (int pass = 0; pass < 2; pass++) {
    hipLaunchKernelGGL(k_pp_copy_u8, blocks_for(nv), dim3(256), 0, s, a, b, nv);
    if (nt) hipLaunchKernelGGL(k_pp_expand, blocks_for(nt), dim3(256), 0, s, in->indices, nt, nv, a, b);
    std::swap(a, b);
  }
  hipLaunchKernelGGL(k_pp_flags_valid, blocks_for(L.n), dim3(256), 0, s, in->indices, nv, nt, a,
                     reinterpret_cast<uint32_t*>(w + L.off_counts), L.n);
  UPF_LAUNCH_CHECK();
  return finish(L, w, in, out, s);
}

}  // extern "C"
