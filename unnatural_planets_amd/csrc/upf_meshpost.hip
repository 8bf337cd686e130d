// upf_meshpost.hip -- mesh post-processing after marching cubes: keep the largest
// connected component (meshRemoveDisconnected, sources/meshGeneration.cpp:41) and the
// water-mesh vertex mask with its two-ring expansion (sources/meshGeneration.cpp:62-95).
//
// Connected components by a lock-free union-find over the triangles' edges (the
// initialisation and compare-and-swap linking of ECL-CC, Jaiganesh and Burtscher, HPDC
// 2018): every vertex starts pointing at the smallest vertex of its triangles; an edge
// finds the roots of its two ends (halving the paths it walks) and links the larger
// root under the smaller one by compare-and-swap, retrying from the new roots when
// another edge linked first.  A label only ever points to a
// smaller id, so a root is the smallest vertex id of its tree.  A flattening pass then
// points every vertex at its root.  Passes repeat until one links nothing and flattens
// nothing (the first pass normally does all the work; the second confirms it), so the
// final label of a vertex is the smallest vertex id of its component, whatever order the
// edges ran in.  The kept component is the one with the most triangles (ties: the
// smaller label).  Both operations end in one order-preserving compaction driven by the
// shared scan: vertex i keeps its order among the kept vertices, triangle t among the
// kept triangles.
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
  L.off_misc = o; o = al256p(o + 16);  // best component key (u64), two changed flags (u32)
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

// Labels change under other workgroups during a union-find pass.  The walks read them
// through the caches (workgroup-scope relaxed accesses) and may see an older value: a
// label only ever moves to an ancestor, so an older value is still an ancestor (or the
// vertex itself while it was a root), and a compare-and-swap against an outdated root
// fails and returns the current one.  Agent-scope loads, which skip the caches, made
// the first pass over the 1.33 M-triangle torus mesh 1.7x slower (and a walk of the
// 2.6 MB label array without the caches' locality, lanes striding over distant
// triangles, 4.7x slower).
__device__ __forceinline__ int32_t ld_label(const int32_t* label, int32_t v) {
  return __hip_atomic_load(label + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void st_label(int32_t* label, int32_t v, int32_t x) {
  __hip_atomic_store(label + v, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// What the atomics wrote (labels, triangle counts, the best key) is read at agent scope
// wherever a stale copy would change the result: cached loads, even in a later kernel,
// have returned a label from before a compare-and-swap of an earlier pass, which left
// a vertex pointing at a non-root after a flattening pass that saw nothing to change
// (1 run in 20 on a 1 M-vertex soup).  The walks of k_pp_unite tolerate such values;
// the flattening, which decides that the labels are final, and the kernels after it do
// not.
template <typename T>
__device__ __forceinline__ T ld_fresh(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_fresh(int32_t* p, int32_t x) {
  __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of v's tree, halving the path on the way: each visited vertex is pointed at its
// grandparent, an ancestor, so no tree changes.  Every step goes to a smaller id.  Only
// vertices that are no longer roots are written (during a pass a root is written by
// uf_unite's compare-and-swap alone, and a vertex never becomes a root again).
__device__ __forceinline__ int32_t uf_root(int32_t* label, int32_t v) {
  int32_t p = ld_label(label, v);
  while (p != v) {
    const int32_t g = ld_label(label, p);
    if (g == p) return p;
    st_label(label, v, g);
    v = g;
    p = ld_label(label, v);
  }
  return v;
}

// Joins the trees of a and b; true when this call linked two roots.  Every failed
// compare-and-swap lowers ra + rb (the failing root is replaced by its new root, a
// smaller id), so the loop ends.  Neighbouring lanes take neighbouring triangles, which
// share vertices and so often the same two roots: a lane whose previous lane holds the
// same pair leaves the link to it (that lane's loop ends only once the two trees are
// one).  60% of the compare-and-swaps of the first pass over a marching-cubes mesh
// failed before.  Must be called by the whole wave's active lanes.
__device__ __forceinline__ bool uf_unite(int32_t* label, int32_t a, int32_t b) {
  int32_t ra = uf_root(label, a), rb = uf_root(label, b);
  {
    const int32_t hi = max(ra, rb), lo = min(ra, rb);
    const int lane = (int)(threadIdx.x & 63);
    const int prev = (lane + 63) & 63;
    const int32_t phi = __shfl(hi, prev), plo = __shfl(lo, prev);
    const uint64_t act = __ballot(1);
    if (lane > 0 && ((act >> prev) & 1ull) && phi == hi && plo == lo) return false;
  }
  while (ra != rb) {
    if (ra < rb) {
      const int32_t t = ra;
      ra = rb;
      rb = t;
    }
    const int32_t old = atomicCAS(label + ra, ra, rb);
    if (old == ra) return true;
    ra = uf_root(label, old);
    rb = uf_root(label, rb);
  }
  return false;
}

// Triangles naming a vertex outside [0, nv) are skipped by every kernel (and never kept).
__device__ __forceinline__ bool tri_ok(const uint32_t* idx, int64_t t, int64_t nv) {
  return idx[3 * t] < nv && idx[3 * t + 1] < nv && idx[3 * t + 2] < nv;
}

// Edges a-b and b-c join all three vertices of a triangle.
__global__ void k_pp_unite(const uint32_t* __restrict__ idx, int64_t nt, int64_t nv, int32_t* label,
                           uint32_t* changed) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nt || !tri_ok(idx, t, nv)) return;
  const int32_t a = (int32_t)idx[3 * t], b = (int32_t)idx[3 * t + 1], c = (int32_t)idx[3 * t + 2];
  const bool ab = uf_unite(label, a, b);
  const bool bc = uf_unite(label, b, c);
  if (ab || bc) *changed = 1u;
}

// Before the first union-find pass, every vertex points at the smallest vertex of the
// triangles it belongs to (ECL-CC's initialisation, by atomicMin): a forest, since every
// label is a smaller id, whose stars already join most of a marching-cubes mesh's
// neighbourhoods.  With it and uf_unite's shared links, a call over the 1.33 M-triangle
// torus mesh takes 0.52-0.53 ms instead of 0.65 (a soup of random triangles over random
// vertex ids, whose neighbouring triangles share no vertices, got slower: 1.5 -> 1.7-1.8
// ms).
__global__ void k_pp_init_min(const uint32_t* __restrict__ idx, int64_t nt, int64_t nv, int32_t* label) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nt || !tri_ok(idx, t, nv)) return;
  const int32_t a = (int32_t)idx[3 * t], b = (int32_t)idx[3 * t + 1], c = (int32_t)idx[3 * t + 2];
  const int32_t m = min(a, min(b, c));
  if (a != m) atomicMin(label + a, m);
  if (b != m) atomicMin(label + b, m);
  if (c != m) atomicMin(label + c, m);
}

// Points every vertex at its root (no links run meanwhile, so the roots stay put).
__global__ void k_pp_flatten(int32_t* label, int64_t nv, uint32_t* changed) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  const int32_t l = ld_fresh(label + v);
  int32_t r = l, p;
  while ((p = ld_fresh(label + r)) != r) r = p;
  if (r != l) {
    st_fresh(label + v, r);
    *changed = 1u;
  }
}

// Triangles per component label.  A mesh is mostly one component, so nearly every
// triangle adds to the same counter, and atomics on one address serialise (one per
// wave of 64 triangles took 238 us for 1.33 M triangles).  A wave counts 1024
// consecutive triangles as runs of equal labels: the distinct labels of each 64 are
// found by ballot (the lowest pending lane's label is broadcast, the lanes holding it
// counted), a run is added to its counter when the label changes, and the four waves of
// a workgroup merge their last runs, so a one-component mesh adds once per workgroup.
constexpr int kCountRun = 16;  // triangles per lane
constexpr int64_t kCountTris = 4 * 64 * kCountRun;  // triangles per workgroup

__global__ void __launch_bounds__(256) k_pp_count(const uint32_t* __restrict__ idx, int64_t nt, int64_t nv,
                                                  const int32_t* __restrict__ label,
                                                  uint32_t* __restrict__ cnt) {
  const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
  const int64_t t0 = ((int64_t)blockIdx.x * 4 + wave) * (64 * kCountRun) + lane;
  int32_t l[kCountRun];
#pragma unroll
  for (int j = 0; j < kCountRun; j++) {
    const int64_t t = t0 + 64 * j;
    l[j] = (t < nt && tri_ok(idx, t, nv)) ? (int32_t)idx[3 * t] : -1;
  }
#pragma unroll
  for (int j = 0; j < kCountRun; j++)
    if (l[j] >= 0) l[j] = ld_fresh(label + l[j]);
  int32_t cur = -1;  // the open run: label and count (wave-uniform)
  uint32_t run = 0;
#pragma unroll
  for (int j = 0; j < kCountRun; j++) {
    uint64_t todo = __ballot(l[j] >= 0);
    while (todo) {
      const int lead = __builtin_ctzll(todo);
      const int32_t key = __shfl(l[j], lead);
      const uint64_t same = __ballot(l[j] == key) & todo;
      const uint32_t n = (uint32_t)__popcll(same);
      if (key == cur) {
        run += n;
      } else {
        if (run && lane == 0) atomicAdd(cnt + cur, run);
        cur = key;
        run = n;
      }
      todo &= ~same;
    }
  }
  __shared__ int32_t s_key[4];
  __shared__ uint32_t s_run[4];
  if (lane == 0) {
    s_key[wave] = cur;
    s_run[wave] = run;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t k = s_key[0];
    uint32_t r = s_run[0];
    for (int w = 1; w < 4; w++) {
      if (s_key[w] == k) {
        r += s_run[w];
      } else {
        if (r) atomicAdd(cnt + k, r);
        k = s_key[w];
        r = s_run[w];
      }
    }
    if (r) atomicAdd(cnt + k, r);
  }
}

__global__ void k_pp_best(const uint32_t* __restrict__ cnt, int64_t nv, unsigned long long* best) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  const uint32_t c = ld_fresh(cnt + v);
  if (c) atomicMax(best, ((unsigned long long)c << 32) | (unsigned long long)(0xFFFFFFFFu - (uint32_t)v));
}

// keep flags in the kScanMc count format: vertex bit 0, triangle bit 8
__global__ void k_pp_flags_component(const uint32_t* __restrict__ idx, int64_t nv, int64_t nt,
                                     const int32_t* __restrict__ label, const unsigned long long* best,
                                     uint32_t* __restrict__ counts, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int32_t keep = (int32_t)(0xFFFFFFFFu - (uint32_t)(ld_fresh(best) & 0xFFFFFFFFull));
  const uint32_t kv = (i < nv && ld_fresh(label + i) == keep) ? 1u : 0u;
  const uint32_t kt = (i < nt && tri_ok(idx, i, nv) && ld_fresh(label + idx[3 * i]) == keep) ? 1u : 0u;
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
                             const uint32_t* __restrict__ toff, int64_t n, const uint4* __restrict__ total,
                             int64_t cap_v, int64_t cap_t, float* __restrict__ opos, float* __restrict__ onrm,
                             uint32_t* __restrict__ oidx) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  // The kept counts are checked against the output here, before the host reads them
  // back: an output too small (or missing) for them is left untouched.
  const uint4 tot = *total;
  if (tot.x > cap_v || tot.y > cap_t || (tot.x && !opos) || (tot.y && !oidx)) return;
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

// scan the keep flags, compact (the kernel checks capacity), read the totals back
int finish(const PostLayout& L, char* w, const upf_mesh_buffers* in, upf_mesh_buffers* out, hipStream_t s) {
  uint32_t* counts = reinterpret_cast<uint32_t*>(w + L.off_counts);
  uint32_t* voff = reinterpret_cast<uint32_t*>(w + L.off_voff);
  uint32_t* toff = reinterpret_cast<uint32_t*>(w + L.off_toff);
  uint4* total = reinterpret_cast<uint4*>(w + L.off_total);
  int st = mc_scan_launch(counts, L.n, reinterpret_cast<uint4*>(w + L.off_tiles), total, voff, nullptr, toff, s,
                          kScanMc);
  if (st) return st;
  if (L.n > 0)
    hipLaunchKernelGGL(k_pp_compact, blocks_for(L.n), dim3(256), 0, s, in->positions, in->normals, in->indices,
                       L.nv, L.nt, counts, voff, toff, L.n, total, out->cap_vertices, out->cap_triangles,
                       out->positions, out->normals, out->indices);
  UPF_LAUNCH_CHECK();
  uint4 h;
  UPF_HIP_CHECK(hipMemcpyAsync(&h, total, sizeof(uint4), hipMemcpyDeviceToHost, s));
  UPF_HIP_CHECK(hipStreamSynchronize(s));
  out->n_vertices = h.x;
  out->n_triangles = h.y;
  if (out->n_vertices > out->cap_vertices || out->n_triangles > out->cap_triangles) return UPF_ERR_CAPACITY;
  if ((out->n_vertices && !out->positions) || (out->n_triangles && !out->indices)) return UPF_ERR_NULL;
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
  uint32_t* changed = misc + 2;  // two flags
  const int64_t nv = L.nv, nt = L.nt;
  if (nt == 0) {  // no triangles: nothing is kept
    out->n_vertices = out->n_triangles = 0;
    return UPF_OK;
  }
  hipLaunchKernelGGL(k_pp_iota, blocks_for(nv), dim3(256), 0, s, label, nv);
  hipLaunchKernelGGL(k_pp_init_min, blocks_for(nt), dim3(256), 0, s, in->indices, nt, nv, label);
  // Two union-find passes per read-back, each with its own changed flag; the loop ends
  // when the second pass of a round changed nothing.  A round that goes on linked two
  // roots in its second pass (after a pass without links the flattening leaves nothing
  // to change), so there are at most nv rounds; the first pass normally joins every
  // component and the second confirms it.
  for (int64_t round = 0;; round++) {
    if (round > nv + 1) return UPF_ERR_HIP;
    hipLaunchKernelGGL(k_pp_zero_u32, dim3(1), dim3(256), 0, s, changed, (int64_t)2);
    for (int p = 0; p < 2; p++) {
      hipLaunchKernelGGL(k_pp_unite, blocks_for(nt), dim3(256), 0, s, in->indices, nt, nv, label, changed + p);
      hipLaunchKernelGGL(k_pp_flatten, blocks_for(nv), dim3(256), 0, s, label, nv, changed + p);
    }
    UPF_LAUNCH_CHECK();
    uint32_t h = 0;
    UPF_HIP_CHECK(hipMemcpyAsync(&h, changed + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    UPF_HIP_CHECK(hipStreamSynchronize(s));
    if (!h) break;
  }
  hipLaunchKernelGGL(k_pp_zero_u32, blocks_for(nv), dim3(256), 0, s, cnt, nv);
  hipLaunchKernelGGL(k_pp_zero_u32, dim3(1), dim3(256), 0, s, misc, (int64_t)2);
  hipLaunchKernelGGL(k_pp_count, dim3((unsigned)((nt + kCountTris - 1) / kCountTris)), dim3(256), 0, s,
                     in->indices, nt, nv, label, cnt);
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
