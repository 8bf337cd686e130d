// upf_oracle_post.hpp -- CPU oracle for the mesh post-processing (TEST INFRASTRUCTURE ONLY).
//
// Restates what libupf documents at upf_mesh_remove_disconnected / upf_mesh_water_mask:
// meshRemoveDisconnected (sources/meshGeneration.cpp:41) as "keep the component with the
// most triangles, ties to the smaller vertex id", and the water vertex mask with its two
// expansion passes (sources/meshGeneration.cpp:62-95).  Outputs are keep flags; the
// compaction is order-preserving.
#pragma once

#include <cstdint>
#include <numeric>
#include <vector>

namespace upfo {

inline int64_t dsu_find(std::vector<int64_t>& p, int64_t v) {
  while (p[v] != v) {
    p[v] = p[p[v]];
    v = p[v];
  }
  return v;
}

inline bool tri_in(const uint32_t* idx, int64_t t, int64_t nv) {
  return idx[3 * t] < (uint64_t)nv && idx[3 * t + 1] < (uint64_t)nv && idx[3 * t + 2] < (uint64_t)nv;
}

// Union by smaller root, so every root is its component's smallest vertex id.
inline void remove_disconnected(int64_t nv, int64_t nt, const uint32_t* idx, uint8_t* keep_v, uint8_t* keep_t) {
  std::vector<int64_t> p(nv);
  std::iota(p.begin(), p.end(), 0);
  for (int64_t t = 0; t < nt; t++) {
    if (!tri_in(idx, t, nv)) continue;
    for (int e = 0; e < 2; e++) {
      const int64_t a = dsu_find(p, idx[3 * t + e]), b = dsu_find(p, idx[3 * t + e + 1]);
      if (a != b) p[std::max(a, b)] = std::min(a, b);
    }
  }
  std::vector<int64_t> cnt(nv, 0);
  for (int64_t t = 0; t < nt; t++)
    if (tri_in(idx, t, nv)) cnt[dsu_find(p, idx[3 * t])]++;
  int64_t best = -1;
  for (int64_t v = 0; v < nv; v++)
    if (cnt[v] > 0 && (best < 0 || cnt[v] > cnt[best])) best = v;
  for (int64_t v = 0; v < nv; v++) keep_v[v] = best >= 0 && dsu_find(p, v) == best;
  for (int64_t t = 0; t < nt; t++) keep_t[t] = best >= 0 && tri_in(idx, t, nv) && dsu_find(p, idx[3 * t]) == best;
}

inline void water_keep(int64_t nv, int64_t nt, const uint32_t* idx, const uint8_t* valid_in, uint8_t* keep_v,
                       uint8_t* keep_t) {
  std::vector<uint8_t> a(valid_in, valid_in + nv), b;
  for (int pass = 0; pass < 2; pass++) {
    b = a;
    for (int64_t t = 0; t < nt; t++) {
      if (!tri_in(idx, t, nv)) continue;
      const uint32_t i0 = idx[3 * t], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
      if (a[i0] || a[i1] || a[i2]) b[i0] = b[i1] = b[i2] = 1;
    }
    a.swap(b);
  }
  for (int64_t v = 0; v < nv; v++) keep_v[v] = a[v];
  for (int64_t t = 0; t < nt; t++)
    keep_t[t] = tri_in(idx, t, nv) && a[idx[3 * t]] && a[idx[3 * t + 1]] && a[idx[3 * t + 2]];
}

}  // namespace upfo
