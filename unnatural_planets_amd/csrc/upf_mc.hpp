// upf_mc.hpp -- marching-cubes workspace layout and launch helpers shared by the
// single-grid entry points and the batched path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "upf_device.hpp"

namespace upf {


struct McDims {
  int nx, ny, nz, gpr;  // gpr = 64-node groups per row
  int64_t nxy;
  int64_t groups;
};

struct McLayout {
  McDims d;
  int64_t tiles;
  size_t off_S, off_counts, off_voff, off_act, off_tiles, off_total, bytes;
};

struct McPtrs {
  uint64_t* S;        // per group: inside bits
  uint32_t* counts;   // per group: packed vertex / triangle counts
  uint32_t* voff;     // per group: first vertex id
  uint2* act;         // active groups: (group id, first triangle id)
  uint4* tiles;       // scan tile sums
  uint4* total;       // (vertices, triangles, active groups)
};

McLayout mc_layout(const DGrid& g);
McPtrs mc_ptrs(const McLayout& L, void* ws);
int mc_grid_ok(const DGrid& g);
// classify + scan, asynchronous; totals stay in the workspace
int mc_classify_scan(const float* d_field, const DGrid& g, float iso, void* d_ws, size_t ws_bytes, hipStream_t s);
// classify from the sign masks already in the workspace (written by the field kernel) + scan
int mc_classify_signs_scan(const DGrid& g, void* d_ws, size_t ws_bytes, hipStream_t s);
// emit over at most max_active active groups (the kernel reads the true count)
int mc_emit_launch(const float* d_field, const DGrid& g, float iso, void* d_ws, int64_t max_active, float* pos,
                   float* nrm, uint32_t* idx, int64_t cap_v, int64_t cap_t, hipStream_t s,
                   int64_t* counts_out = nullptr);
// count-word formats of mc_scan_launch
constexpr int kScanMc = 0;      // vertices bits 0-7, triangles 8-16, active bit 17
constexpr int kScanBlocks = 1;  // vertices bits 0-11, triangles 12-23, active bit 24
// exclusive scan of packed counts (vertices, triangles, active) into voff / toff (dense
// offsets, each optional), act (compacted (index, first triangle) of the active entries,
// optional) and total; tiles_ready: the sums of the scan tiles are already in `tiles`
// (written by the kernel that produced the counts)
int mc_scan_launch(const uint32_t* counts, int64_t groups, uint4* tiles, uint4* total, uint32_t* voff, uint2* act,
                   uint32_t* toff, hipStream_t s, int fmt = kScanMc, bool tiles_ready = false);
int64_t mc_scan_tiles(int64_t groups);

}  // namespace upf
