// upf_nb.hip -- narrow-band marching cubes: coarse pass over block corners, fine
// evaluation and meshing of the active blocks only.
//
// The reference sweeps the dense grid (sources/meshGeneration.cpp:35-39); a planet's
// surface touches a few percent of the cells of a 1024^3 grid, so this path evaluates the
// field at the block corners first (every 8th node per axis, 1/512 of the nodes), keeps
// the blocks whose corners come within `band` of the iso value or straddle it, and
// evaluates and meshes only those.
//
//  coarse    one thread per block corner;
//  activity  one thread per block: a count word with the active bit; the shared scan
//            compacts the active blocks, and a dense map block -> active rank follows;
//  fine      one workgroup per active block: its 9^3 brick of fine nodes into LDS and
//            HBM, then per node the crossing edges it owns (a node owns its +x/+y/+z
//            edges; nodes on the block's upper faces belong to the next block except on
//            the grid's upper boundary) and their exclusive prefix, packed with the
//            node's crossing bits into one uint16 per node; per block the vertex and
//            triangle counts;
//  scan      vertex and triangle bases per active block;
//  emit      one workgroup per active block: vertices at base + prefix + axis rank,
//            and triangles whose edge ids come from the owning block's node table (this
//            block's, or a neighbour's through the dense map);
//  normals   one thread per vertex: the normalized analytic gradient at the vertex.
// Vertex positions use the same interpolation as the dense path on the same per-lane
// node values, so the mesh equals the dense one up to order when the band holds the
// surface.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "upf_device_grad.hpp"
#include "upf_device_row.hpp"
#include "upf_internal.hpp"
#include "upf_mc.hpp"

namespace upf {

#include "upf_mc_device.hpp"

constexpr int kB = UPF_NB_BLOCK;          // cells per block edge
constexpr int kBN = kB + 1;               // nodes per brick edge
constexpr int kBrick = kBN * kBN * kBN;   // 729 nodes
constexpr int kCells = kB * kB * kB;      // 512 cells
constexpr int kNbThreads = 256;
static_assert(kCells == 2 * kNbThreads, "fine/emit give each thread two cells");
static_assert(kBrick <= 3 * kNbThreads, "fine/emit give each thread at most three brick nodes");

struct NbDims {
  int bx, by, bz;  // blocks per axis
  int pad;
  int64_t blocks;
};

__device__ __forceinline__ void block_coords(const NbDims& d, uint32_t b, int& bx, int& by, int& bz) {
  bx = (int)(b % (uint32_t)d.bx);
  const uint32_t r = b / (uint32_t)d.bx;
  by = (int)(r % (uint32_t)d.by);
  bz = (int)(r / (uint32_t)d.by);
}

// Workgroup-wide exclusive sum of one int per thread; *total gets the sum.
__device__ __forceinline__ int wg_exclusive_sum(int v, int* total) {
  __shared__ int s_w[kNbThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int inc = wave_inclusive_sum(v);
  if (lane == 63) s_w[wave] = inc;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kNbThreads / 64; w++) {
    if (w < wave) before += s_w[w];
    all += s_w[w];
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

template <bool EXT>
__global__ __launch_bounds__(256) void k_nb_coarse(const DSpec s, const DGrid g, const NbDims d,
                                                   float* __restrict__ coarse) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int cx = d.bx + 1, cy = d.by + 1;
  const int64_t n = (int64_t)cx * cy * (d.bz + 1);
  if (c >= n) return;
  const int I = (int)(c % cx);
  const int64_t r = c / cx;
  const int J = (int)(r % cy), K = (int)(r / cy);
  coarse[c] = field_eval<EXT>(s, node_coord(g.o[0], g.s[0], I * kB), node_coord(g.o[1], g.s[1], J * kB),
                         node_coord(g.o[2], g.s[2], K * kB));
}

__global__ __launch_bounds__(256) void k_nb_activity(const float* __restrict__ coarse, const NbDims d, float iso,
                                                     float band, uint32_t* __restrict__ counts) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= d.blocks) return;
  int bx, by, bz;
  block_coords(d, (uint32_t)b, bx, by, bz);
  const int cx = d.bx + 1;
  const int64_t cxy = (int64_t)cx * (d.by + 1);
  const int64_t c0 = (int64_t)bz * cxy + (int64_t)by * cx + bx;
  const float v[8] = {coarse[c0], coarse[c0 + 1], coarse[c0 + cx], coarse[c0 + cx + 1],
                      coarse[c0 + cxy], coarse[c0 + cxy + 1], coarse[c0 + cxy + cx], coarse[c0 + cxy + cx + 1]};
  int inside = 0;
  bool near = false;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    inside += v[q] < iso;
    near |= fabsf(v[q] - iso) < band;
  }
  const bool active = near || (inside != 0 && inside != 8);
  counts[b] = active ? (1u << 17) : 0u;  // the kScanMc active bit
}

__global__ void k_nb_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_nb_map(const uint2* __restrict__ act, int64_t nact, int32_t* __restrict__ map) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r < nact) map[act[r].x] = (int32_t)r;
}

template <bool EXT, bool COMPOSED = true>
__global__ __launch_bounds__(kNbThreads) void k_nb_fine(const DSpec s, const DGrid g, const NbDims d, float iso,
                                                        const uint2* __restrict__ act,
                                                        const int32_t* __restrict__ map, float* __restrict__ bricks,
                                                        uint16_t* __restrict__ tabs, uint32_t* __restrict__ bcounts,
                                                        uint32_t* __restrict__ unresolved, int oct_max) {
  __shared__ float s_fe[EXT ? kBrick : 1];
  __shared__ float4 s_slice[4][EXT ? 1 : kRowSlice];
  __shared__ float4 s_gtab[16];
  // the (row, octave) tables of the 4 waves, 2 * kBrickRows * kBrickGroups * oct_max float4
  // each (launch-sized: the spec's largest octave count sets the LDS, hence the occupancy)
  extern __shared__ float4 s_oct[];
  __shared__ uint8_t s_ntris[256];  // triangles per case code
  s_ntris[threadIdx.x] = upf_mc_ntris[threadIdx.x];  // read after the barriers below
  const int64_t a = blockIdx.x;
  const uint32_t b = act[a].x;
  int bx, by, bz;
  block_coords(d, b, bx, by, bz);
  const int gi0 = bx * kB, gj0 = by * kB, gk0 = bz * kB;
  const bool lx = bx == d.bx - 1, ly = by == d.by - 1, lz = bz == d.bz - 1;
  float* __restrict__ brick = bricks + a * kBrick;
  // the brick's values in LDS; the row path stores them over its slices once every wave
  // has finished with its slice (LDS sets the occupancy here)
  float* s_f = EXT ? s_fe : reinterpret_cast<float*>(&s_slice[0][0]);
  static_assert(sizeof(float) * kBrick <= sizeof(float4) * 4 * kRowSlice, "the brick fits the slices");
  if (EXT) {  // elevation modes and further shapes: per node
    for (int n = threadIdx.x; n < kBrick; n += kNbThreads) {
      const int k = n / (kBN * kBN), r = n - k * kBN * kBN;
      const int j = r / kBN, i = r - j * kBN;
      const float v = field_eval<EXT>(s, node_coord(g.o[0], g.s[0], gi0 + i), node_coord(g.o[1], g.s[1], gj0 + j),
                                      node_coord(g.o[2], g.s[2], gk0 + k));
      s_f[n] = v;
      brick[n] = v;
    }
  } else {
    // the brick's 81 rows of 9 nodes (row = j + 9 k) in 12 groups of kBrickRows; wave w
    // takes groups w, w + 4, w + 8 (kBrickGroups of them), its row rho = 7 m + r being
    // brick row 7 (w + 4 m) + r (the 3 rows past 80 repeat row 80 and are not stored);
    // lane 7 * 9 = 63 stands in for node 8 of row 6
    static_assert(kBrickRows * kBrickGroups * 4 >= kBN * kBN, "four waves cover the brick's rows");
    if (threadIdx.x < 16) s_gtab[threadIdx.x] = gradient_entry(threadIdx.x);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const RowCtx ctx = make_row_ctx(s_slice[wave], s_gtab, s_oct + wave * (2 * kBrickRows * kBrickGroups * oct_max),
                                    lane);
    const bool node_lane = lane < kBrickRows * kBN;
    const int rl = node_lane ? lane / kBN : kBrickRows - 1;
    const int i = node_lane ? lane - rl * kBN : kBN - 1;
    const float x = node_coord(g.o[0], g.s[0], gi0 + i);
    auto brick_row = [&](int rho) {
      const int m = rho / kBrickRows;
      return min(kBrickRows * (wave + 4 * m) + (rho - m * kBrickRows), kBN * kBN - 1);
    };
    auto yz = [&](int rho, float& yy, float& zz) {
      const int br = brick_row(rho), kk = br / kBN, jj = br - kk * kBN;
      yy = node_coord(g.o[1], g.s[1], gj0 + jj);
      zz = node_coord(g.o[2], g.s[2], gk0 + kk);
    };
    float y[kBrickGroups], z[kBrickGroups], v[kBrickGroups];
#pragma unroll
    for (int m = 0; m < kBrickGroups; m++) yz(kBrickRows * m + rl, y[m], z[m]);
    brick_field<kBrickGroups, COMPOSED>(ctx, s, rl, yz, x, y, z, v);
    __syncthreads();  // every wave done with its slice
#pragma unroll
    for (int m = 0; m < kBrickGroups; m++) {
      const int br = kBrickRows * (wave + 4 * m) + rl;
      if (node_lane && br < kBN * kBN) {
        const int n = br * kBN + i;
        s_f[n] = v[m];
        brick[n] = v[m];
      }
    }
  }
  __syncthreads();

  // nodes 3t .. 3t+2: crossing bits, owned vertex counts
  int cnt[3], cross[3], tot = 0;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int n = 3 * threadIdx.x + q;
    cnt[q] = cross[q] = 0;
    if (n >= kBrick) continue;
    const int k = n / (kBN * kBN), r = n - k * kBN * kBN;
    const int j = r / kBN, i = r - j * kBN;
    const bool in0 = s_f[n] < iso;
    int c = 0;
    if (i < kB && (s_f[n + 1] < iso) != in0) c |= 1;
    if (j < kB && (s_f[n + kBN] < iso) != in0) c |= 2;
    if (k < kB && (s_f[n + kBN * kBN] < iso) != in0) c |= 4;
    const int ox = (i == kB && !lx), oy = (j == kB && !ly), oz = (k == kB && !lz);
    cross[q] = c;
    if (!(ox | oy | oz)) {
      cnt[q] = __popc(c);
    } else if (c) {
      const int64_t nb = (int64_t)b + ox + (int64_t)oy * d.bx + (int64_t)oz * d.bx * d.by;
      if (map[nb] < 0) atomicAdd(unresolved, 1u);
    }
    tot += cnt[q];
  }
  int nv;
  int pre = wg_exclusive_sum(tot, &nv);
  uint16_t* __restrict__ tab = tabs + a * kBrick;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int n = 3 * threadIdx.x + q;
    if (n < kBrick) tab[n] = (uint16_t)(pre | (cross[q] << 12));
    pre += cnt[q];
  }

  // cells 2t, 2t+1
  int nt = 0;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int c = 2 * threadIdx.x + q;
    const int i = c & 7, j = (c >> 3) & 7, k = c >> 6;
    const int n = (k * kBN + j) * kBN + i;
    const int code = (s_f[n] < iso) | ((s_f[n + 1] < iso) << 1) | ((s_f[n + kBN + 1] < iso) << 2) |
                     ((s_f[n + kBN] < iso) << 3) | ((s_f[n + kBN * kBN] < iso) << 4) |
                     ((s_f[n + kBN * kBN + 1] < iso) << 5) | ((s_f[n + kBN * kBN + kBN + 1] < iso) << 6) |
                     ((s_f[n + kBN * kBN + kBN] < iso) << 7);
    nt += s_ntris[code];
  }
  int ntot;
  (void)wg_exclusive_sum(nt, &ntot);
  if (threadIdx.x == 0) bcounts[a] = (uint32_t)nv | ((uint32_t)ntot << 12) | (1u << 24);
}

__global__ __launch_bounds__(kNbThreads) void k_nb_emit(const DSpec s, const DGrid g, const NbDims d, float iso,
                                                        const uint2* __restrict__ act,
                                                        const int32_t* __restrict__ map,
                                                        const float* __restrict__ bricks,
                                                        const uint16_t* __restrict__ tabs,
                                                        const uint32_t* __restrict__ voff,
                                                        const uint32_t* __restrict__ toff, float* __restrict__ pos,
                                                        float* __restrict__ nrm, uint32_t* __restrict__ idx,
                                                        uint64_t* __restrict__ keys, int64_t cap_v, int64_t cap_t) {
  __shared__ float s_f[kBrick];
  __shared__ uint16_t s_tab[kBrick];
  __shared__ uint8_t s_ntris[256];  // triangles per case code (per-lane table reads stay in LDS)
  __shared__ int32_t s_nrow[8];     // upper neighbour (dx | dy << 1 | dz << 2): its active row, or -1
  __shared__ uint32_t s_nvoff[8];   // and its first vertex id
  const int64_t a = blockIdx.x;
  const uint32_t b = act[a].x;
  int bx, by, bz;
  block_coords(d, b, bx, by, bz);
  const int gi0 = bx * kB, gj0 = by * kB, gk0 = bz * kB;
  const bool lx = bx == d.bx - 1, ly = by == d.by - 1, lz = bz == d.bz - 1;
  static_assert(kNbThreads == 256, "one case code per thread");
  s_ntris[threadIdx.x] = upf_mc_ntris[threadIdx.x];
  // threads 1..7 look up the upper neighbours that own the vertices of the brick's upper
  // faces: the rank read overlaps the brick's reads, the neighbour's first vertex id the
  // vertex section; both are published before the triangle section's barrier
  const int no = (int)threadIdx.x;
  const bool nlook = no >= 1 && no < 8 && !((no & 1) && lx) && !((no & 2) && ly) && !((no & 4) && lz);
  const int32_t nrow =
      nlook ? map[(int64_t)b + (no & 1) + (int64_t)((no >> 1) & 1) * d.bx + (int64_t)(no >> 2) * d.bx * d.by] : -1;
  for (int n = threadIdx.x; n < kBrick; n += kNbThreads) {
    s_f[n] = bricks[a * kBrick + n];
    s_tab[n] = tabs[a * kBrick + n];
  }
  __syncthreads();
  const uint32_t vb = voff[a];
  const uint32_t nvoff = nrow >= 0 ? voff[nrow] : 0u;

  // ---- vertices of the owned nodes 3t .. 3t+2 (normals follow in k_nb_normals)
  for (int q = 0; q < 3; q++) {
    const int n = 3 * threadIdx.x + q;
    if (n >= kBrick) break;
    const int k = n / (kBN * kBN), r = n - k * kBN * kBN;
    const int j = r / kBN, i = r - j * kBN;
    const bool owned = (i < kB || lx) && (j < kB || ly) && (k < kB || lz);
    const int entry = s_tab[n];
    const int c = entry >> 12;
    if (!owned || !c) continue;
    uint32_t vid = vb + (uint32_t)(entry & 0xFFF);
    const int gi[3] = {gi0 + i, gj0 + j, gk0 + k};
    const float p0[3] = {node_coord(g.o[0], g.s[0], gi[0]), node_coord(g.o[1], g.s[1], gi[1]),
                         node_coord(g.o[2], g.s[2], gi[2])};
    const float f0 = s_f[n];
    const int step[3] = {1, kBN, kBN * kBN};
    const int64_t node = ((int64_t)gi[2] * g.n[1] + gi[1]) * g.n[0] + gi[0];
    for (int ax = 0; ax < 3; ax++) {
      if (!((c >> ax) & 1)) continue;
      const float f1 = s_f[n + step[ax]];
      const float t = (iso - f0) / (f1 - f0);
      const float p1a = node_coord(g.o[ax], g.s[ax], gi[ax] + 1);
      float p[3] = {p0[0], p0[1], p0[2]};
      p[ax] = p0[ax] + t * (p1a - p0[ax]);
      if ((int64_t)vid < cap_v) {
        pos[3 * (size_t)vid + 0] = p[0];
        pos[3 * (size_t)vid + 1] = p[1];
        pos[3 * (size_t)vid + 2] = p[2];
        if (keys) keys[vid] = (uint64_t)node * 3u + (uint64_t)ax;
      }
      vid++;
    }
  }

  if (no >= 1 && no < 8) {
    s_nrow[no] = nrow;
    s_nvoff[no] = nvoff;
  }

  // ---- triangles of cells 2t, 2t+1 (wg_exclusive_sum's barrier publishes s_nrow)
  int code[2], nt = 0;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int cc = 2 * threadIdx.x + q;
    const int i = cc & 7, j = (cc >> 3) & 7, k = cc >> 6;
    const int n = (k * kBN + j) * kBN + i;
    code[q] = (s_f[n] < iso) | ((s_f[n + 1] < iso) << 1) | ((s_f[n + kBN + 1] < iso) << 2) |
              ((s_f[n + kBN] < iso) << 3) | ((s_f[n + kBN * kBN] < iso) << 4) |
              ((s_f[n + kBN * kBN + 1] < iso) << 5) | ((s_f[n + kBN * kBN + kBN + 1] < iso) << 6) |
              ((s_f[n + kBN * kBN + kBN] < iso) << 7);
    nt += s_ntris[code[q]];
  }
  int ntot;
  uint32_t tid = toff[a] + (uint32_t)wg_exclusive_sum(nt, &ntot);
  // the triangles' corners: the case's triangulation row in one 16-byte load, the edge
  // owners from a 64-bit immediate (per-element reads of the constant tables at per-lane
  // indices were a chain of dependent memory reads per triangle corner: 528 -> 213 us
  // for this kernel without the whole section)
  for (int q = 0; q < 2; q++) {
    const int cc = 2 * threadIdx.x + q;
    const int i = cc & 7, j = (cc >> 3) & 7, k = cc >> 6;
    const int ntc = s_ntris[code[q]];
    if (!ntc) continue;
    const int4 row = *reinterpret_cast<const int4*>(&c_tris16.v[code[q]][0]);
    const uint32_t rw[4] = {(uint32_t)row.x, (uint32_t)row.y, (uint32_t)row.z, (uint32_t)row.w};
#pragma unroll
    for (int tr = 0; tr < UPF_MC_MAX_TRIS; tr++, tid++) {
      if (tr >= ntc) break;
#pragma unroll
      for (int v = 0; v < 3; v++) {
        const int qq = 3 * tr + v;
        const int e = (int)((rw[qq >> 2] >> (8 * (qq & 3))) & 0xFu);  // 0..11 on a triangle row
        const int own = (int)((kEdgeOwner >> (5 * e)) & 31u);
        const int li = i + (own & 1), lj = j + ((own >> 1) & 1), lk = k + ((own >> 2) & 1);
        const int axis = own >> 3;
        const int ox = (li == kB && !lx), oy = (lj == kB && !ly), oz = (lk == kB && !lz);
        const int nn = ((lk - kB * oz) * kBN + (lj - kB * oy)) * kBN + (li - kB * ox);
        uint32_t id = 0xFFFFFFFFu;
        if (!(ox | oy | oz)) {
          const int entry = s_tab[nn];
          id = vb + (uint32_t)(entry & 0xFFF) + (uint32_t)__popc((entry >> 12) & ((1 << axis) - 1));
        } else {
          const int o = ox | (oy << 1) | (oz << 2);
          const int32_t rnb = s_nrow[o];
          if (rnb >= 0) {
            const int entry = tabs[(int64_t)rnb * kBrick + nn];
            id = s_nvoff[o] + (uint32_t)(entry & 0xFFF) + (uint32_t)__popc((entry >> 12) & ((1 << axis) - 1));
          }
        }
        if ((int64_t)tid < cap_t) idx[3 * (size_t)tid + v] = id;
      }
    }
  }
}

// Normals: the normalized analytic gradient at each vertex, one thread per vertex (few
// nodes of a brick own a vertex, so evaluating the gradient inside k_nb_emit would run
// the whole gradient for mostly idle lanes).
template <bool EXT>
__global__ __launch_bounds__(256) void k_nb_normals(const DSpec s, const float* __restrict__ pos, int64_t n,
                                                    float* __restrict__ nrm) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= n) return;
  ThetaAcc none{nullptr};
  float gr[3];
  (void)field_grad<false, EXT>(s, pos[3 * v], pos[3 * v + 1], pos[3 * v + 2], 0.0f, gr, none);
  const float len = __builtin_sqrtf(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
  const float inv = len > 0.0f ? 1.0f / len : 0.0f;
  nrm[3 * v + 0] = gr[0] * inv;
  nrm[3 * v + 1] = gr[1] * inv;
  nrm[3 * v + 2] = gr[2] * inv;
}

// ------------------------------------------------------------------ host side
struct NbLayout {
  NbDims d;
  int64_t coarse, max_active;
  size_t off_coarse, off_counts, off_tiles, off_total_a, off_act, off_map, off_bricks, off_tabs, off_bcounts,
      off_voff, off_toff, off_total_b, bytes;
};

inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

NbLayout nb_layout(const NbDims& d, int64_t max_active) {
  NbLayout L;
  L.d = d;
  L.coarse = (int64_t)(d.bx + 1) * (d.by + 1) * (d.bz + 1);
  L.max_active = max_active;
  const int64_t tiles = mc_scan_tiles(std::max<int64_t>(d.blocks, std::max<int64_t>(max_active, 1)));
  size_t o = 0;
  L.off_coarse = o; o = al256(o + 4 * (size_t)L.coarse);
  L.off_counts = o; o = al256(o + 4 * (size_t)d.blocks);
  L.off_tiles = o; o = al256(o + sizeof(uint4) * (size_t)tiles);
  L.off_total_a = o; o = al256(o + sizeof(uint4));
  L.off_act = o; o = al256(o + sizeof(uint2) * (size_t)d.blocks);
  L.off_map = o; o = al256(o + 4 * (size_t)d.blocks);
  L.off_bricks = o; o = al256(o + 4 * (size_t)kBrick * max_active);
  L.off_tabs = o; o = al256(o + 2 * (size_t)kBrick * max_active);
  L.off_bcounts = o; o = al256(o + 4 * (size_t)max_active);
  L.off_voff = o; o = al256(o + 4 * (size_t)max_active);
  L.off_toff = o; o = al256(o + 4 * (size_t)max_active);
  L.off_total_b = o; o = al256(o + 2 * sizeof(uint4));  // totals + unresolved counter
  L.bytes = o;
  return L;
}

int nb_dims(const DGrid& g, NbDims* d) {
  for (int a = 0; a < 3; a++)
    if ((g.n[a] - 1) % kB) return UPF_ERR_BAD_GRID;
  d->bx = (g.n[0] - 1) / kB;
  d->by = (g.n[1] - 1) / kB;
  d->bz = (g.n[2] - 1) / kB;
  d->pad = 0;
  d->blocks = (int64_t)d->bx * d->by * d->bz;
  if (d->blocks >= ((int64_t)1 << 31)) return UPF_ERR_BAD_GRID;
  return UPF_OK;
}

}  // namespace upf

using namespace upf;

extern "C" {

size_t upf_nb_workspace_bytes(const upf_grid* grid, int64_t max_active) {
  DGrid g;
  NbDims d;
  if (to_device(grid, &g) || nb_dims(g, &d) || max_active < 0) return 0;
  return nb_layout(d, std::min<int64_t>(max_active, d.blocks)).bytes;
}

int upf_nb_extract(const upf_field_spec* spec, const upf_grid* grid, float iso, float band, upf_mesh_buffers* out,
                   uint64_t* d_keys, void* d_ws, size_t ws_bytes, upf_nb_stats* stats, void* stream) {
  if (!out || !stats) return UPF_ERR_NULL;
  if (!std::isfinite(iso) || !std::isfinite(band) || band < 0.0f) return UPF_ERR_INVALID_VALUE;
  DSpec s;
  DGrid g;
  NbDims d;
  int st = to_device(spec, &s);
  if (st || (st = to_device(grid, &g)) || (st = nb_dims(g, &d))) return st;
  *stats = upf_nb_stats{};
  stats->blocks = d.blocks;
  // the largest active-block capacity this workspace holds
  const NbLayout L0 = nb_layout(d, 0);
  if (!d_ws || ws_bytes < L0.bytes) return UPF_ERR_WORKSPACE;
  const size_t per = 6 * (size_t)kBrick + 12;
  int64_t cap = std::min<int64_t>(d.blocks, (int64_t)((ws_bytes - L0.bytes) / per));
  while (cap > 0 && nb_layout(d, cap).bytes > ws_bytes) cap--;
  const NbLayout L = nb_layout(d, cap);
  char* w = static_cast<char*>(d_ws);
  float* coarse = reinterpret_cast<float*>(w + L.off_coarse);
  uint32_t* counts = reinterpret_cast<uint32_t*>(w + L.off_counts);
  uint4* tiles = reinterpret_cast<uint4*>(w + L.off_tiles);
  uint4* total_a = reinterpret_cast<uint4*>(w + L.off_total_a);
  uint2* act = reinterpret_cast<uint2*>(w + L.off_act);
  int32_t* map = reinterpret_cast<int32_t*>(w + L.off_map);
  float* bricks = reinterpret_cast<float*>(w + L.off_bricks);
  uint16_t* tabs = reinterpret_cast<uint16_t*>(w + L.off_tabs);
  uint32_t* bcounts = reinterpret_cast<uint32_t*>(w + L.off_bcounts);
  uint32_t* voff = reinterpret_cast<uint32_t*>(w + L.off_voff);
  uint32_t* toff = reinterpret_cast<uint32_t*>(w + L.off_toff);
  uint4* total_b = reinterpret_cast<uint4*>(w + L.off_total_b);
  hipStream_t hs = as_stream(stream);

  const auto nblk = [](int64_t n) { return dim3((unsigned)((n + 255) / 256)); };
  const bool modes = spec_ext(s);
  if (modes) hipLaunchKernelGGL(k_nb_coarse<true>, nblk(L.coarse), dim3(256), 0, hs, s, g, d, coarse);
  else hipLaunchKernelGGL(k_nb_coarse<false>, nblk(L.coarse), dim3(256), 0, hs, s, g, d, coarse);
  hipLaunchKernelGGL(k_nb_activity, nblk(d.blocks), dim3(256), 0, hs, coarse, d, iso, band, counts);
  UPF_LAUNCH_CHECK();
  if ((st = mc_scan_launch(counts, d.blocks, tiles, total_a, nullptr, act, nullptr, hs, kScanMc))) return st;
  hipLaunchKernelGGL(k_nb_fill_i32, nblk(d.blocks), dim3(256), 0, hs, map, d.blocks, -1);
  hipLaunchKernelGGL(k_nb_fill_i32, dim3(1), dim3(256), 0, hs, reinterpret_cast<int32_t*>(total_b), (int64_t)8, 0);
  UPF_LAUNCH_CHECK();
  uint4 ha;
  UPF_HIP_CHECK(hipMemcpyAsync(&ha, total_a, sizeof(uint4), hipMemcpyDeviceToHost, hs));
  UPF_HIP_CHECK(hipStreamSynchronize(hs));
  const int64_t nact = ha.z;
  stats->active_blocks = nact;
  if (nact > cap) return UPF_ERR_WORKSPACE;
  out->n_vertices = out->n_triangles = 0;
  if (nact == 0) return UPF_OK;

  hipLaunchKernelGGL(k_nb_map, nblk(nact), dim3(256), 0, hs, act, nact, map);
  if (modes)
    hipLaunchKernelGGL(k_nb_fine<true>, dim3((unsigned)nact), dim3(kNbThreads), 0, hs, s, g, d, iso, act, map, bricks,
                       tabs, bcounts, reinterpret_cast<uint32_t*>(total_b + 1), 0);
  else {
    int oct_max = 1;  // octaves of the largest gradient / value layer (the brick rows' table)
    for (int l = 0; l < s.n_layers; l++)
      if (s.L[l].noise == UPF_NOISE_GRADIENT || s.L[l].noise == UPF_NOISE_VALUE)
        oct_max = std::max(oct_max, s.L[l].fractal == UPF_FRACTAL_NONE ? 1 : s.L[l].octaves);
    const size_t lds = sizeof(float4) * 4 * 2 * kBrickRows * kBrickGroups * oct_max;
    if (spec_composed(s))
      hipLaunchKernelGGL((k_nb_fine<false, true>), dim3((unsigned)nact), dim3(kNbThreads), lds, hs, s, g, d, iso, act,
                         map, bricks, tabs, bcounts, reinterpret_cast<uint32_t*>(total_b + 1), oct_max);
    else
      hipLaunchKernelGGL((k_nb_fine<false, false>), dim3((unsigned)nact), dim3(kNbThreads), lds, hs, s, g, d, iso, act,
                         map, bricks, tabs, bcounts, reinterpret_cast<uint32_t*>(total_b + 1), oct_max);
  }
  UPF_LAUNCH_CHECK();
  if ((st = mc_scan_launch(bcounts, nact, tiles, total_b, voff, nullptr, toff, hs, kScanBlocks))) return st;
  uint4 hb[2];
  UPF_HIP_CHECK(hipMemcpyAsync(hb, total_b, 2 * sizeof(uint4), hipMemcpyDeviceToHost, hs));
  UPF_HIP_CHECK(hipStreamSynchronize(hs));
  stats->n_vertices = hb[0].x;
  stats->n_triangles = hb[0].y;
  stats->unresolved = hb[1].x;
  out->n_vertices = hb[0].x;
  out->n_triangles = hb[0].y;
  if (stats->unresolved) return UPF_ERR_BAND;
  if (out->n_vertices > out->cap_vertices || out->n_triangles > out->cap_triangles) return UPF_ERR_CAPACITY;
  if (out->n_vertices == 0 && out->n_triangles == 0) return UPF_OK;
  if (!out->positions || !out->indices) return UPF_ERR_NULL;
  hipLaunchKernelGGL(k_nb_emit, dim3((unsigned)nact), dim3(kNbThreads), 0, hs, s, g, d, iso, act, map, bricks, tabs,
                     voff, toff, out->positions, out->normals, out->indices, d_keys, out->cap_vertices,
                     out->cap_triangles);
  if (out->normals && out->n_vertices) {
    if (modes)
      hipLaunchKernelGGL(k_nb_normals<true>, nblk(out->n_vertices), dim3(256), 0, hs, s, out->positions,
                         out->n_vertices, out->normals);
    else
      hipLaunchKernelGGL(k_nb_normals<false>, nblk(out->n_vertices), dim3(256), 0, hs, s, out->positions,
                         out->n_vertices, out->normals);
  }
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

}  // extern "C"
