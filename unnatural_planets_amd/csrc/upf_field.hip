// upf_field.hip -- field evaluation kernels (forward and fused forward+adjoint) and
// their C-ABI entry points.
//
// Replaces the per-node callback the reference hands to cage's marching cubes,
// cubes->updateByPosition(Delegate<Real(const Vec3&)>(terrainSdfLand))
// (sources/meshGeneration.cpp:38), and the point-wise terrainSdf* calls
// (sources/coloringModes.cpp:51-57,122-151).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "upf_device_grad.hpp"
#include "upf_device_row.hpp"
#include "upf_device_row_grad.hpp"
#include "upf_internal.hpp"

namespace upf {

// One thread per node.  A workgroup covers 64 consecutive x nodes of 4 rows, so each
// wave writes 256 contiguous bytes.
constexpr int kGridBX = 64;
constexpr int kGridBY = 4;

// Optional fused output: the marching-cubes sign mask of the wave's 64-node group,
// bit = (node inside the row and f < iso), at S[(k * ny + j) * gpr + xb].
__device__ __forceinline__ void write_signs(uint64_t* __restrict__ S, const DGrid& g, int xb, int j, int k,
                                            bool valid, float v, float iso) {
  const uint64_t m = __ballot(valid && v < iso);
  if ((threadIdx.x & 63) == 0) S[((int64_t)k * g.n[1] + j) * ((g.n[0] + 63) / 64) + xb] = m;
}

// The EXT instantiation (elevation modes, further shapes) takes 108 VGPRs unconstrained
// (four waves per SIMD); at six waves (72 VGPRs, nothing spilled) the 512^3 lakes field
// went from 5.44 to 4.78 ms.
template <bool EXT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EXT ? 6 : 8))) void k_eval_grid(const DSpec s, const DGrid g, float* __restrict__ out,
                                                   uint64_t* __restrict__ S, float iso) {
  const int i = blockIdx.x * kGridBX + threadIdx.x;
  const int j = blockIdx.y * kGridBY + threadIdx.y;
  const int k = blockIdx.z;
  if (j >= g.n[1]) return;  // whole waves
  const bool valid = i < g.n[0];
  const int ic = min(i, g.n[0] - 1);
  const float x = node_coord(g.o[0], g.s[0], ic);
  const float y = node_coord(g.o[1], g.s[1], j);
  const float z = node_coord(g.o[2], g.s[2], k);
  const float v = field_eval<EXT>(s, x, y, z);
  if (valid) out[((size_t)k * g.n[1] + j) * g.n[0] + i] = v;
  if (S) write_signs(S, g, blockIdx.x, j, k, valid, v, iso);
}

// Wave-cooperative rows (upf_device_row.hpp): each wave evaluates 64 x nodes of
// kRowsPerWave rows j .. j + kRowsPerWave - 1 of one plane (they share the lattice column
// spans and the x part of the interpolation); lanes past the end of the row stand in for
// the last node so that all 64 lanes take part in the shared lattice work, rows past the
// end of the plane repeat its last row, and only real nodes are stored.  A workgroup
// covers 64 x nodes of kRowWaves row groups.
// Two rows: four rows per wave would share more (the x part over 4 rows) but measured
// 3.5 ms at the same workload: the fully unrolled 4-row layers spill (400-600 B of
// scratch at 80 to 128 VGPRs).
constexpr int kRowsPerWave = 2;
constexpr int kRowWaves = 4;
static_assert(kRowsPerWave <= kRowMaxRows, "the per-octave table holds kRowMaxRows rows");

// Seven waves per SIMD (at most 72 VGPRs): the kernel keeps the VALU busy ~93% of the
// time at five waves and the extra waves fill the issue slots the others leave.  On the
// 512^3 torus workload: five waves 1.80 ms, six 1.56 ms, seven 1.45 ms once the
// lane-derived values are recomputed at their use (lane_of; before, seven waves spilled
// 11 VGPRs to scratch, 2.5 GB of writes per launch), eight 1.69 ms (29 VGPRs spilled).
// The wave's LDS bases come from a scalar wave index and the node index is recomputed
// after the layers, so nothing is spilled.
template <bool COMPOSED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_eval_grid_row(
    const DSpec s, const DGrid g, float* __restrict__ out, uint64_t* __restrict__ S, float iso) {
  constexpr int R = kRowsPerWave;
  __shared__ float4 s_slice[kRowWaves][kRowSlice];
  __shared__ float4 s_gtab[16];
  __shared__ float4 s_oct[kRowWaves][kRowOctTab];
  if (threadIdx.y == 0 && threadIdx.x < 16) s_gtab[threadIdx.x] = gradient_entry(threadIdx.x);
  __syncthreads();
  const int lane = threadIdx.x;
  const int j = __builtin_amdgcn_readfirstlane((blockIdx.y * kRowWaves + threadIdx.y) * R);
  const int k = blockIdx.z;
  if (j >= g.n[1]) return;
  const int rows = min(R, g.n[1] - j);  // real rows of this wave
  const int ic = min((int)blockIdx.x * kGridBX + lane, g.n[0] - 1);
  const float x = node_coord(g.o[0], g.s[0], ic);
  float y[R];
#pragma unroll
  for (int r = 0; r < R; r++) y[r] = node_coord(g.o[1], g.s[1], j + min(r, rows - 1));
  const float z = node_coord(g.o[2], g.s[2], k);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.y);  // LDS bases in scalar registers
  const RowCtx c = make_row_ctx(s_slice[wv], s_gtab, s_oct[wv], lane);
  float v[R];
  row_field<R, COMPOSED>(c, s, x, y, z, v);
  const int i = (int)blockIdx.x * kGridBX + lane_of(c);  // recomputed: not live across the layers
  const bool valid = i < g.n[0];
  const size_t row0 = ((size_t)k * g.n[1] + j) * g.n[0];
#pragma unroll
  for (int r = 0; r < R; r++)
    if (r < rows && valid) out[row0 + (size_t)r * g.n[0] + i] = v[r];
  if (S) {
#pragma unroll
    for (int r = 0; r < R; r++)
      if (r < rows) write_signs(S, g, blockIdx.x, j + r, k, valid, v[r], iso);
  }
}

template <bool EXT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EXT ? 6 : 8))) void k_eval_points(const DSpec s, const float* __restrict__ xyz, int64_t n,
                                                     float* __restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  f[i] = field_eval<EXT>(s, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

// Fused forward + adjoint over a grid-stride loop; every lane of a block runs the same
// number of iterations so the wave reductions inside ThetaAcc see all 64 lanes.
template <bool THETA, bool GRID, bool EXT>
__global__ __launch_bounds__(256) void k_adjoint(const DSpec s, const DGrid g, const float* __restrict__ xyz,
                                                 int64_t n, const float* __restrict__ dLdf, float* __restrict__ fout,
                                                 float* __restrict__ dfdx, float* __restrict__ dLdtheta) {
  __shared__ float th[4][UPF_THETA_COUNT];
  const int wave = threadIdx.x >> 6;
  if (THETA) {
    for (int t = threadIdx.x; t < 4 * UPF_THETA_COUNT; t += 256) (&th[0][0])[t] = 0.0f;
    __syncthreads();
  }
  ThetaAcc acc{th[wave]};
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t base = (int64_t)blockIdx.x * 256; base < n; base += stride) {
    const int64_t idx = base + threadIdx.x;
    const bool valid = idx < n;
    const int64_t i = valid ? idx : n - 1;
    float x, y, z;
    if (GRID) {
      const int64_t nxy = (int64_t)g.n[0] * g.n[1];
      const int kk = (int)(i / nxy);
      const int64_t r = i - (int64_t)kk * nxy;
      const int jj = (int)(r / g.n[0]);
      const int ii = (int)(r - (int64_t)jj * g.n[0]);
      x = node_coord(g.o[0], g.s[0], ii);
      y = node_coord(g.o[1], g.s[1], jj);
      z = node_coord(g.o[2], g.s[2], kk);
    } else {
      x = xyz[3 * i]; y = xyz[3 * i + 1]; z = xyz[3 * i + 2];
    }
    const float w = (THETA && valid) ? dLdf[i] : 0.0f;
    float gr[3];
    const float f = field_grad<THETA, EXT>(s, x, y, z, w, gr, acc);
    if (valid) {
      if (fout) fout[i] = f;
      if (dfdx) { dfdx[3 * i] = gr[0]; dfdx[3 * i + 1] = gr[1]; dfdx[3 * i + 2] = gr[2]; }
    }
  }
  if (THETA) {
    __syncthreads();
    for (int t = threadIdx.x; t < UPF_THETA_COUNT; t += 256) {
      const float v = th[0][t] + th[1][t] + th[2][t] + th[3][t];
      if (v != 0.0f) atomicAdd(dLdtheta + t, v);
    }
  }
}

// Fused forward + adjoint over a grid with wave-cooperative rows: persistent waves
// stride over 64-node row segments (so a block's θ partials are flushed once at the
// end); lanes past the end of a row stand in for its last node with dLdf = 0.  Six
// waves per SIMD (at most 80 VGPRs): on the 512^3 torus workload with dL/dtheta, 9.94 ms
// unconstrained (109 VGPRs), 8.48 ms at five waves, 7.80 ms at six once the lane-derived
// values are recomputed at their use (lane_of; the spills left are in the prologue of
// these persistent waves).
template <bool THETA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_adjoint_rows(const DSpec s, const DGrid g, int64_t items,
                                                      const float* __restrict__ dLdf, float* __restrict__ fout,
                                                      float* __restrict__ dfdx, float* __restrict__ dLdtheta) {
  __shared__ float4 s_slice[4][kRowSlice];
  __shared__ float4 s_gtab[16];
  __shared__ float4 s_oct[4][kRowOctTab];
  __shared__ float th[4][UPF_THETA_COUNT];
  if (threadIdx.x < 16) s_gtab[threadIdx.x] = gradient_entry(threadIdx.x);
  if (THETA)
    for (int t = threadIdx.x; t < 4 * UPF_THETA_COUNT; t += 256) (&th[0][0])[t] = 0.0f;
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const RowCtx c = make_row_ctx(s_slice[wave], s_gtab, s_oct[wave], lane);
  ThetaAcc acc{th[wave]};
  const int gpr = (g.n[0] + 63) / 64;
  const int64_t stride = (int64_t)gridDim.x * 4;
  // item r = (k * ny + j) * gpr + xb, decomposed once; each step adds the stride's own
  // digits (sk, sj, sx) with carries instead of dividing r again
  int64_t r = (int64_t)blockIdx.x * 4 + wave;
  int xb = (int)(r % gpr), j = (int)((r / gpr) % g.n[1]), k = (int)(r / gpr / g.n[1]);
  const int sx = (int)(stride % gpr), sj = (int)((stride / gpr) % g.n[1]), sk = (int)(stride / gpr / g.n[1]);
  for (; r < items; r += stride) {
    // the row's first node (uniform) and this lane's 32-bit offset in it
    const int64_t row0 = ((int64_t)k * g.n[1] + j) * g.n[0];
    const int i = xb * 64 + lane;
    const bool valid = i < g.n[0];
    const int ic = min(i, g.n[0] - 1);
    const float x = node_coord(g.o[0], g.s[0], ic);
    const float y = node_coord(g.o[1], g.s[1], j);
    const float z = node_coord(g.o[2], g.s[2], k);
    const float w = (THETA && valid) ? dLdf[row0 + ic] : 0.0f;
    float gr[3];
    const float f = row_field_grad<THETA>(c, s, x, y, z, w, gr, acc);
    if (valid) {
      const int64_t node = row0 + ic;
      if (fout) fout[node] = f;
      if (dfdx) { dfdx[3 * node] = gr[0]; dfdx[3 * node + 1] = gr[1]; dfdx[3 * node + 2] = gr[2]; }
    }
    xb += sx;
    const int cx = xb >= gpr;
    xb -= cx ? gpr : 0;
    j += sj + cx;
    const int cj = j >= g.n[1];
    j -= cj ? g.n[1] : 0;
    k += sk + cj;
  }
  if (THETA) {
    __syncthreads();
    for (int t = threadIdx.x; t < UPF_THETA_COUNT; t += 256) {
      const float v = th[0][t] + th[1][t] + th[2][t] + th[3][t];
      if (v != 0.0f) atomicAdd(dLdtheta + t, v);
    }
  }
}

namespace {
int adjoint_blocks(int64_t n) {
  // enough workgroups to fill 256 CUs several times; capped so the per-block θ
  // flush stays a few hundred thousand atomics at most
  const int64_t want = (n + 255) / 256;
  return (int)(want < 4096 ? (want > 0 ? want : 1) : 4096);
}
}  // namespace

int field_grid_launch(const DSpec& s, const DGrid& g, float* d_field, int variant, uint64_t* S, float iso,
                      hipStream_t stream) {
  if (variant < 0 || variant > UPF_EVAL_PER_LANE) return UPF_ERR_BAD_SPEC;
  const bool ext = spec_ext(s);
  if (ext) variant = UPF_EVAL_PER_LANE;  // the row form covers the layer stack and the common shapes
  const dim3 blocks((g.n[0] + kGridBX - 1) / kGridBX, (g.n[1] + kGridBY - 1) / kGridBY, g.n[2]);
  if (variant == UPF_EVAL_PER_LANE) {
    if (ext)
      hipLaunchKernelGGL(k_eval_grid<true>, blocks, dim3(kGridBX, kGridBY, 1), 0, stream, s, g, d_field, S, iso);
    else
      hipLaunchKernelGGL(k_eval_grid<false>, blocks, dim3(kGridBX, kGridBY, 1), 0, stream, s, g, d_field, S, iso);
  } else {
    constexpr int rows_per_block = kRowsPerWave * kRowWaves;
    const dim3 rblocks((g.n[0] + kGridBX - 1) / kGridBX, (g.n[1] + rows_per_block - 1) / rows_per_block, g.n[2]);
    if (spec_composed(s))
      hipLaunchKernelGGL(k_eval_grid_row<true>, rblocks, dim3(kGridBX, kRowWaves, 1), 0, stream, s, g, d_field, S, iso);
    else
      hipLaunchKernelGGL(k_eval_grid_row<false>, rblocks, dim3(kGridBX, kRowWaves, 1), 0, stream, s, g, d_field, S, iso);
  }
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

}  // namespace upf

using namespace upf;

extern "C" {

int upf_field_eval_grid_ex(const upf_field_spec* spec, const upf_grid* grid, float* d_field,
                           const upf_eval_opts* opts, void* stream) {
  if (!d_field) return UPF_ERR_NULL;
  DSpec s;
  DGrid g;
  int st = to_device(spec, &s);
  if (st) return st;
  if ((st = to_device(grid, &g))) return st;
  return field_grid_launch(s, g, d_field, opts ? opts->variant : UPF_EVAL_AUTO, nullptr, 0.0f, as_stream(stream));
}

int upf_field_eval_grid(const upf_field_spec* spec, const upf_grid* grid, float* d_field, void* stream) {
  return upf_field_eval_grid_ex(spec, grid, d_field, nullptr, stream);
}

int upf_field_eval_points(const upf_field_spec* spec, const float* d_xyz, int64_t n, float* d_f, float* d_dfdx,
                          void* stream) {
  if (n < 0) return UPF_ERR_BAD_GRID;
  if (n == 0) return UPF_OK;
  if (!d_xyz || !d_f) return UPF_ERR_NULL;
  DSpec s;
  const int st = to_device(spec, &s);
  if (st) return st;
  if (d_dfdx) return upf_field_adjoint(spec, d_xyz, n, nullptr, d_f, d_dfdx, nullptr, stream);
  const dim3 pb((unsigned)((n + 255) / 256));
  if (spec_ext(s))
    hipLaunchKernelGGL(k_eval_points<true>, pb, dim3(256), 0, as_stream(stream), s, d_xyz, n, d_f);
  else
    hipLaunchKernelGGL(k_eval_points<false>, pb, dim3(256), 0, as_stream(stream), s, d_xyz, n, d_f);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

int upf_field_adjoint(const upf_field_spec* spec, const float* d_xyz, int64_t n, const float* d_dLdf, float* d_f,
                      float* d_dfdx, float* d_dLdtheta, void* stream) {
  if (n < 0) return UPF_ERR_BAD_GRID;
  if (n == 0) return UPF_OK;
  if (!d_xyz) return UPF_ERR_NULL;
  if (d_dLdtheta && !d_dLdf) return UPF_ERR_NULL;
  DSpec s;
  const int st = to_device(spec, &s);
  if (st) return st;
  if (d_dLdtheta && s.mode != UPF_ELEV_LAYERS) return UPF_ERR_BAD_SPEC;
  DGrid g{};
  const int nb = adjoint_blocks(n);
  const dim3 gb(nb), tb(256);
  hipStream_t hs = as_stream(stream);
  const bool ext = spec_ext(s);
  if (d_dLdtheta && ext)
    hipLaunchKernelGGL((k_adjoint<true, false, true>), gb, tb, 0, hs, s, g, d_xyz, n, d_dLdf, d_f, d_dfdx, d_dLdtheta);
  else if (d_dLdtheta)
    hipLaunchKernelGGL((k_adjoint<true, false, false>), gb, tb, 0, hs, s, g, d_xyz, n, d_dLdf, d_f, d_dfdx, d_dLdtheta);
  else if (ext)
    hipLaunchKernelGGL((k_adjoint<false, false, true>), gb, tb, 0, hs, s, g, d_xyz, n, d_dLdf, d_f, d_dfdx, d_dLdtheta);
  else
    hipLaunchKernelGGL((k_adjoint<false, false, false>), gb, tb, 0, hs, s, g, d_xyz, n, d_dLdf, d_f, d_dfdx, d_dLdtheta);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

int upf_field_adjoint_grid(const upf_field_spec* spec, const upf_grid* grid, const float* d_dLdf, float* d_f,
                           float* d_dfdx, float* d_dLdtheta, void* stream) {
  if (d_dLdtheta && !d_dLdf) return UPF_ERR_NULL;
  DSpec s;
  DGrid g;
  int st = to_device(spec, &s);
  if (st) return st;
  if ((st = to_device(grid, &g))) return st;
  if (s.mode != UPF_ELEV_LAYERS) return UPF_ERR_BAD_SPEC;  // the parameter adjoint covers the layer stack
  if (spec_ext(s)) {  // a further shape: per lane over the nodes
    const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const dim3 gb(adjoint_blocks(n)), tb(256);
    hipStream_t hs = as_stream(stream);
    if (d_dLdtheta)
      hipLaunchKernelGGL((k_adjoint<true, true, true>), gb, tb, 0, hs, s, g, nullptr, n, d_dLdf, d_f, d_dfdx,
                         d_dLdtheta);
    else
      hipLaunchKernelGGL((k_adjoint<false, true, true>), gb, tb, 0, hs, s, g, nullptr, n, d_dLdf, d_f, d_dfdx,
                         d_dLdtheta);
    UPF_LAUNCH_CHECK();
    return UPF_OK;
  }
  const int64_t items = (int64_t)((g.n[0] + 63) / 64) * g.n[1] * g.n[2];
  const void* kern = d_dLdtheta ? reinterpret_cast<const void*>(&k_adjoint_rows<true>)
                                : reinterpret_cast<const void*>(&k_adjoint_rows<false>);
  const int nb = (int)std::min<int64_t>((items + 3) / 4, (int64_t)resident_blocks(kern, 256));
  if (d_dLdtheta)
    hipLaunchKernelGGL((k_adjoint_rows<true>), dim3(nb), dim3(256), 0, as_stream(stream), s, g, items, d_dLdf, d_f,
                       d_dfdx, d_dLdtheta);
  else
    hipLaunchKernelGGL((k_adjoint_rows<false>), dim3(nb), dim3(256), 0, as_stream(stream), s, g, items, d_dLdf,
                       d_f, d_dfdx, d_dLdtheta);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

}  // extern "C"
