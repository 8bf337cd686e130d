// upf_material.hip -- the full material stack (terrainTile with coloringDefault /
// Barren / Debug) per point and per texel: upf_material_tiles, upf_bake_material.
//
// VALU-bound: a default tile evaluates up to ~40 noise objects, 9 shape evaluations and
// the elevation stack; nothing is read per texel but the triangle id and the triangle's
// three vertices (cached, neighbouring texels share triangles).
#include <hip/hip_runtime.h>

#include <cmath>

#include "upf_internal.hpp"
#include "upf_material.hpp"

namespace upf {

__global__ __launch_bounds__(256) void k_material_tiles(const DMaterial m, const float* __restrict__ pos,
                                                        const float* __restrict__ nrm, int64_t n,
                                                        upf_tile* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  MTile t;
  t.px = pos[3 * i]; t.py = pos[3 * i + 1]; t.pz = pos[3 * i + 2];
  t.nx = nrm[3 * i]; t.ny = nrm[3 * i + 1]; t.nz = nrm[3 * i + 2];
  material_tile(m, t);
  upf_tile o;
  o.albedo[0] = t.r; o.albedo[1] = t.g; o.albedo[2] = t.b;
  o.roughness = t.roughness; o.metallic = t.metallic; o.height = t.height; o.opacity = t.opacity;
  o.elevation = t.elevation; o.slope = t.slope; o.temperature = t.temperature; o.precipitation = t.precipitation;
  o.biome = t.biome; o.type = t.type;
  o.reserved[0] = o.reserved[1] = o.reserved[2] = 0;
  out[i] = o;
}

__device__ __forceinline__ float edge_f(float ax, float ay, float bx, float by, float px, float py) {
  const float u = (bx - ax) * (py - ay);
  const float v = (by - ay) * (px - ax);
  return u - v;
}

// One texel per lane over a 64 x 4 tile: barycentrics of the texel centre in its UV
// triangle (the bake's rule), the tile at the interpolated position with the normalized
// interpolated normal, and the reference's channel layout.
__global__ __launch_bounds__(256) void k_bake_material(const DMaterial m, const float* __restrict__ P,
                                                       const float* __restrict__ N, int64_t nv,
                                                       const uint32_t* __restrict__ I, int64_t nt,
                                                       const float* __restrict__ uv, const int32_t* __restrict__ tri,
                                                       int w, int h, float* __restrict__ albedo,
                                                       float* __restrict__ special, float* __restrict__ height) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= w || y >= h) return;
  const int64_t o = (int64_t)y * w + x;
  const int ac = m.purpose == UPF_PURPOSE_WATER ? 4 : 3;
  const int32_t t = tri[o];
  uint32_t i0 = 0, i1 = 0, i2 = 0;
  const bool ok = t >= 0 && t < nt;
  if (ok) {
    i0 = I[3 * (int64_t)t];
    i1 = I[3 * (int64_t)t + 1];
    i2 = I[3 * (int64_t)t + 2];
  }
  if (!ok || i0 >= nv || i1 >= nv || i2 >= nv) {
    const float nan = __builtin_nanf("");
    for (int c = 0; c < ac; c++) albedo[ac * o + c] = nan;
    special[2 * o] = special[2 * o + 1] = nan;
    height[o] = nan;
    return;
  }
  const float ax = uv[6 * (int64_t)t] * (float)w, ay = uv[6 * (int64_t)t + 1] * (float)h;
  const float bx = uv[6 * (int64_t)t + 2] * (float)w, by = uv[6 * (int64_t)t + 3] * (float)h;
  const float cx = uv[6 * (int64_t)t + 4] * (float)w, cy = uv[6 * (int64_t)t + 5] * (float)h;
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  const float area = edge_f(ax, ay, bx, by, cx, cy);
  const float w0 = edge_f(bx, by, cx, cy, px, py) / area;
  const float w1 = edge_f(cx, cy, ax, ay, px, py) / area;
  const float w2 = edge_f(ax, ay, bx, by, px, py) / area;
  float p[3], n[3];
  for (int a = 0; a < 3; a++) {
    p[a] = w0 * P[3 * (int64_t)i0 + a] + w1 * P[3 * (int64_t)i1 + a] + w2 * P[3 * (int64_t)i2 + a];
    n[a] = w0 * N[3 * (int64_t)i0 + a] + w1 * N[3 * (int64_t)i1 + a] + w2 * N[3 * (int64_t)i2 + a];
  }
  const float nl = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  MTile tl;
  tl.px = p[0]; tl.py = p[1]; tl.pz = p[2];
  if (nl > 0.0f) { tl.nx = n[0] / nl; tl.ny = n[1] / nl; tl.nz = n[2] / nl; }
  else { tl.nx = 0.0f; tl.ny = 1.0f; tl.nz = 0.0f; }
  material_tile(m, tl);
  albedo[ac * o] = tl.r;
  albedo[ac * o + 1] = tl.g;
  albedo[ac * o + 2] = tl.b;
  if (ac == 4) albedo[ac * o + 3] = tl.opacity;
  special[2 * o] = tl.roughness;
  special[2 * o + 1] = tl.metallic;
  height[o] = tl.height;
}

int to_device(const upf_material_spec* in, DMaterial* out) {
  if (!in || !out) return UPF_ERR_NULL;
  int st = to_device(&in->field, &out->field);
  if (st) return st;
  if (in->coloring < 0 || in->coloring >= UPF_COLORING_COUNT) return UPF_ERR_BAD_SPEC;
  if (in->purpose < UPF_PURPOSE_LAND || in->purpose > UPF_PURPOSE_NAVIGATION) return UPF_ERR_BAD_SPEC;
  if (in->purpose == UPF_PURPOSE_LAND && in->field.elevation_scale == 0.0f) return UPF_ERR_INVALID_VALUE;
  out->coloring = in->coloring;
  out->purpose = in->purpose;
  out->poles = in->poles ? 1 : 0;
  out->flowers = in->flowers ? 1 : 0;
  out->seed = in->seed;
  out->pad[0] = out->pad[1] = out->pad[2] = 0;
  return UPF_OK;
}

}  // namespace upf

using namespace upf;

extern "C" {

int upf_material_tiles(const upf_material_spec* spec, const float* d_pos, const float* d_nrm, int64_t n,
                       upf_tile* d_tiles, void* stream) {
  if (n < 0) return UPF_ERR_BAD_GRID;
  DMaterial m;
  const int st = to_device(spec, &m);
  if (st) return st;
  if (n == 0) return UPF_OK;
  if (!d_pos || !d_nrm || !d_tiles) return UPF_ERR_NULL;
  hipLaunchKernelGGL(k_material_tiles, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), m, d_pos,
                     d_nrm, n, d_tiles);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

int upf_bake_material(const upf_material_spec* spec, const float* d_P, const float* d_N, int64_t nv,
                      const uint32_t* d_I, int64_t n, const float* d_uv, const int32_t* d_tri, int32_t w, int32_t h,
                      float* d_albedo, float* d_special, float* d_height, void* stream) {
  if (!spec || !d_tri || !d_albedo || !d_special || !d_height) return UPF_ERR_NULL;
  if (n > 0 && (!d_P || !d_N || !d_I || !d_uv)) return UPF_ERR_NULL;
  if (w < 1 || h < 1 || n < 0 || nv < 0) return UPF_ERR_BAD_GRID;
  DMaterial m;
  const int st = to_device(spec, &m);
  if (st) return st;
  hipLaunchKernelGGL(k_bake_material, dim3((w + 63) / 64, (h + 3) / 4), dim3(256), 0, as_stream(stream), m, d_P, d_N,
                     nv, d_I, n, d_uv, d_tri, w, h, d_albedo, d_special, d_height);
  UPF_LAUNCH_CHECK();
  return UPF_OK;
}

}  // extern "C"
