// upf_oracle_capi.cpp -- C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY).
// Loaded by tests/ through ctypes and by bench.py's CPU-baseline leg; never by libupf.
#include <cstdlib>
#include <memory>

#include "upf_oracle.hpp"
#include "upf_oracle_mc.hpp"
#include "upf_oracle_bake.hpp"
#include "upf_oracle_post.hpp"
#include "upf_oracle_material.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace upfo;

namespace {
template <class T>
void eval_grid(const upf_field_spec* spec, const upf_grid* g, T* out) {
  const SpecT<T> s = spec_to<T>(*spec);
  const int nx = g->dims[0], ny = g->dims[1], nz = g->dims[2];
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const T p[3] = {T(node_coord(g->origin[0], g->spacing[0], i)),
                        T(node_coord(g->origin[1], g->spacing[1], j)),
                        T(node_coord(g->origin[2], g->spacing[2], k))};
        out[((int64_t)k * ny + j) * nx + i] = field_eval(s, p);
      }
}

template <int K>
void seed_theta(SpecT<Dual<K>>& s) {
  int c = 0;
  for (int i = 0; i < UPF_MAX_SHAPE_PARAMS; i++) s.shape_params[i].d[c++] = 1;
  for (int l = 0; l < UPF_MAX_LAYERS; l++) s.layers[l].weight.d[c++] = 1;
  for (int l = 0; l < UPF_MAX_LAYERS; l++)
    for (int k = 0; k < UPF_MAX_OCTAVES; k++) s.layers[l].freq[k].d[c++] = 1;
  for (int l = 0; l < UPF_MAX_LAYERS; l++)
    for (int k = 0; k < UPF_MAX_OCTAVES; k++) s.layers[l].amp[k].d[c++] = 1;
}
}  // namespace

extern "C" {

int upfo_version(void) { return UPF_VERSION; }

uint32_t upfo_hash3(uint32_t seed, int32_t x, int32_t y, int32_t z) { return hash3(seed, x, y, z); }
uint32_t upfo_hash1(uint32_t k) { return hash1(k); }

// One octave of raw noise at lattice coordinates q (no frequency scaling).
double upfo_noise_f64(int noise, uint32_t seed, const double* q, double jitter) {
  switch (noise) {
    case UPF_NOISE_GRADIENT: return noise_gradient<double>(seed, q);
    case UPF_NOISE_VALUE: return noise_value<double>(seed, q);
    case UPF_NOISE_CELLULAR: return noise_cellular<double>(seed, q, jitter);
    case UPF_NOISE_SIMPLEX: return noise_simplex<double>(seed, q);
    default: return NAN;
  }
}
float upfo_noise_f32(int noise, uint32_t seed, const float* q, float jitter) {
  switch (noise) {
    case UPF_NOISE_GRADIENT: return noise_gradient<float>(seed, q);
    case UPF_NOISE_VALUE: return noise_value<float>(seed, q);
    case UPF_NOISE_CELLULAR: return noise_cellular<float>(seed, q, jitter);
    case UPF_NOISE_SIMPLEX: return noise_simplex<float>(seed, q);
    default: return NAN;
  }
}

void upfo_eval_points_f32(const upf_field_spec* spec, const float* xyz, int64_t n, float* out) {
  const SpecT<float> s = spec_to<float>(*spec);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) out[i] = field_eval(s, xyz + 3 * i);
}

void upfo_eval_points_f64(const upf_field_spec* spec, const float* xyz, int64_t n, double* out) {
  const SpecT<double> s = spec_to<double>(*spec);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    const double p[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    out[i] = field_eval(s, p);
  }
}

void upfo_eval_grid_f32(const upf_field_spec* spec, const upf_grid* g, float* out) { eval_grid<float>(spec, g, out); }
void upfo_eval_grid_f64(const upf_field_spec* spec, const upf_grid* g, double* out) { eval_grid<double>(spec, g, out); }

int upfo_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void upfo_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

// Forward-mode df/dx (FP64 duals) at n points: f -> out_f[n], grad -> out_g[n][3].
void upfo_grad_points(const upf_field_spec* spec, const float* xyz, int64_t n, double* out_f, double* out_g) {
  const SpecT<Dual<3>> s = spec_to<Dual<3>>(*spec);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    Dual<3> p[3] = {Dual<3>(xyz[3 * i]), Dual<3>(xyz[3 * i + 1]), Dual<3>(xyz[3 * i + 2])};
    for (int a = 0; a < 3; a++) p[a].d[a] = 1;
    const Dual<3> f = field_eval(s, p);
    out_f[i] = f.v;
    for (int a = 0; a < 3; a++) out_g[3 * i + a] = f.d[a];
  }
}

// sum_i dLdf[i] * df_i/dθ (FP64 duals) -> out[UPF_THETA_COUNT] in upf_theta_grad order.
void upfo_theta_grad(const upf_field_spec* spec, const float* xyz, int64_t n, const float* dLdf, double* out) {
  SpecT<Dual<UPF_THETA_COUNT>> s = spec_to<Dual<UPF_THETA_COUNT>>(*spec);
  seed_theta(s);
  for (int j = 0; j < UPF_THETA_COUNT; j++) out[j] = 0;
  for (int64_t i = 0; i < n; i++) {
    const Dual<UPF_THETA_COUNT> p[3] = {Dual<UPF_THETA_COUNT>(xyz[3 * i]), Dual<UPF_THETA_COUNT>(xyz[3 * i + 1]),
                                        Dual<UPF_THETA_COUNT>(xyz[3 * i + 2])};
    const Dual<UPF_THETA_COUNT> f = field_eval(s, p);
    for (int j = 0; j < UPF_THETA_COUNT; j++) out[j] += (double)dLdf[i] * f.d[j];
  }
}

// FP32 flops of one field evaluation at point xyz (counting rule in upf_oracle.hpp).
long long upfo_flops_at(const upf_field_spec* spec, const float* xyz) {
  const SpecT<CountedFloat> s = spec_to<CountedFloat>(*spec);
  const CountedFloat p[3] = {xyz[0], xyz[1], xyz[2]};
  CountedFloat::count() = 0;
  (void)field_eval(s, p);
  return CountedFloat::count();
}

// Oracle restatement of the per-octave expansion documented at upf_layer.
void upfo_layer_expand(int fractal, int octaves, float frequency, float lacunarity, float gain,
                       float* freq, float* amp) {
  const int oct = fractal == UPF_FRACTAL_NONE ? 1 : octaves;
  const float g = std::fabs(gain);
  float sum = 1.0f, a = g;
  for (int i = 1; i < oct; i++) { sum += a; a *= g; }
  float f = frequency, w = 1.0f / sum;
  if (fractal == UPF_FRACTAL_NONE) w = 1.0f;
  for (int k = 0; k < UPF_MAX_OCTAVES; k++) {
    if (k < oct) { freq[k] = f; amp[k] = w; f *= lacunarity; w *= gain; }
    else { freq[k] = 0.0f; amp[k] = 0.0f; }
  }
}

// ---- marching cubes
void upfo_mc_table(int* ntris, int* tris) {
  const McTable& t = table();
  for (int c = 0; c < 256; c++) {
    ntris[c] = t.ntris[c];
    for (int i = 0; i < 15; i++) tris[c * 15 + i] = t.tris[c][i];
  }
}

void upfo_case_codes(const float* field, const upf_grid* g, float iso, uint8_t* out) {
  GridView v{field, g->dims[0], g->dims[1], g->dims[2],
             {g->origin[0], g->origin[1], g->origin[2]}, {g->spacing[0], g->spacing[1], g->spacing[2]}};
  const int cx = v.nx - 1, cy = v.ny - 1, cz = v.nz - 1;
  for (int k = 0; k < cz; k++)
    for (int j = 0; j < cy; j++)
      for (int i = 0; i < cx; i++) out[((int64_t)k * cy + j) * cx + i] = (uint8_t)case_code(v, i, j, k, iso);
}

void* upfo_mc_run(const float* field, const upf_grid* g, float iso, int with_normals) {
  GridView v{field, g->dims[0], g->dims[1], g->dims[2],
             {g->origin[0], g->origin[1], g->origin[2]}, {g->spacing[0], g->spacing[1], g->spacing[2]}};
  return new McMesh(mc_extract(v, iso, with_normals != 0));
}
void upfo_mc_counts(void* h, int64_t* nv, int64_t* nt) {
  const McMesh* m = static_cast<McMesh*>(h);
  *nv = (int64_t)m->positions.size() / 3;
  *nt = (int64_t)m->indices.size() / 3;
}
void upfo_mc_copy(void* h, float* pos, float* nrm, uint32_t* idx) {
  const McMesh* m = static_cast<McMesh*>(h);
  if (pos) std::memcpy(pos, m->positions.data(), m->positions.size() * sizeof(float));
  if (nrm && !m->normals.empty()) std::memcpy(nrm, m->normals.data(), m->normals.size() * sizeof(float));
  if (idx) std::memcpy(idx, m->indices.data(), m->indices.size() * sizeof(uint32_t));
}
void upfo_mc_free(void* h) { delete static_cast<McMesh*>(h); }

// ---- texel bake
int upfo_atlas_side(int64_t n, int w, int h) { return atlas_side(n, w, h); }
void upfo_atlas(int64_t n, int w, int h, float* uv) { atlas(n, w, h, uv); }
void upfo_rasterize(const float* uv, int64_t n, int w, int h, int32_t* tri) { rasterize(uv, n, w, h, tri); }
void upfo_bake(const upf_bake_spec* b, const float* P, const float* N, const uint32_t* I, const float* uv,
               const int32_t* tri, int w, int h, float* albedo, float* bump, float* rough) {
  bake(*b, P, N, I, uv, tri, w, h, albedo, bump, rough);
}

void upfo_image_dilate(float* img, int w, int h, int channels, int rounds) { dilate_rounds(img, w, h, channels, rounds); }
void upfo_image_to_u8(const float* img, int64_t n, uint8_t* out) {
  for (int64_t i = 0; i < n; i++) out[i] = to_u8(img[i]);
}

void upfo_image_height_to_normal(const uint8_t* in, int w, int h, float strength, uint8_t* out) {
  height_to_normal(in, w, h, strength, out);
}
void upfo_image_special_to_pbr(const uint8_t* in, int64_t n, uint8_t* out) { special_to_pbr(in, n, out); }

// ---- the craters' Voronoi: the nearest point to each of n positions
void upfo_voronoi_nearest(uint32_t seed, int ppc, float cell_size, const float* pos, int64_t n, float* out) {
  for (int64_t i = 0; i < n; i++) voronoi_nearest(seed, ppc, cell_size, pos + 3 * i, out + 3 * i);
}

void upfo_voronoi_cell_points(uint32_t seed, int ppc, float cell_size, const int32_t* cell, float* out) {
  voronoi_cell_points(seed, ppc, cell_size, cell, out);
}

// ---- the material stack: terrainTile at n points -> upf_tile[n]
void upfo_material_tiles(const upf_material_spec* spec, const float* pos, const float* nrm, int64_t n, upf_tile* out) {
  const mat::Spec s = mat::spec_from(*spec);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; i++) {
    mat::Tile t;
    for (int a = 0; a < 3; a++) { t.pos[a] = pos[3 * i + a]; t.nrm[a] = nrm[3 * i + a]; }
    mat::tile(s, t);
    upf_tile& o = out[i];
    for (int a = 0; a < 3; a++) o.albedo[a] = t.albedo[a];
    o.roughness = t.roughness; o.metallic = t.metallic; o.height = t.height; o.opacity = t.opacity;
    o.elevation = t.elevation; o.slope = t.slope; o.temperature = t.temperature; o.precipitation = t.precipitation;
    o.biome = t.biome; o.type = t.type;
    o.reserved[0] = o.reserved[1] = o.reserved[2] = 0;
  }
}
// Generator::pixel with the full material (sources/textures.cpp:24-43): the tile at the
// texel's barycentric position with the normalized barycentric normal.
void upfo_bake_material(const upf_material_spec* spec, const float* P, const float* N, const uint32_t* I,
                        const float* uv, const int32_t* tri, int w, int h, float* albedo, float* special,
                        float* height) {
  const mat::Spec s = mat::spec_from(*spec);
  const int ac = spec->purpose == UPF_PURPOSE_WATER ? 4 : 3;
#pragma omp parallel for schedule(dynamic, 1)
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const int64_t o = (int64_t)y * w + x;
      const int32_t t = tri[o];
      if (t < 0) {
        for (int c = 0; c < ac; c++) albedo[ac * o + c] = NAN;
        special[2 * o] = special[2 * o + 1] = height[o] = NAN;
        continue;
      }
      const float ax = uv[6 * (int64_t)t] * (float)w, ay = uv[6 * (int64_t)t + 1] * (float)h;
      const float bx = uv[6 * (int64_t)t + 2] * (float)w, by = uv[6 * (int64_t)t + 3] * (float)h;
      const float cx = uv[6 * (int64_t)t + 4] * (float)w, cy = uv[6 * (int64_t)t + 5] * (float)h;
      const float px = (float)x + 0.5f, py = (float)y + 0.5f;
      const float area = edge_fn(ax, ay, bx, by, cx, cy);
      const float wt[3] = {edge_fn(bx, by, cx, cy, px, py) / area, edge_fn(cx, cy, ax, ay, px, py) / area,
                           edge_fn(ax, ay, bx, by, px, py) / area};
      const uint32_t iv[3] = {I[3 * (int64_t)t], I[3 * (int64_t)t + 1], I[3 * (int64_t)t + 2]};
      mat::Tile tl;
      float n[3];
      for (int a = 0; a < 3; a++) {
        tl.pos[a] = wt[0] * P[3 * iv[0] + a] + wt[1] * P[3 * iv[1] + a] + wt[2] * P[3 * iv[2] + a];
        n[a] = wt[0] * N[3 * iv[0] + a] + wt[1] * N[3 * iv[1] + a] + wt[2] * N[3 * iv[2] + a];
      }
      const float nl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      if (nl > 0.0f) for (int a = 0; a < 3; a++) tl.nrm[a] = n[a] / nl;
      else { tl.nrm[0] = 0.0f; tl.nrm[1] = 1.0f; tl.nrm[2] = 0.0f; }
      mat::tile(s, tl);
      for (int a = 0; a < 3; a++) albedo[ac * o + a] = tl.albedo[a];
      if (ac == 4) albedo[ac * o + 3] = tl.opacity;
      special[2 * o] = tl.roughness;
      special[2 * o + 1] = tl.metallic;
      height[o] = tl.height;
    }
}

// One generate* step of the stack on prepared tiles (for the per-step pins): the tile's
// inputs (elevation, slope, temperature, precipitation, biome, type and the colour
// accumulated so far) come in through `io`, the step's outputs go back there.
// step: 0 elevation, 1 precipitation, 2 temperature, 3 slope, 4 biome, 5 type, 6 bedrock,
// 7 cliffs, 8 mica, 9 dirt, 10 sand, 11 grass, 12 boulders, 13 water, 14 flowers, 15 ice,
// 16 snow, 17 visualization, 18 finalization.
int upfo_material_step(const upf_material_spec* spec, int step, const float* pos, const float* nrm, int64_t n,
                       upf_tile* io) {
  if (step < 0 || step > 18) return 1;
  const mat::Spec s = mat::spec_from(*spec);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; i++) {
    mat::Tile t;
    upf_tile& o = io[i];
    for (int a = 0; a < 3; a++) { t.pos[a] = pos[3 * i + a]; t.nrm[a] = nrm[3 * i + a]; t.albedo[a] = o.albedo[a]; }
    t.roughness = o.roughness; t.metallic = o.metallic; t.height = o.height; t.opacity = o.opacity;
    t.elevation = o.elevation; t.slope = o.slope; t.temperature = o.temperature; t.precipitation = o.precipitation;
    t.biome = o.biome; t.type = o.type;
    switch (step) {
      case 0: t.elevation = s.purpose == UPF_PURPOSE_LAND ? mat::elev_sdf(s, t.pos) : elevation_eval(s.field, t.pos); break;
      case 1: mat::precipitation(s, t); break;
      case 2: mat::temperature(s, t); break;
      case 3: mat::slope(s, t); break;
      case 4: mat::biome(t); break;
      case 5: mat::type(t); break;
      case 6: mat::bedrock(s, t); break;
      case 7: mat::cliffs(s, t); break;
      case 8: mat::mica(s, t); break;
      case 9: mat::dirt(s, t); break;
      case 10: mat::sand(s, t); break;
      case 11: mat::grass(s, t); break;
      case 12: mat::boulders(s, t); break;
      case 13: mat::water(s, t); break;
      case 14: mat::flowers(s, t); break;
      case 15: mat::ice(s, t); break;
      case 16: mat::snow(s, t); break;
      case 17: mat::visualization(t); break;
      default: mat::finalize(t); break;
    }
    for (int a = 0; a < 3; a++) o.albedo[a] = t.albedo[a];
    o.roughness = t.roughness; o.metallic = t.metallic; o.height = t.height; o.opacity = t.opacity;
    o.elevation = t.elevation; o.slope = t.slope; o.temperature = t.temperature; o.precipitation = t.precipitation;
    o.biome = t.biome; o.type = t.type;
  }
  return 0;
}
// Voronoi evaluate(position, normal).points[0] of the stack (the plane-projected nearest
// point, sources/voronoi.cpp:58-72) at n positions.
void upfo_voronoi_plane(uint32_t seed, int ppc, float cell, const float* pos, const float* nrm, int64_t n, float* out) {
  for (int64_t i = 0; i < n; i++) mat::voronoi_plane(seed, ppc, cell, pos + 3 * i, nrm + 3 * i, out + 3 * i);
}

// single noise objects of the stack, for the pins: kind, fractal, octaves, frequency,
// lacunarity, gain, distance, operation
float upfo_material_noise(int kind, uint32_t seed, float freq, int fractal, int octaves, float lac, float gain, int dist,
                          int ret, const float* p) {
  mat::Cfg c{(mat::Kind)kind, freq, fractal, octaves, lac, gain, (mat::Dist)dist, (mat::Ret)ret};
  return mat::noise(c, seed, p);
}
void upfo_rgb_to_hsv(const float* c, float* o) { mat::rgb_to_hsv(c, o); }
void upfo_hsv_to_rgb(const float* c, float* o) { mat::hsv_to_rgb(c, o); }

// ---- mesh post-processing (keep flags)
void upfo_mesh_remove_disconnected(int64_t nv, int64_t nt, const uint32_t* idx, uint8_t* keep_v, uint8_t* keep_t) {
  remove_disconnected(nv, nt, idx, keep_v, keep_t);
}
void upfo_mesh_water_keep(int64_t nv, int64_t nt, const uint32_t* idx, const uint8_t* valid, uint8_t* keep_v,
                          uint8_t* keep_t) {
  water_keep(nv, nt, idx, valid, keep_v, keep_t);
}

// ---- the project's pow and atan (upf_math.hpp), for the pins against Python's math
void upfo_powf_n(const float* x, const float* e, float* out, int64_t n) {
  for (int64_t i = 0; i < n; i++) out[i] = upf_math::powf_(x[i], e[i]);
}
void upfo_atanf_n(const float* x, float* out, int64_t n) {
  for (int64_t i = 0; i < n; i++) out[i] = upf_math::atanf_(x[i]);
}

// ---- sources/math.cpp helpers (FP64), for closed-form checks
double upfo_rescale(double v, double ia, double ib, double oa, double ob) { return rescale<double>(v, ia, ib, oa, ob); }
double upfo_sharp_edge(double v, double p) { return sharp_edge<double>(v, p); }
double upfo_terrace(double x, double s) { return terrace<double>(x, s); }
double upfo_smooth_min(double a, double b, double k) { return smooth_min<double>(a, b, k); }
double upfo_smooth_max(double a, double b, double k) { return smooth_max<double>(a, b, k); }

}  // extern "C"
