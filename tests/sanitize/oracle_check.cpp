// oracle_check.cpp -- drives the CPU oracle's C entry points (oracle/upf_oracle_capi.cpp)
// over the shapes, modes, marching cubes, bake, material stack and post-processing, in a
// build with AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile).
// Any out-of-bounds access, use after free, leak or undefined operation aborts the run;
// tests/test_sanitizers.py runs it.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../../include/upf.h"

extern "C" {
void upfo_eval_points_f32(const upf_field_spec*, const float*, int64_t, float*);
void upfo_eval_points_f64(const upf_field_spec*, const float*, int64_t, double*);
void upfo_eval_grid_f32(const upf_field_spec*, const upf_grid*, float*);
void upfo_grad_points(const upf_field_spec*, const float*, int64_t, double*, double*);
void upfo_theta_grad(const upf_field_spec*, const float*, int64_t, const float*, double*);
long long upfo_flops_at(const upf_field_spec*, const float*);
void* upfo_mc_run(const float*, const upf_grid*, float, int);
void upfo_mc_counts(void*, int64_t*, int64_t*);
void upfo_mc_copy(void*, float*, float*, uint32_t*);
void upfo_mc_free(void*);
void upfo_case_codes(const float*, const upf_grid*, float, uint8_t*);
void upfo_atlas(int64_t, int, int, float*);
void upfo_rasterize(const float*, int64_t, int, int, int32_t*);
void upfo_bake(const upf_bake_spec*, const float*, const float*, const uint32_t*, const float*, const int32_t*, int, int,
               float*, float*, float*);
void upfo_image_dilate(float*, int, int, int, int);
void upfo_image_to_u8(const float*, int64_t, uint8_t*);
void upfo_image_height_to_normal(const uint8_t*, int, int, float, uint8_t*);
void upfo_material_tiles(const upf_material_spec*, const float*, const float*, int64_t, upf_tile*);
void upfo_bake_material(const upf_material_spec*, const float*, const float*, const uint32_t*, const float*,
                        const int32_t*, int, int, float*, float*, float*);
void upfo_mesh_remove_disconnected(int64_t, int64_t, const uint32_t*, uint8_t*, uint8_t*);
void upfo_mesh_water_keep(int64_t, int64_t, const uint32_t*, const uint8_t*, uint8_t*, uint8_t*);
void upfo_voronoi_nearest(uint32_t, int, float, const float*, int64_t, float*);
void upfo_powf_n(const float*, const float*, float*, int64_t);
void upfo_atanf_n(const float*, float*, int64_t);
void upfo_set_num_threads(int);
}

namespace {
upf_layer layer(int noise, int fractal, int octaves, float f, float gain, float w, uint32_t seed, int blend) {
  upf_layer L = {};
  L.noise = noise; L.fractal = fractal; L.octaves = fractal == UPF_FRACTAL_NONE ? 1 : octaves; L.blend = blend;
  L.seed = seed; L.weight = w; L.jitter = 0.9f;
  float amp = 1.0f, sum = 1.0f, g = std::fabs(gain), a = g;
  for (int i = 1; i < L.octaves; i++) { sum += a; a *= g; }
  amp = fractal == UPF_FRACTAL_NONE ? 1.0f : 1.0f / sum;
  for (int k = 0; k < L.octaves; k++) { L.freq[k] = f; L.amp[k] = amp; f *= 2.0f; amp *= gain; }
  return L;
}
upf_grid grid(float lo, float hi, int n) {
  upf_grid g = {};
  for (int a = 0; a < 3; a++) { g.origin[a] = lo; g.spacing[a] = (hi - lo) / (float)(n - 1); g.dims[a] = n; }
  return g;
}
}  // namespace

int main() {
  upfo_set_num_threads(2);
  std::mt19937 rng(7);
  std::uniform_real_distribution<float> U(-1.6f, 1.6f);
  std::vector<float> pts(3 * 500);
  for (float& v : pts) v = U(rng);
  // every shape with a three-layer stack (add, mask, lerp), every compose
  for (int shape = 0; shape < UPF_SHAPE_COUNT; shape++) {
    if (shape == UPF_SHAPE_SINE_NET) continue;  // needs the caller's weights (below)
    upf_field_spec s = {};
    s.shape = shape; s.compose = shape % UPF_COMPOSE_COUNT; s.elevation_scale = 0.5f;
    for (int i = 0; i < UPF_MAX_SHAPE_PARAMS; i++) s.shape_params[i] = 0.4f + 0.1f * i;
    s.n_layers = 3;
    s.layers[0] = layer(UPF_NOISE_GRADIENT, UPF_FRACTAL_RIDGED, 3, 2.0f, 0.5f, 0.05f, 1, UPF_BLEND_ADD);
    s.layers[1] = layer(UPF_NOISE_VALUE, UPF_FRACTAL_NONE, 1, 0.7f, 0.5f, 0.5f, 2, UPF_BLEND_MASK);
    s.layers[2] = layer(UPF_NOISE_CELLULAR, UPF_FRACTAL_FBM, 2, 3.0f, 0.5f, 0.05f, 3, UPF_BLEND_LERP);
    std::vector<float> f(500);
    std::vector<double> f64(500), g64(1500), th(UPF_THETA_COUNT);
    upfo_eval_points_f32(&s, pts.data(), 500, f.data());
    upfo_eval_points_f64(&s, pts.data(), 500, f64.data());
    upfo_grad_points(&s, pts.data(), 50, f64.data(), g64.data());
    upfo_theta_grad(&s, pts.data(), 20, f.data(), th.data());
    (void)upfo_flops_at(&s, pts.data());
  }
  {  // the sine network with arbitrary weights
    std::vector<float> w(UPF_SINE_NET_WEIGHTS);
    for (float& v : w) v = U(rng);
    upf_field_spec s = {};
    s.shape = UPF_SHAPE_SINE_NET; s.shape_params[0] = 0.5f; s.shape_params[1] = 1.0f; s.shape_params[2] = 1.4f;
    s.shape_weights = w.data(); s.elevation_scale = 1.0f;
    std::vector<float> f(500);
    upfo_eval_points_f32(&s, pts.data(), 500, f.data());
  }
  // the elevation modes (metre scale), each with its reference layers
  const int mode_layers[UPF_ELEV_COUNT] = {0, 0, 1, 2, 4, 4, 2};
  std::vector<float> mp(3 * 300);
  std::uniform_real_distribution<float> M(-2000.0f, 2000.0f);
  for (float& v : mp) v = M(rng);
  for (int mode = 1; mode < UPF_ELEV_COUNT; mode++) {
    upf_field_spec s = {};
    s.shape = UPF_SHAPE_SPHERE; s.shape_params[0] = 1100.0f; s.elevation_scale = 0.1f; s.elevation_mode = mode;
    s.n_layers = mode_layers[mode];
    for (int l = 0; l < s.n_layers; l++) s.layers[l] = layer(UPF_NOISE_SIMPLEX, UPF_FRACTAL_FBM, 3, 0.001f, 0.5f, 1.0f, 9 + l, 0);
    if (mode == UPF_ELEV_CRATERS) { s.layers[0].octaves = 2; s.layers[0].freq[0] = 300.0f; }
    std::vector<float> f(300);
    upfo_eval_points_f32(&s, mp.data(), 300, f.data());
  }
  // marching cubes, post-processing, bake and material on a small sphere planet
  upf_field_spec s = {};
  s.shape = UPF_SHAPE_SPHERE; s.shape_params[0] = 1.0f; s.elevation_scale = 1.0f; s.n_layers = 1;
  s.layers[0] = layer(UPF_NOISE_GRADIENT, UPF_FRACTAL_FBM, 4, 2.0f, 0.5f, 0.05f, 1234, 0);
  const upf_grid g = grid(-1.3f, 1.3f, 40);
  std::vector<float> field(40 * 40 * 40);
  upfo_eval_grid_f32(&s, &g, field.data());
  std::vector<uint8_t> codes(39 * 39 * 39);
  upfo_case_codes(field.data(), &g, 0.0f, codes.data());
  void* h = upfo_mc_run(field.data(), &g, 0.0f, 1);
  int64_t nv = 0, nt = 0;
  upfo_mc_counts(h, &nv, &nt);
  std::vector<float> P(3 * nv), N(3 * nv);
  std::vector<uint32_t> I(3 * nt);
  upfo_mc_copy(h, P.data(), N.data(), I.data());
  upfo_mc_free(h);
  std::vector<uint8_t> kv(nv), kt(nt), valid(nv);
  upfo_mesh_remove_disconnected(nv, nt, I.data(), kv.data(), kt.data());
  for (int64_t v = 0; v < nv; v++) valid[v] = P[3 * v + 2] < 0.2f;
  upfo_mesh_water_keep(nv, nt, I.data(), valid.data(), kv.data(), kt.data());
  const int W = 256, H = 192;
  std::vector<float> uv(6 * nt);
  upfo_atlas(nt, W, H, uv.data());
  std::vector<int32_t> tri(W * H);
  upfo_rasterize(uv.data(), nt, W, H, tri.data());
  upf_bake_spec b = {};
  b.field = s;
  b.detail = layer(UPF_NOISE_GRADIENT, UPF_FRACTAL_FBM, 3, 16.0f, 0.5f, 1.0f, 977, 0);
  b.bands[0] = -0.01f; b.bands[1] = 0.005f; b.bands[2] = 0.025f; b.band_width = 0.005f;
  b.cliff[0] = 0.05f; b.cliff[1] = 0.25f; b.detail_albedo = 0.3f;
  std::vector<float> albedo(3 * W * H), bump(W * H), rough(W * H);
  upfo_bake(&b, P.data(), N.data(), I.data(), uv.data(), tri.data(), W, H, albedo.data(), bump.data(), rough.data());
  upfo_image_dilate(albedo.data(), W, H, 3, 7);
  std::vector<uint8_t> u8(3 * W * H), nmap(3 * W * H);
  upfo_image_to_u8(albedo.data(), 3 * W * H, u8.data());
  upfo_image_height_to_normal(u8.data(), W, H, 1.0f, nmap.data());
  for (int coloring = 0; coloring < UPF_COLORING_COUNT; coloring++)
    for (int purpose = UPF_PURPOSE_LAND; purpose <= UPF_PURPOSE_NAVIGATION; purpose++) {
      upf_material_spec m = {};
      m.field.shape = UPF_SHAPE_SPHERE; m.field.shape_params[0] = 1100.0f; m.field.elevation_scale = 0.1f;
      m.field.elevation_mode = UPF_ELEV_LAKES; m.field.n_layers = 4;
      for (int l = 0; l < 4; l++) m.field.layers[l] = layer(UPF_NOISE_VALUE, UPF_FRACTAL_FBM, 3, 0.0013f, 0.5f, 1.0f, l, 0);
      m.coloring = coloring; m.purpose = purpose; m.poles = 1; m.flowers = 1; m.seed = 5;
      std::vector<upf_tile> tiles(300);
      std::vector<float> nrm(mp.size());
      for (size_t i = 0; i < mp.size(); i += 3) {
        const float l = std::sqrt(mp[i] * mp[i] + mp[i + 1] * mp[i + 1] + mp[i + 2] * mp[i + 2]);
        for (int a = 0; a < 3; a++) nrm[i + a] = mp[i + a] / l;
      }
      upfo_material_tiles(&m, mp.data(), nrm.data(), 300, tiles.data());
      std::vector<float> Pm(P.size());
      for (size_t i = 0; i < P.size(); i++) Pm[i] = P[i] * 1100.0f;
      const int ac = purpose == UPF_PURPOSE_WATER ? 4 : 3;
      std::vector<float> a(ac * W * H), sp(2 * W * H), hm(W * H);
      upfo_bake_material(&m, Pm.data(), N.data(), I.data(), uv.data(), tri.data(), W, H, a.data(), sp.data(), hm.data());
    }
  std::vector<float> vp(3 * 50), vo(3 * 50);
  for (float& v : vp) v = M(rng);
  upfo_voronoi_nearest(11, 2, 300.0f, vp.data(), 50, vo.data());
  std::vector<float> x(1000), e(1000), r(1000);
  for (int i = 0; i < 1000; i++) { x[i] = std::fabs(U(rng)); e[i] = 0.3f + 0.005f * i; }
  upfo_powf_n(x.data(), e.data(), r.data(), 1000);
  upfo_atanf_n(x.data(), r.data(), 1000);
  std::printf("oracle_check: ok (%lld vertices, %lld triangles)\n", (long long)nv, (long long)nt);
  return 0;
}
