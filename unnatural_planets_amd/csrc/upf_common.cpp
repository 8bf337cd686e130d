// upf_common.cpp -- spec validation and conversion, layer expansion, the flop table
// and status strings of libupf (host code).
#include <atomic>
#include <cmath>
#include <cstring>

#include "upf_internal.hpp"

namespace upf {

namespace {
bool finite(float v) { return std::isfinite(v); }

int shape_params_used(int shape) {
  switch (shape) {
    case UPF_SHAPE_NONE: return 0;
    case UPF_SHAPE_SPHERE: return 1;
    case UPF_SHAPE_BOX: return 4;
    case UPF_SHAPE_TORUS: return 2;
    case UPF_SHAPE_CAPSULE: return 2;
    case UPF_SHAPE_CYLINDER: return 2;
    case UPF_SHAPE_PLANE: return 4;
    case UPF_SHAPE_OCTAHEDRON: return 2;
    case UPF_SHAPE_DOUBLE_TORUS: return 3;
    case UPF_SHAPE_TORUS_CROSS: return 3;
    case UPF_SHAPE_H2O: case UPF_SHAPE_BOWL: case UPF_SHAPE_INSIDE_CUBE: case UPF_SHAPE_WORMHOLE:
    case UPF_SHAPE_PIPE: case UPF_SHAPE_HEMISPHERES: return 0;
    case UPF_SHAPE_H3O: case UPF_SHAPE_H4O: case UPF_SHAPE_GEAR: case UPF_SHAPE_MANDELBULB: case UPF_SHAPE_MOBIUS:
    case UPF_SHAPE_TWISTED_PLANE: case UPF_SHAPE_BELT: case UPF_SHAPE_TRIANGULAR_PRISM:
    case UPF_SHAPE_HEXAGONAL_PRISM: case UPF_SHAPE_TWISTED_HEX_PRISM: case UPF_SHAPE_TETRAHEDRON:
    case UPF_SHAPE_KNOT: return 0;
    case UPF_SHAPE_FIBERS: return 3;
    case UPF_SHAPE_ASTEROID: return 2;
    case UPF_SHAPE_SINE_NET: return 3;
    default: return -1;
  }
}
}  // namespace

int to_device(const upf_field_spec* in, DSpec* out) {
  if (!in || !out) return UPF_ERR_NULL;
  if (shape_params_used(in->shape) < 0) return UPF_ERR_BAD_SPEC;
  if (in->compose < 0 || in->compose >= UPF_COMPOSE_COUNT) return UPF_ERR_BAD_SPEC;
  if (in->n_layers < 0 || in->n_layers > UPF_MAX_LAYERS) return UPF_ERR_BAD_SPEC;
  const int mode = in->elevation_mode;
  if (mode < 0 || mode >= UPF_ELEV_COUNT) return UPF_ERR_BAD_SPEC;
  if (mode != UPF_ELEV_LAYERS && in->n_layers != elevation_mode_layers(mode)) return UPF_ERR_BAD_SPEC;
  for (int i = 0; i < UPF_MAX_SHAPE_PARAMS; i++)
    if (!finite(in->shape_params[i])) return UPF_ERR_INVALID_VALUE;
  if (!finite(in->elevation_scale)) return UPF_ERR_INVALID_VALUE;
  if (in->compose == UPF_COMPOSE_ELEVATION && in->elevation_scale == 0.0f) return UPF_ERR_INVALID_VALUE;
  if (in->shape == UPF_SHAPE_SINE_NET) {  // weights, and nonzero scale and divisors
    if (!in->shape_weights) return UPF_ERR_NULL;
    if (in->shape_params[0] == 0.0f || in->shape_params[1] == 0.0f || in->shape_params[2] == 0.0f)
      return UPF_ERR_INVALID_VALUE;
  }
  std::memset(out, 0, sizeof(*out));
  out->shape = in->shape;
  out->compose = in->compose;
  out->n_layers = in->n_layers;
  for (int i = 0; i < UPF_MAX_SHAPE_PARAMS; i++) out->sp[i] = in->shape_params[i];
  out->escale = in->elevation_scale;
  out->mode = mode;
  out->w = in->shape == UPF_SHAPE_SINE_NET ? in->shape_weights : nullptr;
  out->all_add = 1;
  for (int l = 0; l < in->n_layers; l++) {
    const upf_layer& a = in->layers[l];
    DLayer& b = out->L[l];
    if (mode == UPF_ELEV_CRATERS && l == 0) {  // the impact Voronoi: seed, points per cell, cell size
      if (a.octaves < 1 || a.octaves > UPF_MAX_OCTAVES) return UPF_ERR_BAD_SPEC;
      if (!finite(a.freq[0]) || !(a.freq[0] > 0.0f)) return UPF_ERR_INVALID_VALUE;
      b.octaves = a.octaves;
      b.seed = a.seed;
      b.freq[0] = a.freq[0];
      continue;
    }
    if (a.noise < 1 || a.noise >= UPF_NOISE_COUNT) return UPF_ERR_BAD_SPEC;
    if (a.fractal < 0 || a.fractal >= UPF_FRACTAL_COUNT) return UPF_ERR_BAD_SPEC;
    if (a.blend < 0 || a.blend >= UPF_BLEND_COUNT) return UPF_ERR_BAD_SPEC;
    const int oct = a.fractal == UPF_FRACTAL_NONE ? 1 : a.octaves;
    if (oct < 1 || oct > UPF_MAX_OCTAVES) return UPF_ERR_BAD_SPEC;
    if (!finite(a.weight) || !finite(a.jitter)) return UPF_ERR_INVALID_VALUE;
    for (int k = 0; k < oct; k++)
      if (!finite(a.freq[k]) || !finite(a.amp[k])) return UPF_ERR_INVALID_VALUE;
    b.noise = a.noise;
    b.fractal = a.fractal;
    b.octaves = oct;
    b.blend = a.blend;
    b.seed = a.seed;
    b.weight = a.weight;
    b.jitter = a.jitter;
    b.js = a.jitter / 1023.0f;
    b.j0 = a.jitter * 0.5f;
    for (int k = 0; k < UPF_MAX_OCTAVES; k++) {
      b.freq[k] = k < oct ? a.freq[k] : 0.0f;
      b.amp[k] = k < oct ? a.amp[k] : 0.0f;
    }
    if (a.blend != UPF_BLEND_ADD) out->all_add = 0;
  }
  return UPF_OK;
}

int resident_blocks(const void* kernel, int threads) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu <= 0)
    per_cu = 1;
  return per_cu * device_cus();
}

int device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 256;
  if (dev < 64) {
    const int c = cache[dev].load(std::memory_order_relaxed);
    if (c > 0) return c;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  if (dev < 64) cache[dev].store(cus, std::memory_order_relaxed);
  return cus;
}

int to_device(const upf_grid* in, DGrid* out) {
  if (!in || !out) return UPF_ERR_NULL;
  int64_t total = 1;
  for (int a = 0; a < 3; a++) {
    if (in->dims[a] < 2 || in->dims[a] > 65535) return UPF_ERR_BAD_GRID;
    if (!finite(in->origin[a]) || !finite(in->spacing[a]) || !(in->spacing[a] > 0.0f)) return UPF_ERR_BAD_GRID;
    total *= in->dims[a];
    out->o[a] = in->origin[a];
    out->s[a] = in->spacing[a];
    out->n[a] = in->dims[a];
  }
  if (total > (int64_t)1 << 34) return UPF_ERR_BAD_GRID;
  out->pad = 0;
  return UPF_OK;
}

}  // namespace upf

// ---------------------------------------------------------------- flop table
// FP32 flops per sample by the counting rule (FMA = 2, add/sub/mul/div/sqrt = 1,
// compare/select/floor/convert/integer = 0), following the forward formulas in
// upf_device.hpp term by term.
namespace {
double noise_flops(int noise) {
  switch (noise) {
    // q = p*f (3) + t (3) + t-1 (3) + 3 fades (3*7) + 8 corner dots (8) + 7 lerps (7*3)
    case UPF_NOISE_GRADIENT: return 3 + 3 + 3 + 21 + 8 + 21;
    // q (3) + t (3) + fades (21) + lerps (21)
    case UPF_NOISE_VALUE: return 3 + 3 + 21 + 21;
    // q (3) + js, j0 (2) + k - j0 per axis value (9) + 27 cells x (3 fma + 3 sub + 5 d2)
    // + sqrt (1)
    case UPF_NOISE_CELLULAR: return 3 + 2 + 9 + 27 * (6 + 3 + 5) + 1;
    // q (3), skew (3), skewed point (3), unskew (3), corner 0 offset (6), corners 1-3
    // (6 + 6 + 3), 4 corners x (r^2 5, t 1, t^4 2, dot 1, product 1, sum 1), scale 1
    case UPF_NOISE_SIMPLEX: return 3 + 3 + 3 + 3 + 6 + 15 + 4 * 11 + 1;
    default: return 0;
  }
}

// Host scalar that runs the extended-shape template (upf_shapes.hpp) and tallies, by the
// counting rule, the flops of the float evaluation (fwd) and of the value-and-gradient
// evaluation with three derivative directions (grad): + and - cost 1 / 4, * 1 / 10,
// / 1 / 10, sqrt 1 / 5 (sqrt, 1/(2r), three products); sign, abs, min and max are free.
struct OpCount {
  float v;
};
struct OpTally {
  double fwd = 0, grad = 0;
};
thread_local OpTally g_tally;
inline OpCount op(float v, double f, double g) {
  g_tally.fwd += f;
  g_tally.grad += g;
  return {v};
}
float value_of(const OpCount& a) { return a.v; }
OpCount lift(const OpCount&, float v) { return {v}; }
OpCount operator+(OpCount a, OpCount b) { return op(a.v + b.v, 1, 4); }
OpCount operator-(OpCount a, OpCount b) { return op(a.v - b.v, 1, 4); }
OpCount operator*(OpCount a, OpCount b) { return op(a.v * b.v, 1, 10); }
OpCount operator/(OpCount a, OpCount b) { return op(a.v / b.v, 1, 10); }
OpCount operator-(OpCount a) { return {-a.v}; }
OpCount operator+(OpCount a, float b) { return op(a.v + b, 1, 1); }
OpCount operator-(OpCount a, float b) { return op(a.v - b, 1, 1); }
OpCount operator*(OpCount a, float b) { return op(a.v * b, 1, 4); }
OpCount operator/(OpCount a, float b) { return op(a.v / b, 1, 4); }
OpCount operator+(float a, OpCount b) { return op(a + b.v, 1, 1); }
OpCount operator-(float a, OpCount b) { return op(a - b.v, 1, 1); }
OpCount operator*(float a, OpCount b) { return op(a * b.v, 1, 4); }
OpCount sq_root(OpCount a) { return op(std::sqrt(a.v), 1, 5); }
// transcendental functions counted as 1 (and 5 for value and gradient)
OpCount pow_v(OpCount a, float e) { return op(std::pow(a.v, e), 1, 5); }
OpCount floor_v(OpCount a) { return {std::floor(a.v)}; }
OpCount sin_v(OpCount a) { return op(std::sin(a.v), 1, 5); }
OpCount fma_s(float w, OpCount h, OpCount acc) { return op(std::fma(w, h.v, acc.v), 2, 8); }
OpCount cos_v(OpCount a) { return op(std::cos(a.v), 1, 5); }
OpCount atan2_v(OpCount y, OpCount x) { return op(std::atan2(y.v, x.v), 1, 12); }
OpCount acos_v(OpCount a) { return op(std::acos(a.v), 1, 7); }
OpCount log_v(OpCount a) { return op(std::log(a.v), 1, 5); }

// the point the flop count of a data-dependent shape is taken at (the oracle's count
// uses the same point)
constexpr float kCountPoint[3] = {0.3f, 0.2f, -0.4f};

// metre-scale count point for the further shapes (inside the Mandelbulb's escape radius,
// off every axis)
constexpr float kCountPointM[3] = {300.0f, 200.0f, -400.0f};

OpTally shape_ext_tally(int shape, const float* sp) {
  g_tally = OpTally{};
  if (shape == UPF_SHAPE_ASTEROID) {
    // |p| 5, direction 6, two simplex octaves (value: 75 each, gradient: 138 each), the
    // octave sum 3, the radial composition 3
    g_tally.fwd = 5 + 6 + 2 * (3 + 75) + 3 + 3;
    g_tally.grad = 5 * 4 + 6 * 10 + 2 * (3 + 138) + 3 + 3 * 4 + 18;
    return g_tally;
  }
  if (shape == UPF_SHAPE_SINE_NET) {  // a point inside the unit ball of q (the full network)
    static const float zero_w[UPF_SINE_NET_WEIGHTS] = {};
    const float s = sp[0];
    (void)upf::sine_net_t<OpCount>(zero_w, s, sp[1], sp[2], OpCount{0.3f / s}, OpCount{0.2f / s}, OpCount{-0.4f / s});
    return g_tally;
  }
  if (shape >= UPF_SHAPE_H3O) {
    (void)upf::shape_more_t<OpCount>(shape, sp, OpCount{kCountPointM[0]}, OpCount{kCountPointM[1]},
                                     OpCount{kCountPointM[2]}, OpCount{0.0f});
    return g_tally;
  }
  (void)upf::shape_ext<OpCount>(shape, sp, OpCount{kCountPoint[0]}, OpCount{kCountPoint[1]}, OpCount{kCountPoint[2]},
                           OpCount{0.0f});
  return g_tally;
}

double shape_flops(int shape, const float* sp) {
  switch (shape) {
    case UPF_SHAPE_NONE: return 0;
    case UPF_SHAPE_SPHERE: return 7;
    case UPF_SHAPE_BOX: return 11;
    case UPF_SHAPE_TORUS: return 10;
    case UPF_SHAPE_CAPSULE: return 9;
    default: return shape_ext_tally(shape, sp).fwd;
  }
}

double layer_flops(const upf_layer& L) {
  const int oct = L.fractal == UPF_FRACTAL_NONE ? 1 : L.octaves;
  return oct * (noise_flops(L.noise) + (L.fractal == UPF_FRACTAL_RIDGED ? 4 : (L.fractal == UPF_FRACTAL_FBM ? 2 : 0)));
}

// An elevation mode, tallied through its template (upf_modes.hpp) on the full path
// (every layer value -0.6, which takes the lakes' mountains branch), each layer call
// costing its layer's flops; the craters' Voronoi adds 17 flops per candidate point.
double mode_flops(const upf_field_spec* s) {
  g_tally = OpTally{};
  const auto layer = [&](int l, OpCount, OpCount, OpCount) {
    g_tally.fwd += layer_flops(s->layers[l]);
    return OpCount{-0.6f};
  };
  const upf_layer& v = s->layers[0];
  (void)upf::elevation_mode_v<OpCount>(s->elevation_mode, v.seed, v.octaves, v.freq[0] > 0.0f ? v.freq[0] : 1.0f,
                                       s->n_layers, OpCount{kCountPoint[0]}, OpCount{kCountPoint[1]},
                                       OpCount{kCountPoint[2]}, layer, OpCount{0.0f});
  double f = g_tally.fwd;
  if (s->elevation_mode == UPF_ELEV_CRATERS) f += 27.0 * v.octaves * 17 + 1;
  return f;
}

double elevation_flops(const upf_field_spec* s) {
  if (s->elevation_mode != UPF_ELEV_LAYERS) return mode_flops(s);
  double f = 0;
  for (int l = 0; l < s->n_layers; l++) {
    const upf_layer& L = s->layers[l];
    const int oct = L.fractal == UPF_FRACTAL_NONE ? 1 : L.octaves;
    const double per = noise_flops(L.noise) +
                       (L.fractal == UPF_FRACTAL_RIDGED ? 4 : (L.fractal == UPF_FRACTAL_FBM ? 2 : 0));
    f += oct * per + 1;  // weight
    f += L.blend == UPF_BLEND_ADD ? 1 : (L.blend == UPF_BLEND_MASK ? 1 : 3);
  }
  return f;
}
// Fused forward + adjoint (k_adjoint with the θ accumulation on), tallied from the
// formulas of upf_device_grad.hpp by the same rule.  Each θ component costs its own
// products plus one add for the sum over samples (the cross-lane reduction is a
// per-wave tree, counted once per sample).
double noise_d_flops(int noise) {
  switch (noise) {
    // q 3, t 3, t-1 3, fades 21, corner gradients 16, corner dots 40, value lerps 18,
    // gradient blend 63, dn/du 19, fade' 12, chain 6, final lerp 3
    case UPF_NOISE_GRADIENT: return 3 + 3 + 3 + 21 + 16 + 40 + 18 + 63 + 19 + 12 + 6 + 3;
    // q 3, t 3, fades 21, lerps 18, dx 18, dy 10, dz 6, final lerp 3
    case UPF_NOISE_VALUE: return 3 + 3 + 21 + 18 + 18 + 10 + 6 + 3;
    // forward cellular 393, 1/r, 3 products
    case UPF_NOISE_CELLULAR: return noise_flops(UPF_NOISE_CELLULAR) + 1 + 3;
    // forward, then per corner t^3 1, -8 t^3 (g.d) 2, t^4 g 3, k d 3, sums 6; scale 3
    case UPF_NOISE_SIMPLEX: return noise_flops(UPF_NOISE_SIMPLEX) + 4 * 15 + 3;
    default: return 0;
  }
}

double shape_d_flops(int shape, const float* sp) {
  switch (shape) {
    case UPF_SHAPE_NONE: return 0;
    case UPF_SHAPE_SPHERE: return 11;
    case UPF_SHAPE_BOX: return 18;
    case UPF_SHAPE_TORUS: return 18;
    case UPF_SHAPE_CAPSULE: return 15;
    default: return shape_ext_tally(shape, sp).grad;
  }
}

int shape_nparams(int shape) {
  switch (shape) {
    case UPF_SHAPE_SPHERE: return 1;
    case UPF_SHAPE_BOX: return 4;
    case UPF_SHAPE_TORUS: return 2;
    case UPF_SHAPE_CAPSULE: return 2;
    default: return 0;
  }
}
}  // namespace

extern "C" {

int upf_version(void) { return UPF_VERSION; }

double upf_adjoint_flops_per_sample(const upf_field_spec* s) {
  if (!s || s->elevation_mode != UPF_ELEV_LAYERS) return 0;
  const bool has_shape = s->compose != UPF_COMPOSE_ELEVATION_RAW;
  const bool has_elev = s->compose == UPF_COMPOSE_LAND || s->compose == UPF_COMPOSE_NAVIGATION ||
                        s->compose == UPF_COMPOSE_ELEVATION_RAW;
  double f = 0;
  if (has_shape) f += shape_d_flops(s->shape, s->shape_params) + 3 + 3 * shape_nparams(s->shape);
  if (!has_elev || s->n_layers == 0) return f + 1;
  bool all_add = true;
  for (int l = 0; l < s->n_layers; l++) all_add &= s->layers[l].blend == UPF_BLEND_ADD;
  if (!all_add || s->compose == UPF_COMPOSE_NAVIGATION) f += elevation_flops(s) + 4 * s->n_layers;
  for (int l = 0; l < s->n_layers; l++) {
    const upf_layer& L = s->layers[l];
    const int oct = L.fractal == UPF_FRACTAL_NONE ? 1 : L.octaves;
    // per octave: value accumulate 2, position-gradient scale 2 + 3 fma, q.grad 5,
    // d/dfreq 2 + 1, d/damp 1 + 1, ridged term and slope 3
    double per = noise_d_flops(L.noise) + 2 + 2 + 6 + 5 + 3;
    if (L.fractal != UPF_FRACTAL_NONE) per += 2;
    if (L.fractal == UPF_FRACTAL_RIDGED) per += 3;
    f += oct * per;
    f += 2 + 2 + 4;  // coefficients, elevation accumulate, d/dweight
  }
  return f + 2;
}

const char* upf_status_string(int status) {
  switch (status) {
    case UPF_OK: return "ok";
    case UPF_ERR_INVALID_VALUE: return "invalid value (NaN or Inf)";
    case UPF_ERR_BAD_SPEC: return "bad spec (unknown id or count out of range)";
    case UPF_ERR_CAPACITY: return "output capacity too small";
    case UPF_ERR_WORKSPACE: return "workspace missing or too small";
    case UPF_ERR_BAD_GRID: return "bad grid";
    case UPF_ERR_HIP: return "HIP runtime error";
    case UPF_ERR_NULL: return "null pointer";
    case UPF_ERR_EMPTY_MESH: return "empty mesh";
    case UPF_ERR_BAND: return "narrow band too narrow: the surface leaves the active blocks";
    default: return "unknown status";
  }
}

int upf_layer_init(upf_layer* layer, int32_t noise, int32_t fractal, int32_t octaves, float frequency,
                   float lacunarity, float gain, float weight, uint32_t seed, float jitter, int32_t blend) {
  if (!layer) return UPF_ERR_NULL;
  if (noise < 1 || noise >= UPF_NOISE_COUNT || fractal < 0 || fractal >= UPF_FRACTAL_COUNT || blend < 0 ||
      blend >= UPF_BLEND_COUNT)
    return UPF_ERR_BAD_SPEC;
  const int oct = fractal == UPF_FRACTAL_NONE ? 1 : octaves;
  if (oct < 1 || oct > UPF_MAX_OCTAVES) return UPF_ERR_BAD_SPEC;
  if (!std::isfinite(frequency) || !std::isfinite(lacunarity) || !std::isfinite(gain) || !std::isfinite(weight) ||
      !std::isfinite(jitter))
    return UPF_ERR_INVALID_VALUE;
  std::memset(layer, 0, sizeof(*layer));
  layer->noise = noise;
  layer->fractal = fractal;
  layer->octaves = oct;
  layer->blend = blend;
  layer->seed = seed;
  layer->weight = weight;
  layer->jitter = jitter;
  // fractal bounding 1 / sum |gain|^k, accumulated in FP32 in octave order
  const float g = std::fabs(gain);
  float sum = 1.0f, a = g;
  for (int i = 1; i < oct; i++) { sum += a; a *= g; }
  float w = fractal == UPF_FRACTAL_NONE ? 1.0f : 1.0f / sum;
  float f = frequency;
  for (int k = 0; k < oct; k++) {
    layer->freq[k] = f;
    layer->amp[k] = w;
    f *= lacunarity;
    w *= gain;
  }
  return UPF_OK;
}

int upf_spec_set_elevation_mode(upf_field_spec* s, int32_t mode, uint32_t seed) {
  if (!s) return UPF_ERR_NULL;
  if (mode < 0 || mode >= UPF_ELEV_COUNT) return UPF_ERR_BAD_SPEC;
  s->elevation_mode = mode;
  std::memset(s->layers, 0, sizeof(s->layers));
  upf_layer* L = s->layers;
  // (noise, fractal, octaves, frequency, lacunarity, gain) per sources/elevationModes.cpp
  switch (mode) {
    case UPF_ELEV_SIMPLE:  // :16-27
      upf_layer_init(&L[0], UPF_NOISE_SIMPLEX, UPF_FRACTAL_RIDGED, 6, 0.0005f, 2.0f, 0.4f, 1.0f, seed, 0.0f, 0);
      break;
    case UPF_ELEV_LEGACY:  // :36-54
      upf_layer_init(&L[0], UPF_NOISE_VALUE, UPF_FRACTAL_FBM, 4, 0.0005f, 2.0f, 0.5f, 1.0f, seed, 0.0f, 0);
      upf_layer_init(&L[1], UPF_NOISE_VALUE, UPF_FRACTAL_FBM, 4, 1.0f, 2.0f, 0.5f, 1.0f, seed + 1, 0.0f, 0);
      break;
    case UPF_ELEV_LAKES:    // :131-141, then the mountains :70-100
    case UPF_ELEV_ISLANDS:  // :153-163
      upf_layer_init(&L[0], UPF_NOISE_VALUE, UPF_FRACTAL_FBM, 4, 0.0013f, 2.0f, 0.5f, 1.0f, seed, 0.0f, 0);
      upf_layer_init(&L[1], UPF_NOISE_GRADIENT, UPF_FRACTAL_NONE, 1, 0.0015f, 2.0f, 0.5f, 1.0f, seed + 1, 0.0f, 0);
      upf_layer_init(&L[2], UPF_NOISE_SIMPLEX, UPF_FRACTAL_RIDGED, 4, 0.001f, 1.5f, -0.4f, 1.0f, seed + 2, 0.0f, 0);
      upf_layer_init(&L[3], UPF_NOISE_GRADIENT, UPF_FRACTAL_FBM, 3, 0.002f, 2.0f, 0.3f, 1.0f, seed + 3, 0.0f, 0);
      break;
    case UPF_ELEV_CRATERS:  // :175-190: Voronoi cell 300 m, 2 points per cell; value noise
      L[0].noise = UPF_NOISE_CELLULAR;
      L[0].octaves = 2;
      L[0].seed = seed;
      L[0].weight = 1.0f;
      L[0].freq[0] = 300.0f;
      L[0].amp[0] = 1.0f;
      upf_layer_init(&L[1], UPF_NOISE_VALUE, UPF_FRACTAL_NONE, 1, 1.0f, 2.0f, 0.5f, 1.0f, seed + 1, 0.0f, 0);
      break;
    default:
      break;
  }
  s->n_layers = upf::elevation_mode_layers(mode);
  return UPF_OK;
}

int upf_spec_validate(const upf_field_spec* spec) {
  upf::DSpec d;
  return upf::to_device(spec, &d);
}

double upf_flops_per_sample(const upf_field_spec* s) {
  if (!s) return 0;
  const double sh = shape_flops(s->shape, s->shape_params);
  switch (s->compose) {
    case UPF_COMPOSE_LAND: return sh + elevation_flops(s) + 2;
    case UPF_COMPOSE_WATER: return sh;
    case UPF_COMPOSE_NAVIGATION: return sh + elevation_flops(s) + 2;
    case UPF_COMPOSE_ELEVATION: return sh + 1;
    case UPF_COMPOSE_ELEVATION_RAW: return elevation_flops(s);
    default: return 0;
  }
}

}  // extern "C"
