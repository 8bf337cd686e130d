// abi_check.cpp -- drives libupf's host-side C entry points (spec validation and
// conversion, layer expansion, elevation-mode set-up, the flop tables, status strings;
// unnatural_planets_amd/csrc/upf_common.cpp) with valid and randomly corrupted specs, in
// a build whose host code carries AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/sanitize/Makefile).  No GPU is touched.  tests/test_sanitizers.py runs it.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "../../include/upf.h"

int main() {
  std::mt19937 rng(11);
  int ok = 0, rejected = 0;
  for (int status = -2; status < 14; status++)
    if (!upf_status_string(status)) return 1;
  for (int trial = 0; trial < 20000; trial++) {
    upf_field_spec s = {};
    s.shape = (int)(rng() % (UPF_SHAPE_COUNT + 2)) - 1;
    s.compose = (int)(rng() % (UPF_COMPOSE_COUNT + 1));
    s.elevation_scale = (trial % 17) ? 0.1f : 0.0f;
    for (int i = 0; i < UPF_MAX_SHAPE_PARAMS; i++) s.shape_params[i] = (float)(rng() % 2000) * 0.001f;
    const int mode = (int)(rng() % (UPF_ELEV_COUNT + 1));
    if (mode > 0 && mode < UPF_ELEV_COUNT) {
      (void)upf_spec_set_elevation_mode(&s, mode, rng());
    } else {
      s.n_layers = (int)(rng() % (UPF_MAX_LAYERS + 2));
      for (int l = 0; l < UPF_MAX_LAYERS; l++)
        (void)upf_layer_init(&s.layers[l], (int)(rng() % 6), (int)(rng() % 4) - 1, (int)(rng() % 10), 1.5f, 2.0f,
                             (trial % 3) ? 0.5f : -0.4f, 0.05f, rng(), 0.9f, (int)(rng() % 4));
      s.elevation_mode = mode;
    }
    if (trial % 5 == 0) {  // corrupt a few bytes of the struct (not the weights pointer)
      unsigned char* b = reinterpret_cast<unsigned char*>(&s);
      for (int k = 0; k < 3; k++) {
        const size_t at = rng() % sizeof(s);
        if (at >= offsetof(upf_field_spec, shape_weights) && at < offsetof(upf_field_spec, shape_weights) + 8) continue;
        b[at] = (unsigned char)rng();
      }
    }
    if (s.shape == UPF_SHAPE_SINE_NET) s.shape_weights = nullptr;  // must be refused, never read
    const int st = upf_spec_validate(&s);
    if (st == UPF_OK) {
      ok++;
      (void)upf_flops_per_sample(&s);
      (void)upf_adjoint_flops_per_sample(&s);
    } else {
      rejected++;
    }
  }
  upf_layer L;
  if (upf_layer_init(nullptr, 1, 1, 4, 1, 2, 0.5f, 1, 0, 1, 0) == UPF_OK) return 1;
  if (upf_layer_init(&L, 1, 1, 99, 1, 2, 0.5f, 1, 0, 1, 0) == UPF_OK) return 1;
  if (upf_spec_validate(nullptr) == UPF_OK) return 1;
  std::printf("abi_check: ok (%d valid specs, %d rejected)\n", ok, rejected);
  return ok > 100 && rejected > 100 ? 0 : 1;
}
