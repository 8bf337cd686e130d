// mesh_generate_generic.cpp -- the reference-side binding of the integration guide,
// compiled against include/upf.h and linked with libupf.so: makeSpec (the spec in place
// of the mode function pointers) and meshGenerateGeneric (sources/meshGeneration.cpp:
// 32-45: the grid sweep and marching cubes, then meshRemoveDisconnected).  cage's
// Holder<Mesh> and CAGE_THROW_ERROR are replaced by a plain struct and std exceptions;
// the calls into libupf are the guide's, line for line.
//
//   mesh_generate_generic <out.bin> <box resolution> [static]
//
// writes int64 n_vertices, n_triangles, then positions, normals (float[V][3]) and
// indices (uint32[T][3]).  With "static" the sweep runs upf_field_mesh_async_static
// (no workspace) before the same post-processing.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "upf.h"

namespace {

struct Mesh {
  std::vector<float> positions, normals;
  std::vector<uint32_t> indices;
};

void check(int r) {
  if (r) throw std::runtime_error(upf_status_string(r));
}

// The reference selects terrainShapeFnc / terrainElevationFnc by mode name
// (sources/modesConfiguration.cpp:40-183); here one spec holds the choice.
upf_field_spec makeSpec(uint32_t seed) {
  upf_field_spec s = {};
  s.shape = UPF_SHAPE_SPHERE;    // sdfSphere, sources/sdf.cpp:20-23
  s.shape_params[0] = 1100;      // radius in metres
  s.elevation_scale = 1.0f / 10; // meshElevationRatio, modesConfiguration.cpp:227
  s.n_layers = 1;                // a generic fBm layer (elevationModes.cpp:16-27 form)
  check(upf_layer_init(&s.layers[0], UPF_NOISE_GRADIENT, UPF_FRACTAL_FBM, 6, 0.0005f, 2.0f, 0.4f, 1000.0f, seed,
                       1.0f, UPF_BLEND_ADD));
  s.compose = UPF_COMPOSE_LAND;
  check(upf_spec_validate(&s));
  return s;
}

Mesh meshGenerateGeneric(upf_field_spec spec, uint32_t compose, int boxResolution, bool noWorkspace) {
  const float boxSize = 4000;  // sources/meshGeneration.cpp:19-28
  spec.compose = compose;
  upf_grid g = {};
  for (int a = 0; a < 3; a++) {
    g.origin[a] = boxSize * -0.5f;
    g.spacing[a] = boxSize / (boxResolution - 1);
    g.dims[a] = boxResolution;
  }
  hipStream_t st;
  if (hipStreamCreate(&st) != hipSuccess) throw std::runtime_error("hipStreamCreate");
  const size_t nodes = size_t(boxResolution) * boxResolution * boxResolution;
  float* field;
  hipMalloc(&field, nodes * 4);
  const size_t wsBytes = upf_mc_workspace_bytes(&g);
  void* ws;
  hipMalloc(&ws, wsBytes);
  int64_t nv = 0, nt = 0;
  upf_mesh_buffers mb = {};
  if (!noWorkspace) {
    // UPF_ERR_INVALID_VALUE = the reference's !valid() throw
    check(upf_field_mesh_count(&spec, &g, 0.0f, field, nullptr, ws, wsBytes, &nv, &nt, st));
    hipMalloc(&mb.positions, nv * 12);
    hipMalloc(&mb.normals, nv * 12);
    hipMalloc(&mb.indices, nt * 12);
    mb.cap_vertices = mb.n_vertices = nv;
    mb.cap_triangles = mb.n_triangles = nt;
    check(upf_mc_emit(field, &g, 0.0f, &mb, ws, wsBytes, st));
  } else {
    // the allocation-free step: buffers sized from one counting call, then the sweep
    // with the library's scratch
    check(upf_field_mesh_count(&spec, &g, 0.0f, field, nullptr, ws, wsBytes, &nv, &nt, st));
    hipMalloc(&mb.positions, nv * 12);
    hipMalloc(&mb.normals, nv * 12);
    hipMalloc(&mb.indices, nt * 12);
    mb.cap_vertices = mb.n_vertices = nv;
    mb.cap_triangles = mb.n_triangles = nt;
    int64_t* dCounts;
    hipMalloc(&dCounts, 16);
    check(upf_field_mesh_async_static(&spec, &g, 0.0f, field, nullptr, &mb, dCounts, st));
    int64_t counts[2];
    hipMemcpyAsync(counts, dCounts, 16, hipMemcpyDeviceToHost, st);
    hipStreamSynchronize(st);
    hipFree(dCounts);
    if (counts[0] != nv || counts[1] != nt) throw std::runtime_error("static sweep counts differ");
  }
  // meshRemoveDisconnected on the device; libupf's normals and winding already point
  // outward, so meshFlipNormals is not needed
  upf_mesh_buffers kept = {};
  hipMalloc(&kept.positions, nv * 12);
  hipMalloc(&kept.normals, nv * 12);
  hipMalloc(&kept.indices, nt * 12);
  kept.cap_vertices = nv;
  kept.cap_triangles = nt;
  const size_t postBytes = upf_mesh_post_workspace_bytes(nv, nt);
  void* post;
  hipMalloc(&post, postBytes);
  check(upf_mesh_remove_disconnected(&mb, &kept, post, postBytes, st));
  Mesh poly;
  poly.positions.resize(kept.n_vertices * 3);
  poly.normals.resize(kept.n_vertices * 3);
  poly.indices.resize(kept.n_triangles * 3);
  hipMemcpyAsync(poly.positions.data(), kept.positions, kept.n_vertices * 12, hipMemcpyDeviceToHost, st);
  hipMemcpyAsync(poly.normals.data(), kept.normals, kept.n_vertices * 12, hipMemcpyDeviceToHost, st);
  hipMemcpyAsync(poly.indices.data(), kept.indices, kept.n_triangles * 12, hipMemcpyDeviceToHost, st);
  hipStreamSynchronize(st);
  for (void* p : {(void*)field, ws, post, (void*)mb.positions, (void*)mb.normals, (void*)mb.indices,
                  (void*)kept.positions, (void*)kept.normals, (void*)kept.indices})
    hipFree(p);
  hipStreamDestroy(st);
  return poly;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <out.bin> <box resolution> [static]\n", argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[2]);
  const bool noWorkspace = argc > 3 && std::string(argv[3]) == "static";
  try {
    const Mesh m = meshGenerateGeneric(makeSpec(1234), UPF_COMPOSE_LAND, n, noWorkspace);
    if (m.indices.empty()) throw std::runtime_error("generated empty base land mesh");  // meshGeneration.cpp:52-53
    FILE* f = std::fopen(argv[1], "wb");
    if (!f) throw std::runtime_error("cannot open the output file");
    const int64_t counts[2] = {(int64_t)m.positions.size() / 3, (int64_t)m.indices.size() / 3};
    std::fwrite(counts, sizeof(counts), 1, f);
    std::fwrite(m.positions.data(), 4, m.positions.size(), f);
    std::fwrite(m.normals.data(), 4, m.normals.size(), f);
    std::fwrite(m.indices.data(), 4, m.indices.size(), f);
    std::fclose(f);
    std::printf("mesh_generate_generic: %lld vertices, %lld triangles\n", (long long)counts[0], (long long)counts[1]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mesh_generate_generic failed: %s\n", e.what());
    return 1;
  }
  return 0;
}
