// upf_internal.hpp -- host-side helpers shared by the libupf translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/upf.h"
#include "upf_device.hpp"

namespace upf {

// Validates `in` and converts it to the kernel-side DSpec.  Returns a upf_status.
int to_device(const upf_field_spec* in, DSpec* out);

// Validates a grid and converts it.
int to_device(const upf_grid* in, DGrid* out);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Grid evaluation (variant = upf_eval_variant); with S non-null the kernel also writes
// each 64-node group's marching-cubes sign mask (bit = f < iso) at S[(k*ny + j)*gpr + xb].
int field_grid_launch(const DSpec& s, const DGrid& g, float* d_field, int variant, uint64_t* S, float iso,
                      hipStream_t stream);

// Compute units of the current device (cached per device; 256 if the query fails).
int device_cus();

// Workgroups of `threads` threads of `kernel` that are resident on the current device at
// once (the occupancy its registers and LDS allow, times the compute units): the size of
// a persistent grid whose waves each take an equal share, so that no wave starts its
// share after others have finished theirs.
int resident_blocks(const void* kernel, int threads);

// Workgroups are dispatched round-robin over the XCDs (workgroup L runs on XCD L % 8),
// and each XCD has its own L2.  Work item of workgroup L when n items are split into 8
// contiguous ranges, one per XCD: neighbouring items (which share cache lines of their
// inputs and outputs) then run on the same XCD.  Launch kXcds * ceil(n / kXcds)
// workgroups; items >= n exit.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_item(int64_t L, int64_t n) {
  const int64_t per = (n + kXcds - 1) / kXcds;
  return (L % kXcds) * per + L / kXcds;
}
inline int64_t xcd_grid(int64_t n) { return kXcds * ((n + kXcds - 1) / kXcds); }

#define UPF_HIP_CHECK(expr)                      \
  do {                                           \
    if ((expr) != hipSuccess) return UPF_ERR_HIP; \
  } while (0)

#define UPF_LAUNCH_CHECK()                             \
  do {                                                 \
    if (hipGetLastError() != hipSuccess) return UPF_ERR_HIP; \
  } while (0)

}  // namespace upf
