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

#define UPF_HIP_CHECK(expr)                      \
  do {                                           \
    if ((expr) != hipSuccess) return UPF_ERR_HIP; \
  } while (0)

#define UPF_LAUNCH_CHECK()                             \
  do {                                                 \
    if (hipGetLastError() != hipSuccess) return UPF_ERR_HIP; \
  } while (0)

}  // namespace upf
