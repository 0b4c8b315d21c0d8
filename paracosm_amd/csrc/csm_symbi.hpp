// SymBi index (DCS/DCG counters and d1/d2 bitmaps) on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "csm_kernels.hpp"

namespace csm {

struct SymbiIndex {
  uint32_t built = 0;
};

}  // namespace csm
