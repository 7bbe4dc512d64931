// count.h -- launchers of the shell-count kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "primitives.h"

namespace sjcs {

constexpr int kCountThreads = 256;
constexpr int kCountWaves = kCountThreads / 64;
constexpr uint32_t kHalosPerBlock = 64;

// Dynamic LDS the plain kernel needs for nrads shells (0 when the shells do
// not fit and the kernel counts straight into HBM).
size_t count_lds_bytes(int nrads);

// counts[row*nrads + k] (= or +=) the shell-k count of every halo against
// the indexed particles.  `inbox`, when not null, receives += the number of
// in-box pair tests.  No allocation; everything goes to stream s.
hipError_t count_halos(const IndexView& ix, const HaloView& hv, const Shells& sh,
                       uint64_t* counts, bool accumulate, unsigned long long* inbox,
                       hipStream_t s);

}  // namespace sjcs
