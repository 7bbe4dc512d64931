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

// LDS budget of the per-group shell histograms of one workgroup.
constexpr size_t kMaxGroupBinBytes = 32 * 1024;

// Lanes per halo the count kernel uses for nrads shells when `requested`
// lanes (8, 16 or 64; anything else = automatic) are asked for.
int count_group_for(int nrads, int requested);

// Dynamic LDS of the count kernel for nrads shells at `group` lanes/halo.
size_t count_lds_bytes(int nrads, int group);

// counts[row*nrads + k] (= or +=) the shell-k count of every halo against
// the indexed particles, `group` lanes per halo (64: one halo per
// wavefront; 0: automatic), lanes walking `seg` (4 or 8) particles of a row
// each.  With `inbox` not null, nothing is counted and inbox receives += the
// number of in-box pair tests.  No allocation.
hipError_t count_halos(const IndexView& ix, const HaloView& hv, const Shells& sh,
                       uint64_t* counts, bool accumulate, unsigned long long* inbox,
                       int group, int seg, hipStream_t s);

}  // namespace sjcs
