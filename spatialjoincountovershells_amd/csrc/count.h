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

// Particle staging in the count kernel: a workgroup stages at most
// stage_cap particles of the cell rows of its halos' union box, from at
// most kStageRows rows.  count_stage_bytes: the LDS it adds.
constexpr int kStageRows = kCountThreads;
constexpr uint32_t kStageCapMax = 8192;
constexpr size_t kMaxCountLds = 160 * 1024;
size_t count_stage_bytes(uint32_t stage_cap);

// counts[i*nrads + k] (= or +=) the shell-k count of sorted halo i against
// the indexed particles (u32: the caller keeps a running total below 2^32,
// one particle counting at most once per periodic image), `group` lanes per
// halo (64: one halo per wavefront; 0: automatic), lanes walking `seg` (4 or
// 8) particles of a row each.  With `inbox` not null, nothing is counted and
// inbox receives += the number of in-box pair tests.  stage_cap > 0: the
// count pass of 8 or 16 lanes per halo stages particles in LDS (at most
// stage_cap per workgroup of stage_halos halos, 0 = 64).  No allocation.
hipError_t count_halos(const IndexView& ix, const HaloView& hv, const Shells& sh,
                       uint32_t* counts, bool accumulate, unsigned long long* inbox,
                       int group, int seg, hipStream_t s, uint32_t stage_cap = 0,
                       uint32_t stage_halos = 0);

// c64[i] = (store ? 0 : c64[i]) + c32[i] for i < total.
hipError_t widen_counts(const uint32_t* c32, uint64_t* c64, uint64_t total, bool store,
                        hipStream_t s);

// inv[orig[i]] = i for i < n (orig a permutation of [0, n)).
hipError_t invert_rows(const uint32_t* orig, uint32_t* inv, uint64_t n, hipStream_t s);

// dst[r*K + k] = c64[inv[r]*K + k] + c32[inv[r]*K + k] (a null source adds
// 0): counts from sorted halo order back to the halos' input order, inv[r]
// being input row r's sorted position (invert_rows).  Requires K*64 < 2^32.
hipError_t readout_counts(const uint32_t* c32, const uint64_t* c64, const uint32_t* inv,
                          uint64_t* dst, uint64_t n, int K, hipStream_t s);
// The same rows from the u32 counts alone, as u32.
hipError_t readout_counts32(const uint32_t* c32, const uint32_t* inv, uint32_t* dst, uint64_t n,
                            int K, hipStream_t s);

}  // namespace sjcs
