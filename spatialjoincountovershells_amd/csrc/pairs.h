// pairs.h -- device helpers shared by the count kernels: the reference's
// distance and first-match shell, the safe radius below which the box test
// is implied, bounds-checked particle loads.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

#pragma clang fp contract(off)

namespace sjcs {

constexpr int kShellScan = 0;     // any thresholds: first match by linear scan
constexpr int kShellTable = 1;    // sorted: table lower bound + upward scan
constexpr int kShellOneStep = 2;  // sorted, one threshold per bucket at most

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// First k with d < r2[k] (src/objects/ArraySTR/MultiRangeCount.c:107-113),
// as the byte offset 4k of its histogram bin, for a d < r2[nrads-1] (the
// caller's gate); the scan mode takes any d and returns -1 when no shell
// matches.
template <int kShell>
__device__ __forceinline__ int shell_off(float d, const float* r2, const ShellEntry* lut,
                                         const Shells& sh) {
  if constexpr (kShell == kShellScan) {
    for (int k = 0; k < sh.nrads; ++k)
      if (d < r2[k]) return 4 * k;
    return -1;
  } else {
    // d < r2max bounds the bucket above; bits below the base clamp to 0
    const int b = max((int)(__float_as_uint(d) - sh.lut_base) >> sh.lut_shift, 0);
    const ShellEntry e = lut[b];
    int off = (int)e.off + (d < e.t ? 0 : 4);
    if constexpr (kShell == kShellTable)
      while (!(d < r2[off >> 2])) off += 4;  // ends: d < r2[nrads-1], r2 ascending
    return off;
  }
}

// FP32 distance exactly as DPPS 0x77 rounds it (include/objects/RTree/Distance.h:11-13).
__device__ __forceinline__ float dpps_distance(float px, float py, float pz, float cx,
                                               float cy, float cz) {
  const float dx = px - cx;
  const float dy = py - cy;
  const float dz = pz - cz;
  const float xx = dx * dx;
  const float yy = dy * dy;
  const float zz = dz * dz;
  const float sxy = xx + yy;
  return sxy + zz;
}

// gfx9 buffer descriptor word 3 (raw buffer, 32-bit data); out-of-range
// loads return 0, so lanes past the end need no branch.
constexpr int kBufferFlags = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const float* p, uint64_t n) {
  const uint64_t bytes = n * 4;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0,
                                           (int)(bytes < 0xffffffffull ? bytes : 0xffffffffull),
                                           kBufferFlags);
}

__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
}

// Below the returned value an FP32 distance proves the closed box test of
// the query box [fl(s - rbox), fl(s + rbox)] (src/objects/RTree/Mbr.c:
// 171-182): with m = max(|sx|, |sy|, |sz|) and A = rbox - (m + rbox) 2^-21,
// dist < A^2 (1 - 2^-19) gives |p - s| < A (1 + 2^-24) on every axis (each
// squared term is at most the FP32 sum), which keeps p inside the box with
// room for the roundings of the box edges and of p - s.  A non-finite
// centre gets 0: every pair is box-tested.
__device__ __forceinline__ float safe_dist2(float sx, float sy, float sz, float rbox) {
  const float m = fmaxf(fmaxf(fabsf(sx), fabsf(sy)), fabsf(sz));
  const float a = rbox - (m + rbox) * 0x1p-21f;
  return a > 0.0f ? (a * a) * (1.0f - 0x1p-19f) : 0.0f;
}

}  // namespace sjcs
