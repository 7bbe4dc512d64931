// count.hip -- per-halo periodic multi-shell count kernels.
//
// Reference: multiRangeCountWithPriodicBoundArraySTR and its recursive
// walk (src/objects/ArraySTR/MultiRangeCount.c:13-115), driven per halo by
// threadJoin (src/astr_rcount/main.c:795-826).
//
// Plain kernel: one wavefront per halo.  For the halo and each periodic
// image of it, the query box's cell range is cut into (y, z) rows whose x
// runs are contiguous in the sorted particle arrays; the wave scans the row
// lengths and its 64 lanes stride over the concatenated candidates with
// coalesced SoA loads.  Every candidate gets the reference's box test,
// FP32 distance and first-match shell, and lands in a per-wave LDS
// histogram that is flushed to the u64 counts when the halo is done.
#include "count.h"

#pragma clang fp contract(off)

namespace sjcs {
namespace {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// First k with d < r2[k], or -1 (src/objects/ArraySTR/MultiRangeCount.c:107-113).
template <bool kSorted>
__device__ __forceinline__ int shell_of(float d, const float* r2, const uint16_t* lut,
                                        const Shells& sh) {
  if (kSorted) {
    if (!(d < sh.r2max)) return -1;  // also rejects NaN
    int b = (int)(__float_as_uint(d) - sh.lut_base);
    b = b < 0 ? 0 : (b >> sh.lut_shift);
    b = b < sh.nlut - 1 ? b : sh.nlut - 1;
    int k = lut[b];
    while (!(d < r2[k])) ++k;  // ends: d < r2[nrads-1] and r2 ascending
    return k;
  } else {
    for (int k = 0; k < sh.nrads; ++k)
      if (d < r2[k]) return k;
    return -1;
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

template <bool kSorted, bool kLds, bool kInbox>
__global__ __launch_bounds__(kCountThreads) void count_plain_kernel(
    IndexView ix, HaloView hv, Shells sh, uint64_t* __restrict__ counts,
    uint32_t halos_per_block, int accumulate, unsigned long long* __restrict__ inbox_total) {
  extern __shared__ __align__(16) uint32_t smem[];
  const int K = sh.nrads;
  const int kpad = (K + 3) & ~3;
  float* s_r2 = reinterpret_cast<float*>(smem);
  uint16_t* s_lut = reinterpret_cast<uint16_t*>(smem + kpad);
  uint32_t* s_rows = kLds ? smem + kpad + kLutSize / 2 : smem;
  uint32_t* s_bins = s_rows + kCountWaves * 128;

  if (kLds) {
    for (int i = threadIdx.x; i < K; i += kCountThreads) s_r2[i] = sh.r2[i];
    if (kSorted)
      for (int i = threadIdx.x; i < sh.nlut; i += kCountThreads) s_lut[i] = sh.lut[i];
  }
  __syncthreads();
  const float* r2 = kLds ? s_r2 : sh.r2;
  const uint16_t* lut = kLds ? s_lut : sh.lut;

  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  uint32_t* rowx = s_rows + wave * 128;
  uint32_t* rowb = rowx + 64;
  uint32_t* bins = s_bins + wave * K;

  // Workgroups go round-robin over the 8 XCDs; give each XCD one contiguous
  // range of the (Morton-sorted) halos so its L2 sees neighbouring queries.
  const uint32_t per = gridDim.x >> 3;
  const uint32_t blk = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint64_t h0 = (uint64_t)blk * halos_per_block;
  uint64_t h1 = h0 + halos_per_block;
  if (h1 > hv.n) h1 = hv.n;

  const float rbox = sh.rbox;
  const float inv = ix.grid.inv_cell;
  const int G = ix.grid.g;
  const int gm1 = G - 1;
  uint64_t inbox = 0;

  for (uint64_t hs = h0 + wave; hs < h1; hs += kCountWaves) {
    const float cx = hv.x[hs], cy = hv.y[hs], cz = hv.z[hs];
    const uint32_t row = hv.orig[hs];
    if (kLds) {
      for (int k = lane; k < K; k += 64) bins[k] = 0u;
      wave_sync();
    }
    // A NaN centre makes every distance NaN: nothing is ever counted.
    const bool live = (cx == cx) && (cy == cy) && (cz == cz);
    if (live) {
      // Periodic images (src/objects/ArraySTR/MultiRangeCount.c:31-71,
      // src/objects/RTree/Mbr.c:284-313): an axis whose base box leaves
      // [0, 1000] is shifted by +1000 (low side wins) or -1000.
      float shift[3] = {0.0f, 0.0f, 0.0f};
      const float c[3] = {cx, cy, cz};
      int oob = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float lo = c[d] - rbox, hi = c[d] + rbox;
        if (lo < kBoxLo) {
          shift[d] = kBoxHi;
          oob |= 1 << d;
        } else if (hi > kBoxHi) {
          shift[d] = -kBoxHi;
          oob |= 1 << d;
        }
      }
      for (int img = 0; img < 8; ++img) {
        if (img & ~oob) continue;
        const float sx = (img & 1) ? cx + shift[0] : cx;
        const float sy = (img & 2) ? cy + shift[1] : cy;
        const float sz = (img & 4) ? cz + shift[2] : cz;
        const float lx = sx - rbox, hx = sx + rbox;
        const float ly = sy - rbox, hy = sy + rbox;
        const float lz = sz - rbox, hz = sz + rbox;
        const int x0 = cell_lo(lx, inv, gm1), x1 = cell_hi(hx, inv, gm1);
        const int y0 = cell_lo(ly, inv, gm1), y1 = cell_hi(hy, inv, gm1);
        const int z0 = cell_lo(lz, inv, gm1), z1 = cell_hi(hz, inv, gm1);
        const int ny = y1 - y0 + 1;
        const int nrows = ny * (z1 - z0 + 1);
        for (int rb = 0; rb < nrows; rb += 64) {
          const int r = rb + lane;
          uint32_t beg = 0, len = 0;
          if (r < nrows) {
            const uint32_t yy = (uint32_t)(y0 + r % ny), zz = (uint32_t)(z0 + r / ny);
            const uint32_t kb = (zz * (uint32_t)G + yy) * (uint32_t)G;
            const uint32_t k0 = kb + (uint32_t)x0, k1 = kb + (uint32_t)x1;
            beg = k0 ? ix.cell_end[k0 - 1] : 0u;
            len = ix.cell_end[k1] - beg;
          }
          const uint32_t incl = wave_incl_sum(len, lane);
          const uint32_t total = __shfl(incl, 63, 64);
          if (total == 0) continue;
          rowx[lane] = incl - len;
          rowb[lane] = beg;
          wave_sync();
          const int nr = nrows - rb < 64 ? nrows - rb : 64;
          for (uint32_t c0 = 0; c0 < total; c0 += 64) {
            const uint32_t i = c0 + (uint32_t)lane;
            if (i < total) {
              int lo = 0, hi = nr;
              while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (rowx[mid] <= i) lo = mid; else hi = mid;
              }
              const uint32_t p = rowb[lo] + (i - rowx[lo]);
              const float px = ix.x[p], py = ix.y[p], pz = ix.z[p];
              // closed, NaN-permissive box (src/objects/RTree/Mbr.c:171-182)
              if (!(px < lx) && !(hx < px) && !(py < ly) && !(hy < py) && !(pz < lz) &&
                  !(hz < pz)) {
                if (kInbox) {  // measurement pass: count the pair test only
                  ++inbox;
                  continue;
                }
                const float dist = dpps_distance(px, py, pz, sx, sy, sz);
                const int k = shell_of<kSorted>(dist, r2, lut, sh);
                if (k >= 0) {
                  if (kLds)
                    atomicAdd(&bins[k], 1u);
                  else
                    atomicAdd(reinterpret_cast<unsigned long long*>(counts) +
                                  (uint64_t)row * K + k, 1ull);
                }
              }
            }
          }
          wave_sync();
        }
      }
    }
    if (kLds && !kInbox) {
      wave_sync();
      uint64_t* dst = counts + (uint64_t)row * K;
      for (int k = lane; k < K; k += 64) {
        const uint64_t v = bins[k];
        if (accumulate) dst[k] += v; else dst[k] = v;
      }
      wave_sync();
    }
  }
  if (kInbox) {
    unsigned long long v = inbox;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0 && v) atomicAdd(inbox_total, v);
  }
}

template <bool kSorted, bool kLds, bool kInbox>
void launch_plain(const IndexView& ix, const HaloView& hv, const Shells& sh,
                  uint64_t* counts, int accumulate, unsigned long long* inbox,
                  uint32_t nblocks, size_t lds, hipStream_t s) {
  count_plain_kernel<kSorted, kLds, kInbox><<<nblocks, kCountThreads, lds, s>>>(
      ix, hv, sh, counts, kHalosPerBlock, accumulate, inbox);
}

}  // namespace

constexpr size_t kNoLdsBytes = 4 * kCountWaves * 128;

size_t count_lds_bytes(int nrads) {
  if (nrads > kMaxLdsShells) return 0;
  const size_t kpad = (size_t)((nrads + 3) & ~3);
  return 4 * (kpad + kLutSize / 2 + kCountWaves * 128 + (size_t)kCountWaves * nrads);
}

hipError_t count_halos(const IndexView& ix, const HaloView& hv, const Shells& sh,
                       uint64_t* counts, bool accumulate, unsigned long long* inbox,
                       hipStream_t s) {
  if (hv.n == 0 || sh.nrads <= 0) return hipSuccess;
  const bool lds = sh.nrads <= kMaxLdsShells;
  const bool measure = inbox != nullptr;
  if (!measure && !accumulate && (!lds || ix.n == 0)) {
    // the HBM-atomic path adds into the counts, and an empty chunk writes
    // nothing: either way a first chunk must start from zero
    hipError_t e = fill_u64(counts, 0ull, hv.n * (uint64_t)sh.nrads, s);
    if (e != hipSuccess) return e;
  }
  if (ix.n == 0) return hipSuccess;
  uint64_t blocks = ceil_div(hv.n, kHalosPerBlock);
  blocks = (blocks + 7) & ~7ull;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  const size_t ldsb = count_lds_bytes(sh.nrads);
  const int acc = accumulate ? 1 : 0;
  const uint32_t nb = (uint32_t)blocks;
  const bool srt = sh.sorted != 0;
  const bool in = inbox != nullptr;
  if (srt && lds && !in) launch_plain<true, true, false>(ix, hv, sh, counts, acc, inbox, nb, ldsb, s);
  else if (srt && lds && in) launch_plain<true, true, true>(ix, hv, sh, counts, acc, inbox, nb, ldsb, s);
  else if (!srt && lds && !in) launch_plain<false, true, false>(ix, hv, sh, counts, acc, inbox, nb, ldsb, s);
  else if (!srt && lds && in) launch_plain<false, true, true>(ix, hv, sh, counts, acc, inbox, nb, ldsb, s);
  else if (srt && !lds && !in) launch_plain<true, false, false>(ix, hv, sh, counts, 1, inbox, nb, kNoLdsBytes, s);
  else if (srt && !lds && in) launch_plain<true, false, true>(ix, hv, sh, counts, 1, inbox, nb, kNoLdsBytes, s);
  else if (!srt && !lds && !in) launch_plain<false, false, false>(ix, hv, sh, counts, 1, inbox, nb, kNoLdsBytes, s);
  else launch_plain<false, false, true>(ix, hv, sh, counts, 1, inbox, nb, kNoLdsBytes, s);
  return hipGetLastError();
}

}  // namespace sjcs
