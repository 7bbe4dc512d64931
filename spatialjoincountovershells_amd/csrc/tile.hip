// tile.hip -- one wavefront per tile of nearby halos, candidates shared.
//
// Reference per pair: the periodic images of each halo
// (src/objects/ArraySTR/MultiRangeCount.c:31-71, src/objects/RTree/Mbr.c:
// 284-313), the closed NaN-permissive box test (Mbr.c:171-182), the FP32
// distance (include/objects/RTree/Distance.h:11-13) and the first-match shell
// (MultiRangeCount.c:107-113).
//
// For each image the wavefront takes the union of the boxes of the tile's
// halos that have that image, cuts every cell row of it to the union of their
// spheres, and deals the rows' particles out 64 at a time (lane = candidate,
// found by a binary search over the rows' prefix sums).  Each candidate is
// then tested against every halo of the tile in turn, with the halo's image
// centre, safe radius and box read from LDS by broadcast; hits go to the
// halo's LDS histogram, flushed to the counts (rows of consecutive halos are
// contiguous) when the tile is done.
#include "tile.h"

#include "pairs.h"
#include "primitives.h"

#pragma clang fp contract(off)

namespace sjcs {
namespace {

constexpr int kTileThreads = 256;
constexpr int kTileWaves = kTileThreads / 64;
constexpr int kHaloWords = 12;  // sx sy sz safe lx hx ly hy lz hz (padded)

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

template <int kShell, bool kInbox>
__global__ __launch_bounds__(kTileThreads) void count_tile_kernel(
    IndexView ix, HaloView hv, Shells sh, TileView tv, uint32_t* __restrict__ counts,
    int accumulate, unsigned long long* __restrict__ inbox_total) {
  extern __shared__ __align__(16) uint32_t smem[];
  const int K = sh.nrads;
  const int kpad = (K + 3) & ~3;
  // LDS: [lut 2*kLutSize][r2 kpad][per wave: halos 64*kHaloWords, rows 64*2,
  // bins kTileBinWords]
  ShellEntry* s_lut = reinterpret_cast<ShellEntry*>(smem);
  float* s_r2 = reinterpret_cast<float*>(smem + 2 * kLutSize);
  for (int i = threadIdx.x; i < K; i += kTileThreads) s_r2[i] = sh.r2[i];
  for (int i = threadIdx.x; i < sh.nlut; i += kTileThreads) s_lut[i] = sh.lut[i];
  __syncthreads();

  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const uint32_t t = blockIdx.x * kTileWaves + wave;
  if (t >= tv.n) return;  // no block-wide barrier follows
  uint32_t* wbase = smem + 2 * kLutSize + kpad + wave * (64 * kHaloWords + 128 + kTileBinWords);
  float* hd = reinterpret_cast<float*>(wbase);
  uint32_t* rows = wbase + 64 * kHaloWords;  // [64] first particle, [64] prefix
  uint32_t* bins = rows + 128;

  const uint32_t h0 = tv.start[t];
  const int nt = (int)tv.len[t];
  const float rbox = sh.rbox;
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
  if (lane < nt) {
    cx = hv.x[h0 + lane];
    cy = hv.y[h0 + lane];
    cz = hv.z[h0 + lane];
  }
  if (!kInbox)
    for (int i = lane; i < nt * K; i += 64) bins[i] = 0u;
  // a NaN centre makes every distance NaN: that halo counts nothing, but its
  // box passes every coordinate on the NaN axes, and the in-box measurement
  // counts those pair tests as the reference makes them
  const bool live = lane < nt && (kInbox || ((cx == cx) && (cy == cy) && (cz == cz)));
  float shift[3] = {0.0f, 0.0f, 0.0f};
  int oob = 0;
  {
    const float c[3] = {cx, cy, cz};
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
  }
  const uint32_t any_oob = (__ballot(live && (oob & 1)) ? 1u : 0u) |
                           (__ballot(live && (oob & 2)) ? 2u : 0u) |
                           (__ballot(live && (oob & 4)) ? 4u : 0u);

  const __amdgpu_buffer_rsrc_t rx = make_rsrc(ix.x, ix.n);
  const __amdgpu_buffer_rsrc_t ry = make_rsrc(ix.y, ix.n);
  const __amdgpu_buffer_rsrc_t rz = make_rsrc(ix.z, ix.n);
  const float inv = ix.grid.inv_cell, invx = ix.grid.inv_cell_x, cs = ix.grid.cell;
  const int G = ix.grid.g, gm1 = G - 1, GX = ix.grid.gx, gxm1 = GX - 1;
  uint64_t inbox = 0;

  for (uint32_t img = 0; img < 8; ++img) {
    if (img & ~any_oob) continue;
    const bool valid = live && !(img & ~(uint32_t)oob);
    const uint64_t vmask = __ballot(valid);
    if (!vmask) continue;
    const float sx = (img & 1) ? cx + shift[0] : cx;
    const float sy = (img & 2) ? cy + shift[1] : cy;
    const float sz = (img & 4) ? cz + shift[2] : cz;
    const float lx = sx - rbox, hx = sx + rbox;
    const float ly = sy - rbox, hy = sy + rbox;
    const float lz = sz - rbox, hz = sz + rbox;
    wave_sync();  // the previous image's readers of hd are done
    if (valid) {
      float* q = hd + lane * kHaloWords;
      q[0] = sx; q[1] = sy; q[2] = sz; q[3] = kInbox ? 0.0f : safe_dist2(sx, sy, sz, rbox);
      q[4] = lx; q[5] = hx; q[6] = ly; q[7] = hy; q[8] = lz; q[9] = hz;
    }
    // union of the valid halos' boxes (cell range) and of their centres
    // (sphere cut); cell_of is monotone, so the union's corner cells bound
    // every member's cells; a NaN edge (measurement only) opens its axis
    const float ulx = wave_min(valid ? (lx == lx ? lx : -INFINITY) : INFINITY);
    const float uhx = wave_max(valid ? (hx == hx ? hx : INFINITY) : -INFINITY);
    const float uly = wave_min(valid ? (ly == ly ? ly : -INFINITY) : INFINITY);
    const float uhy = wave_max(valid ? (hy == hy ? hy : INFINITY) : -INFINITY);
    const float ulz = wave_min(valid ? (lz == lz ? lz : -INFINITY) : INFINITY);
    const float uhz = wave_max(valid ? (hz == hz ? hz : INFINITY) : -INFINITY);
    const float cmnx = wave_min(valid ? sx : INFINITY), cmxx = wave_max(valid ? sx : -INFINITY);
    const float cmny = wave_min(valid ? sy : INFINITY), cmxy = wave_max(valid ? sy : -INFINITY);
    const float cmnz = wave_min(valid ? sz : INFINITY), cmxz = wave_max(valid ? sz : -INFINITY);
    const int x0 = cell_of(ulx, invx, gxm1), x1 = cell_of(uhx, invx, gxm1);
    const int y0 = cell_of(uly, inv, gm1), y1 = cell_of(uhy, inv, gm1);
    const int z0 = cell_of(ulz, inv, gm1), z1 = cell_of(uhz, inv, gm1);
    const int ny = y1 - y0 + 1;
    const int nrows = ny * (z1 - z0 + 1);
    wave_sync();  // hd written

    for (int rb = 0; rb < nrows; rb += 64) {
      const int r = rb + lane;
      uint32_t beg = 0, len = 0;
      if (r < nrows) {
        const int zq = r / ny;
        const int yy = y0 + (r - zq * ny), zz = z0 + zq;
        int xa = x0, xb = x1;
        bool keep = true;
        if (!kInbox) {
          // nearest point of the row's (y, z) cell rectangle (widened by
          // 1e-3, edge cells open) to the centres' bounding box bounds
          // every member's |py - sy|, |pz - sz|; a counted pair of any
          // member has |px - sx| < w
          const float ylo = yy > 0 ? (float)yy * cs - 1e-3f : -INFINITY;
          const float yhi = yy < gm1 ? (float)(yy + 1) * cs + 1e-3f : INFINITY;
          const float zlo = zz > 0 ? (float)zz * cs - 1e-3f : -INFINITY;
          const float zhi = zz < gm1 ? (float)(zz + 1) * cs + 1e-3f : INFINITY;
          const float dy = fmaxf(fmaxf(ylo - cmxy, cmny - yhi), 0.0f);
          const float dz = fmaxf(fmaxf(zlo - cmxz, cmnz - zhi), 0.0f);
          const float rem = sh.r2max_hi - (dy * dy + dz * dz) * 0.99999f;
          keep = rem > 0.0f;
          if (keep) {
            const float w = __builtin_amdgcn_sqrtf(rem) * 1.00001f + 1e-3f;
            xa = max(xa, cell_of(cmnx - w, invx, gxm1));
            xb = min(xb, cell_of(cmxx + w, invx, gxm1));
          }
        }
        if (keep && xa <= xb) {
          const uint32_t kb = ((uint32_t)zz * (uint32_t)G + (uint32_t)yy) * (uint32_t)GX;
          const uint32_t k0 = kb + (uint32_t)xa, k1 = kb + (uint32_t)xb;
          beg = k0 ? ix.cell_end[k0 - 1] : 0u;
          len = ix.cell_end[k1] - beg;
        }
      }
      const uint32_t incl = wave_incl_sum(len, lane);
      const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
      rows[lane] = beg;
      rows[64 + lane] = incl - len;
      wave_sync();
      for (uint32_t c0 = 0; c0 < total; c0 += 64) {
        const uint32_t q = c0 + (uint32_t)lane;
        const bool inq = q < total;
        // the last row whose prefix is <= q holds candidate q
        int j = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
          if (rows[64 + j + step] <= q) j += step;
        const uint32_t idx = rows[j] + (q - rows[64 + j]);
        const uint32_t boff = inq ? idx * 4u : 0xfffffffcu;  // out of range: loads 0
        const float px = buf_load(rx, boff), py = buf_load(ry, boff), pz = buf_load(rz, boff);
        uint64_t hm = vmask;
        while (hm) {
          const int h = __builtin_ctzll(hm);
          hm &= hm - 1;
          const float4 c = *reinterpret_cast<const float4*>(hd + h * kHaloWords);
          if constexpr (kInbox) {
            const float4 b0 = *reinterpret_cast<const float4*>(hd + h * kHaloWords + 4);
            const float2 b1 = *reinterpret_cast<const float2*>(hd + h * kHaloWords + 8);
            const bool in = inq && !(px < b0.x) && !(b0.y < px) && !(py < b0.z) &&
                            !(b0.w < py) && !(pz < b1.x) && !(b1.y < pz);
            inbox += in ? 1 : 0;
          } else {
            const float dist = dpps_distance(px, py, pz, c.x, c.y, c.z);
            bool hit = inq & (dist < sh.r2max);
            if (hit && !(dist < c.w)) {
              const float4 b0 = *reinterpret_cast<const float4*>(hd + h * kHaloWords + 4);
              const float2 b1 = *reinterpret_cast<const float2*>(hd + h * kHaloWords + 8);
              hit = !(px < b0.x) && !(b0.y < px) && !(py < b0.z) && !(b0.w < py) &&
                    !(pz < b1.x) && !(b1.y < pz);
            }
            if (hit) {
              const int o = shell_off<kShell>(dist, s_r2, s_lut, sh);
              atomicAdd(bins + h * K + (o >> 2), 1u);
            }
          }
        }
      }
      wave_sync();  // rows[] is rewritten by the next batch
    }
  }
  if (!kInbox) {
    wave_sync();
    uint32_t* dst = counts + (uint64_t)h0 * K;
    for (int i = lane; i < nt * K; i += 64) {
      const uint32_t v = bins[i];
      if (accumulate) dst[i] += v; else dst[i] = v;
    }
  } else {
    unsigned long long v = inbox;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0 && v) atomicAdd(inbox_total, v);
  }
}

__global__ void permute_halos_kernel(const float* __restrict__ sx, const float* __restrict__ sy,
                                     const float* __restrict__ sz,
                                     const uint32_t* __restrict__ srow, float* __restrict__ dx,
                                     float* __restrict__ dy, float* __restrict__ dz,
                                     uint32_t* __restrict__ drow,
                                     const uint32_t* __restrict__ perm, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t p = perm[i];
    dx[i] = sx[p];
    dy[i] = sy[p];
    dz[i] = sz[p];
    drow[i] = srow[p];
  }
}

}  // namespace

int tile_cap(int nrads) {
  if (nrads <= 0 || nrads > kMaxLdsShells) return 0;
  const int c = kTileBinWords / nrads;
  return c < kTileMaxHalos ? c : kTileMaxHalos;
}

hipError_t count_tiles(const IndexView& ix, const HaloView& hv, const Shells& sh,
                       const TileView& tv, uint32_t* counts, bool accumulate,
                       unsigned long long* inbox, hipStream_t s) {
  if (tv.n == 0 || sh.nrads <= 0) return hipSuccess;
  if (!sh.sorted || tile_cap(sh.nrads) < 1) return hipErrorInvalidValue;
  const bool measure = inbox != nullptr;
  if (ix.n == 0) {
    if (measure) return hipSuccess;
    // nothing to count: a first chunk still stores zeros, launched below
  }
  if (ix.n > 0x3fffffffull) return hipErrorInvalidValue;  // 32-bit byte offsets
  const size_t kpad = (size_t)((sh.nrads + 3) & ~3);
  const size_t lds = 4 * (2 * (size_t)kLutSize + kpad +
                          kTileWaves * (64 * (size_t)kHaloWords + 128 + kTileBinWords));
  const uint32_t blocks = (uint32_t)ceil_div(tv.n, kTileWaves);
  const int acc = accumulate ? 1 : 0;
  if (sh.one_step) {
    if (measure)
      count_tile_kernel<kShellOneStep, true><<<blocks, kTileThreads, lds, s>>>(ix, hv, sh, tv, counts, acc, inbox);
    else
      count_tile_kernel<kShellOneStep, false><<<blocks, kTileThreads, lds, s>>>(ix, hv, sh, tv, counts, acc, inbox);
  } else {
    if (measure)
      count_tile_kernel<kShellTable, true><<<blocks, kTileThreads, lds, s>>>(ix, hv, sh, tv, counts, acc, inbox);
    else
      count_tile_kernel<kShellTable, false><<<blocks, kTileThreads, lds, s>>>(ix, hv, sh, tv, counts, acc, inbox);
  }
  return hipGetLastError();
}

hipError_t permute_halos(const float* sx, const float* sy, const float* sz, const uint32_t* srow,
                         float* dx, float* dy, float* dz, uint32_t* drow, const uint32_t* perm,
                         uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = ceil_div(n, 256);
  if (blocks > 65536) blocks = 65536;
  permute_halos_kernel<<<(uint32_t)blocks, 256, 0, s>>>(sx, sy, sz, srow, dx, dy, dz, drow, perm, n);
  return hipGetLastError();
}

}  // namespace sjcs
