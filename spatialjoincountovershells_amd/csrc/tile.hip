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
// spheres, and deals the rows' particles out in segments of eight
// consecutive particles, 64 segments at a time (lane = segment, its row
// found by a binary search over the rows' segment prefix sums).  Each
// candidate is then tested against every halo of the tile in turn, with the halo's image
// centre, safe radius and box read from LDS by broadcast; hits go to the
// halo's LDS histogram, flushed to the counts (rows of consecutive halos are
// contiguous) when the tile is done.  Halos of a tile at the very same
// position have the very same counts: one of them is evaluated and its
// histogram written to all of them (the in-box measurement weighs its pair
// tests by their number).
//
// Tiles differ in cost by orders of magnitude (a halo at a big clump sees
// 1e5 particles more than one in the field).  A pre-pass estimates every
// tile's pair tests and cuts heavy tiles into parts (a part takes every
// nparts-th round of 64 segments and adds its histograms atomically); the
// count is a persistent kernel whose wavefronts take work items (tile, part)
// from a queue until it is empty.
#include "tile.h"

#include <type_traits>

#include "pairs.h"
#include "primitives.h"

#pragma clang fp contract(off)

namespace sjcs {
namespace {

constexpr int kHaloWords = 12;  // sx sy sz safe lx hx ly hy lz hz mult (padded)
constexpr int kWaveWords = 64 * kHaloWords + 192 + 64 + 64 + kTileBinWords;
constexpr int kTileSeg = 8;     // consecutive particles a lane takes per round
constexpr uint64_t kItemPairs = 1ull << 22;  // pair tests a work item aims at
constexpr uint32_t kMaxParts = 256;
constexpr uint32_t kTileGrid = 3072;         // persistent wavefronts
constexpr int kHaloBatch = 8;   // halos whose LDS reads go out together

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

// Pre-pass, one wavefront per tile: an estimate of the tile's pair tests
// (its halos times the particles in the rows of the union of their base
// boxes; the lanes take the halos, then the rows), the number of parts it
// is cut into (written to item[t]; item[n] = 0 for the scan that follows),
// and, when the counts are stored rather than added, zeros in the rows of
// tiles of several parts (their parts add atomically).
constexpr int kCostWaves = 4;

__global__ __launch_bounds__(64 * kCostWaves) void tile_cost_kernel(
    IndexView ix, HaloView hv, Shells sh, TileView tv, uint32_t* __restrict__ item,
    uint32_t* __restrict__ counts, int zero_multi) {
  const uint32_t t = blockIdx.x * kCostWaves + (threadIdx.x >> 6);  // same on the wave
  const int lane = threadIdx.x & 63;
  if (t > tv.n) return;
  if (t == tv.n) {
    if (lane == 0) item[t] = 0;
    return;
  }
  const uint32_t h0 = tv.start[t], nt = tv.len[t];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = lane; i < nt; i += 64) {
    const float c[3] = {hv.x[h0 + i], hv.y[h0 + i], hv.z[h0 + i]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (c[d] == c[d]) {  // a NaN centre counts nothing
        lo[d] = fminf(lo[d], c[d] - sh.rbox);
        hi[d] = fmaxf(hi[d], c[d] + sh.rbox);
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = wave_min(lo[d]);
    hi[d] = wave_max(hi[d]);
  }
  uint64_t cand = 0;
  if (lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2] && ix.n) {
    const int gm1 = ix.grid.g - 1, gxm1 = ix.grid.gx - 1;
    const int x0 = cell_of(lo[0], ix.grid.inv_cell_x, gxm1), x1 = cell_of(hi[0], ix.grid.inv_cell_x, gxm1);
    const int y0 = cell_of(lo[1], ix.grid.inv_cell, gm1), y1 = cell_of(hi[1], ix.grid.inv_cell, gm1);
    const int z0 = cell_of(lo[2], ix.grid.inv_cell, gm1), z1 = cell_of(hi[2], ix.grid.inv_cell, gm1);
    const int ny = y1 - y0 + 1, nrows = ny * (z1 - z0 + 1);
    for (int r = lane; r < nrows; r += 64) {
      const int zz = z0 + r / ny, yy = y0 + r % ny;
      const uint32_t kb = ((uint32_t)zz * (uint32_t)ix.grid.g + (uint32_t)yy) * (uint32_t)ix.grid.gx;
      const uint32_t k0 = kb + (uint32_t)x0, k1 = kb + (uint32_t)x1;
      cand += ix.cell_end[k1] - (k0 ? ix.cell_end[k0 - 1] : 0u);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cand += __shfl_xor(cand, off, 64);
  uint64_t parts = (cand * nt + kItemPairs - 1) / kItemPairs;
  if (parts < 1) parts = 1;
  if (parts > kMaxParts) parts = kMaxParts;
  if (lane == 0) item[t] = (uint32_t)parts;
  if (zero_multi && parts > 1) {
    for (uint32_t i = 0; i < nt; ++i) {
      uint32_t* dst = counts + (uint64_t)(hv.slot ? hv.slot[h0 + i] : h0 + i) * sh.nrads;
      for (int k = lane; k < sh.nrads; k += 64) dst[k] = 0u;
    }
  }
}

template <int kShell, bool kInbox>
__global__ __launch_bounds__(64) void count_tile_kernel(
    IndexView ix, HaloView hv, Shells sh, TileView tv, const uint32_t* __restrict__ item_start,
    uint32_t* __restrict__ queue, uint32_t* __restrict__ counts, int accumulate,
    unsigned long long* __restrict__ inbox_total) {
  extern __shared__ __align__(16) uint32_t smem[];
  const int K = sh.nrads;
  const int kpad = (K + 3) & ~3;
  // LDS: [r2 kpad][shell table, nlut entries rounded to even][halos
  // 64*kHaloWords][rows 64*3][valid-halo list 64][representatives 64][bins].
  // The histograms are sized for the launch's longest tile, so the table's
  // copy (at most 8 KB) costs few wavefronts per CU and saves each hit an
  // L1 round trip.
  float* s_r2 = reinterpret_cast<float*>(smem);
  for (int i = threadIdx.x; i < K; i += 64) s_r2[i] = sh.r2[i];
  ShellEntry* s_lut = reinterpret_cast<ShellEntry*>(smem + kpad);
  for (int i = threadIdx.x; i < sh.nlut; i += 64) s_lut[i] = sh.lut[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  uint32_t* wbase = smem + kpad + 2 * ((sh.nlut + 1) & ~1);
  float* hd = reinterpret_cast<float*>(wbase);
  // per row j of a batch: [64] base (segment q starts at base + q * kTileSeg),
  // [64] first segment, [64] end particle
  uint32_t* rows = wbase + 64 * kHaloWords;
  uint32_t* hl = rows + 192;                 // the image's valid halos
  uint32_t* s_rep = hl + 64;                 // first halo of the tile at the same spot
  uint32_t* bins = s_rep + 64;
  const uint32_t nitems = item_start[tv.n];
  uint64_t inbox = 0;

  for (;;) {
    // next work item: tile t, part p of np
    uint32_t it = 0;
    if (lane == 0) it = atomicAdd(queue, 1u);
    it = __builtin_amdgcn_readfirstlane((uint32_t)__shfl((int)it, 0, 64));
    if (it >= nitems) break;
    uint32_t t = 0;
    for (uint32_t lo = 0, hi = tv.n; lo < hi;) {  // the last t with item_start[t] <= it
      const uint32_t mid = (lo + hi) >> 1;
      if (item_start[mid] <= it) { t = mid; lo = mid + 1; } else { hi = mid; }
    }
    const uint32_t p = it - item_start[t];
    const uint32_t np = item_start[t + 1] - item_start[t];

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
    // the first live halo of the tile at this halo's position (NaN never
    // matches), and how many halos each such representative stands for
    int rep = lane;
    for (int j = 0; j < nt; ++j) {
      const float xj = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, cx), j));
      const float yj = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, cy), j));
      const float zj = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, cz), j));
      const bool lj = __builtin_amdgcn_readlane((int)live, j) != 0;
      if (rep == lane && j < lane && lj && live && xj == cx && yj == cy && zj == cz) rep = j;
    }
    s_rep[lane] = (uint32_t)rep;
    hl[lane] = 0u;
    wave_sync();
    if (live) atomicAdd(&hl[rep], 1u);
    wave_sync();
    const uint32_t mult = hl[lane];
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

    for (uint32_t img = 0; img < 8; ++img) {
      if (img & ~any_oob) continue;
      const bool valid = live && rep == lane && !(img & ~(uint32_t)oob);
      const uint64_t vmask = __ballot(valid);
      if (!vmask) continue;
      const float sx = (img & 1) ? cx + shift[0] : cx;
      const float sy = (img & 2) ? cy + shift[1] : cy;
      const float sz = (img & 4) ? cz + shift[2] : cz;
      const float lx = sx - rbox, hx = sx + rbox;
      const float ly = sy - rbox, hy = sy + rbox;
      const float lz = sz - rbox, hz = sz + rbox;
      wave_sync();  // the previous image's readers of hd are done
      const int nv = __popcll(vmask);
      if (valid) {
        hl[__builtin_amdgcn_mbcnt_hi((uint32_t)(vmask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)vmask, 0u))] = lane;
        float* q = hd + lane * kHaloWords;
        q[0] = sx; q[1] = sy; q[2] = sz; q[3] = kInbox ? 0.0f : safe_dist2(sx, sy, sz, rbox);
        q[4] = lx; q[5] = hx; q[6] = ly; q[7] = hy; q[8] = lz; q[9] = hz;
        q[10] = __builtin_bit_cast(float, mult);
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
        // rows cut into segments of kTileSeg consecutive particles, dealt
        // out 64 per round (a part takes every np-th round)
        const uint32_t nseg = (len + kTileSeg - 1) / kTileSeg;
        const uint32_t sincl = wave_incl_sum(nseg, lane);
        const uint32_t stotal = (uint32_t)__builtin_amdgcn_readlane((int)sincl, 63);
        const uint32_t sexcl = sincl - nseg;
        rows[lane] = beg - sexcl * (uint32_t)kTileSeg;  // wraps modulo 2^32; the sums do not
        rows[64 + lane] = sexcl;
        rows[128 + lane] = beg + len;
        wave_sync();
        for (uint32_t c0 = p * 64u; c0 < stotal; c0 += np * 64u) {
          const uint32_t q = c0 + (uint32_t)lane;
          // the last row whose first segment is <= q holds segment q (an
          // empty row shares its first segment with the row after it)
          int j = 0;
#pragma unroll
          for (int step = 32; step > 0; step >>= 1)
            if (rows[64 + j + step] <= q) j += step;
          const uint32_t start = rows[j] + q * (uint32_t)kTileSeg;
          const uint32_t slen = q < stotal ? min((uint32_t)kTileSeg, rows[128 + j] - start) : 0u;
          // a slot past the segments loads whatever its offset names
          // (bounds-checked by the buffer descriptor) and never uses it
          const uint32_t boff = start * 4u;
          float px[kTileSeg], py[kTileSeg], pz[kTileSeg];
#pragma unroll
          for (int u = 0; u < kTileSeg; ++u) {
            px[u] = buf_load(rx, boff + 4u * u);
            py[u] = buf_load(ry, boff + 4u * u);
            pz[u] = buf_load(rz, boff + 4u * u);
          }
          // the valid halos, NB at a time (8 while at least 8 are left, then
          // 4, 2, 1: nv is the same on every lane, and a tile of copies of
          // one position has a single valid halo): their LDS reads go out
          // together once per round, then for each of the lane's candidates
          // the distances, the shell-table reads and the bin increments
          auto batch = [&](auto nbc, int hb) {
            constexpr int NB = decltype(nbc)::value;
            int hh[NB];
            float4 c[NB];
#pragma unroll
            for (int u = 0; u < NB; ++u) hh[u] = (int)hl[hb + u];
#pragma unroll
            for (int u = 0; u < NB; ++u)
              c[u] = *reinterpret_cast<const float4*>(hd + hh[u] * kHaloWords);
#pragma unroll
            for (int k = 0; k < kTileSeg; ++k) {
              const bool inq = (uint32_t)k < slen;
              if constexpr (kInbox) {
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                  const float4 b0 = *reinterpret_cast<const float4*>(hd + hh[u] * kHaloWords + 4);
                  const float4 b1 = *reinterpret_cast<const float4*>(hd + hh[u] * kHaloWords + 8);
                  const bool in = inq && !(px[k] < b0.x) && !(b0.y < px[k]) && !(py[k] < b0.z) &&
                                  !(b0.w < py[k]) && !(pz[k] < b1.x) && !(b1.y < pz[k]);
                  inbox += in ? __builtin_bit_cast(uint32_t, b1.z) : 0u;
                }
              } else {
                float dist[NB];
                bool hit[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                  dist[u] = dpps_distance(px[k], py[k], pz[k], c[u].x, c[u].y, c[u].z);
                  hit[u] = inq & (dist[u] < sh.r2max);
                  if (hit[u] && !(dist[u] < c[u].w)) {
                    const float4 b0 = *reinterpret_cast<const float4*>(hd + hh[u] * kHaloWords + 4);
                    const float2 b1 = *reinterpret_cast<const float2*>(hd + hh[u] * kHaloWords + 8);
                    hit[u] = !(px[k] < b0.x) && !(b0.y < px[k]) && !(py[k] < b0.z) &&
                             !(b0.w < py[k]) && !(pz[k] < b1.x) && !(b1.y < pz[k]);
                  }
                }
                ShellEntry e[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                  const int b = (int)(__float_as_uint(dist[u]) - sh.lut_base) >> sh.lut_shift;
                  e[u] = s_lut[hit[u] ? max(b, 0) : 0];
                }
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                  if (hit[u]) {
                    int o = (int)e[u].off + (dist[u] < e[u].t ? 0 : 4);
                    if constexpr (kShell == kShellTable)
                      while (!(dist[u] < s_r2[o >> 2])) o += 4;  // ends: dist < r2[nrads-1]
                    atomicAdd(bins + hh[u] * K + (o >> 2), 1u);
                  }
                }
              }
            }
          };
          int hb = 0;
          for (; hb + kHaloBatch <= nv; hb += kHaloBatch)
            batch(std::integral_constant<int, kHaloBatch>{}, hb);
          if (nv - hb >= 4) {
            batch(std::integral_constant<int, 4>{}, hb);
            hb += 4;
          }
          if (nv - hb >= 2) {
            batch(std::integral_constant<int, 2>{}, hb);
            hb += 2;
          }
          if (nv - hb >= 1) batch(std::integral_constant<int, 1>{}, hb);
        }
        wave_sync();  // rows[] is rewritten by the next batch
      }
    }
    if (!kInbox) {
      wave_sync();
      for (int h = 0; h < nt; ++h) {
        uint32_t* dst = counts + (uint64_t)(hv.slot ? hv.slot[h0 + h] : h0 + h) * K;
        const uint32_t* src = bins + s_rep[h] * K;
        for (int k = lane; k < K; k += 64) {
          const uint32_t v = src[k];
          if (np > 1) {
            if (v) atomicAdd(dst + k, v);  // parts share the rows (zeroed when storing)
          } else if (accumulate) {
            dst[k] += v;
          } else {
            dst[k] = v;
          }
        }
      }
      wave_sync();  // bins are cleared by the next item
    }
  }  // work items
  if (kInbox) {
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
    if (drow) drow[i] = srow[p];
  }
}

}  // namespace

int tile_cap(int nrads) {
  if (nrads <= 0 || nrads > kMaxLdsShells) return 0;
  const int c = kTileBinWords / nrads;
  return c < kTileMaxHalos ? c : kTileMaxHalos;
}

uint64_t tile_work_words(uint32_t ntiles) {
  return (uint64_t)ntiles + 1 + 64 + scan_workspace_words((uint64_t)ntiles + 1);
}

hipError_t count_tiles(const IndexView& ix, const HaloView& hv, const Shells& sh,
                       const TileView& tv, uint32_t* counts, bool accumulate,
                       unsigned long long* inbox, uint32_t* work, hipStream_t s) {
  if (tv.n == 0 || sh.nrads <= 0) return hipSuccess;
  if (!sh.sorted || tile_cap(sh.nrads) < 1) return hipErrorInvalidValue;
  const bool measure = inbox != nullptr;
  if (measure && ix.n == 0) return hipSuccess;
  if (ix.n > 0x3fffffffull) return hipErrorInvalidValue;  // 32-bit byte offsets
  uint32_t* item = work;                       // ntiles + 1
  uint32_t* queue = work + tv.n + 1;           // 1 (own 256-byte line)
  uint32_t* scan_ws = queue + 64;
  tile_cost_kernel<<<(tv.n + 1 + kCostWaves - 1) / kCostWaves, 64 * kCostWaves, 0, s>>>(
      ix, hv, sh, tv, item, counts, (!measure && !accumulate) ? 1 : 0);
  hipError_t e = exclusive_sum_u32(item, item, (uint64_t)tv.n + 1, scan_ws, s);
  if (e == hipSuccess) e = fill_u32(queue, 0u, 1, s);
  if (e != hipSuccess) return e;
  // histograms for the longest tile only: shorter tiles leave LDS for more
  // wavefronts per CU
  if (tv.max_len < 1 || tv.max_len > (uint32_t)tile_cap(sh.nrads)) return hipErrorInvalidValue;
  const size_t kpad = (size_t)((sh.nrads + 3) & ~3);
  const size_t bin_words = ((size_t)tv.max_len * (size_t)sh.nrads + 3) & ~(size_t)3;
  const size_t lut_words = 2 * (size_t)((sh.nlut + 1) & ~1);
  const size_t lds = 4 * (kpad + lut_words + (size_t)kWaveWords - kTileBinWords + bin_words);
  const int acc = accumulate ? 1 : 0;
  if (sh.one_step) {
    if (measure)
      count_tile_kernel<kShellOneStep, true><<<kTileGrid, 64, lds, s>>>(ix, hv, sh, tv, item, queue, counts, acc, inbox);
    else
      count_tile_kernel<kShellOneStep, false><<<kTileGrid, 64, lds, s>>>(ix, hv, sh, tv, item, queue, counts, acc, inbox);
  } else {
    if (measure)
      count_tile_kernel<kShellTable, true><<<kTileGrid, 64, lds, s>>>(ix, hv, sh, tv, item, queue, counts, acc, inbox);
    else
      count_tile_kernel<kShellTable, false><<<kTileGrid, 64, lds, s>>>(ix, hv, sh, tv, item, queue, counts, acc, inbox);
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
