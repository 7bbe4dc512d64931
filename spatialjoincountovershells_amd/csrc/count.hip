// count.hip -- per-halo periodic multi-shell count kernels.
//
// Reference: multiRangeCountWithPriodicBoundArraySTR and its recursive
// walk (src/objects/ArraySTR/MultiRangeCount.c:13-115), driven per halo by
// threadJoin (src/astr_rcount/main.c:795-826).
//
// One halo per GROUP lanes of a wavefront (GROUP 64: the plain
// one-halo-per-wavefront kernel; 16 or 8: four or eight halos in flight per
// wavefront).
// For the halo and each of its periodic images the query box's cell range is
// cut into (y, z) rows, whose x runs are contiguous in the sorted particle
// arrays.  Row r of a batch belongs to lane r of the group.  Rows are cut
// into segments of SEG consecutive particles; segments are dealt out GROUP
// at a time, a lane finds the row of its segment from a mark + max-scan over
// the group (one LDS round trip), then walks the segment's particles with
// immediate-offset buffer loads.  Each candidate gets the reference's box
// test, FP32 distance and first-match shell, and lands in the group's LDS
// histogram, flushed to the u64 counts when the halo is done.
#include "count.h"
#include "pairs.h"

#pragma clang fp contract(off)

namespace sjcs {
namespace {

// DPP row_shr:n within each row of 16 lanes; lanes whose source falls before
// the row start read 0.
template <int N>
__device__ __forceinline__ uint32_t row_shr(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x110 | N, 0xf, 0xf, true);
}

// DPP row_bcast:15 (lane 15 of each row to the next row: rows 1 and 3
// enabled) and row_bcast:31 (lane 31 to rows 2 and 3); disabled rows read 0.
__device__ __forceinline__ uint32_t row_bcast15(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
}
__device__ __forceinline__ uint32_t row_bcast31(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
}

// Inclusive sum / max over the GROUP lanes of a group (values >= 0), in
// registers: DPP shifts within rows of 16 lanes, then (GROUP 64) the row
// broadcasts carry each row's total into the rows above.  For GROUP 8 a DPP
// row holds two groups, so a step that would reach into the group below is
// masked.
template <int GROUP>
__device__ __forceinline__ uint32_t grp_scan_add(uint32_t v, int gl) {
  if constexpr (GROUP == 16 || GROUP == 64) {
    v += row_shr<1>(v);
    v += row_shr<2>(v);
    v += row_shr<4>(v);
    v += row_shr<8>(v);
    if constexpr (GROUP == 64) {
      v += row_bcast15(v);
      v += row_bcast31(v);
    }
  } else if constexpr (GROUP == 8) {
    uint32_t t = row_shr<1>(v);
    v += gl >= 1 ? t : 0u;
    t = row_shr<2>(v);
    v += gl >= 2 ? t : 0u;
    t = row_shr<4>(v);
    v += gl >= 4 ? t : 0u;
  } else {
#pragma unroll
    for (int off = 1; off < GROUP; off <<= 1) {
      uint32_t t = __shfl_up(v, off, GROUP);
      if (gl >= off) v += t;
    }
  }
  return v;
}

template <int GROUP>
__device__ __forceinline__ uint32_t grp_scan_max(uint32_t v, int gl) {
  if constexpr (GROUP == 16 || GROUP == 64) {
    v = max(v, row_shr<1>(v));
    v = max(v, row_shr<2>(v));
    v = max(v, row_shr<4>(v));
    v = max(v, row_shr<8>(v));
    if constexpr (GROUP == 64) {
      v = max(v, row_bcast15(v));
      v = max(v, row_bcast31(v));
    }
  } else if constexpr (GROUP == 8) {
    uint32_t t = row_shr<1>(v);
    v = max(v, gl >= 1 ? t : 0u);
    t = row_shr<2>(v);
    v = max(v, gl >= 2 ? t : 0u);
    t = row_shr<4>(v);
    v = max(v, gl >= 4 ? t : 0u);
  } else {
#pragma unroll
    for (int off = 1; off < GROUP; off <<= 1) {
      uint32_t t = __shfl_up(v, off, GROUP);
      if (gl >= off) v = max(v, t);
    }
  }
  return v;
}

// Value of lane j (group-relative) of this lane's group.
template <int GROUP>
__device__ __forceinline__ uint32_t grp_lane(uint32_t v, uint32_t j) {
  return (uint32_t)__shfl((int)v, (int)j, GROUP);
}

// Value of the last lane of this lane's group (an inclusive scan's total):
// 16 lanes by a DPP row_newbcast:15, 64 lanes by a lane read.
template <int GROUP>
__device__ __forceinline__ uint32_t grp_last(uint32_t v) {
  if constexpr (GROUP == 16)
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x15F, 0xf, 0xf, false);
  else if constexpr (GROUP == 64)
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
  else
    return grp_lane<GROUP>(v, GROUP - 1);
}

// Max over the wave of a value that is uniform within each group, read from
// each group's first lane (a scalar: loops bounded by it stay uniform).
template <int GROUP>
__device__ __forceinline__ uint32_t wave_max_of_groups(uint32_t v) {
  uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
#pragma unroll
  for (int g = GROUP; g < 64; g += GROUP)
    m = max(m, (uint32_t)__builtin_amdgcn_readlane((int)v, g));
  return m;
}

// Inclusive sum over a workgroup of kCountThreads, one value per thread
// (s_tmp: kCountWaves words); *total receives the sum of all.
__device__ __forceinline__ uint32_t block_incl_sum(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)v, off, 64);
    if (lane >= off) v += t;
  }
  if (lane == 63) s_tmp[wave] = v;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kCountWaves; ++w) {
    const uint32_t t = s_tmp[w];
    before += w < wave ? t : 0u;
    all += t;
  }
  __syncthreads();
  *total = all;
  return v + before;
}

template <int GROUP, int SEG, int kShell, bool kLds, bool kInbox, bool kStage>
__global__ __launch_bounds__(kCountThreads) void count_group_kernel(
    IndexView ix, HaloView hv, Shells sh, uint32_t* __restrict__ counts,
    uint32_t halos_per_block, int accumulate, unsigned long long* __restrict__ inbox_total,
    uint32_t stage_cap, uint32_t copies, uint32_t ks) {
  constexpr int H = 64 / GROUP;
  // Sorted shells: thresholds are finite or +inf, so a live halo's boxes are
  // NaN-free and a counted pair needs dist < r2max; each row's x run is cut
  // to the sphere's chord (with margins), which drops no counted pair.  The
  // in-box measurement and the any-threshold scan walk the whole box.
  constexpr bool kSphere = kShell != kShellScan && !kInbox;
  extern __shared__ __align__(16) uint32_t smem[];
  const int K = sh.nrads;
  const int kpad = (K + 3) & ~3;
  // LDS: [lut 2*kLutSize][r2 kpad][marks 4 waves x 128][bins 4 waves x H x
  // copies x ks].  A halo's histogram is kept in `copies` interleaved copies
  // (a power of two; lane gl adds into copy gl % copies), ks words apart: ks
  // is odd when copies > 1, so one shell's copies sit in different LDS
  // banks.  With 64 lanes on one halo the hits of a round crowd into its
  // outer shells, and one copy would take them a lane at a time.
  ShellEntry* s_lut = reinterpret_cast<ShellEntry*>(smem);
  float* s_r2 = reinterpret_cast<float*>(smem + 2 * kLutSize);
  uint32_t* s_mark = kLds ? smem + 2 * kLutSize + kpad : smem;
  uint32_t* s_bins = s_mark + kCountWaves * 128;

  if (kLds) {
    for (int i = threadIdx.x; i < K; i += kCountThreads) s_r2[i] = sh.r2[i];
    if (kShell != kShellScan)
      for (int i = threadIdx.x; i < sh.nlut; i += kCountThreads) s_lut[i] = sh.lut[i];
  }
  __syncthreads();
  const float* r2 = kLds ? s_r2 : sh.r2;
  const ShellEntry* lut = kLds ? s_lut : sh.lut;

  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int grp = lane / GROUP;
  const int gl = lane % GROUP;
  uint32_t* mark = s_mark + wave * 128 + grp * 2 * GROUP;
  // one copy below 64 lanes per halo: constants, so those kernels keep
  // their plain histogram code
  if constexpr (GROUP != 64) {
    copies = 1u;
    ks = (uint32_t)K;
  }
  const uint32_t nbins = copies * ks;
  uint32_t* bins = s_bins + (wave * H + grp) * nbins;
  char* bins_bytes = reinterpret_cast<char*>(bins + (gl & (copies - 1)) * ks);

  const __amdgpu_buffer_rsrc_t rx = make_rsrc(ix.x, ix.n);
  const __amdgpu_buffer_rsrc_t ry = make_rsrc(ix.y, ix.n);
  const __amdgpu_buffer_rsrc_t rz = make_rsrc(ix.z, ix.n);

  // Workgroups go round-robin over the 8 XCDs; give each XCD one contiguous
  // range of the (Morton-sorted) halos so its L2 sees neighbouring queries.
  const uint32_t per = gridDim.x >> 3;
  const uint32_t blk = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint64_t h0 = (uint64_t)blk * halos_per_block;
  uint64_t h1 = h0 + halos_per_block;
  if (h1 > hv.n) h1 = hv.n;

  const float rbox = sh.rbox;
  const float inv = ix.grid.inv_cell;
  const float invx = ix.grid.inv_cell_x;
  const float cs = ix.grid.cell;
  const int G = ix.grid.g;
  const int gm1 = G - 1;
  const int GX = ix.grid.gx;
  const int gxm1 = GX - 1;
  uint64_t inbox = 0;

  // Particle staging (kStage, count pass with ascending shells): the
  // workgroup copies the particles of the union of its live halos' base
  // boxes into LDS -- the union's cell rows, each row's x run with
  // coalesced loads, one wavefront per row -- and its halos' base images
  // read their candidates from there.  s_delta[r] maps a particle index of
  // union row r to its LDS slot.  Images across the periodic boundary, and
  // workgroups whose union has more than kStageRows rows or stage_cap
  // particles, read from HBM/L2 as without staging.
  // LDS after the histograms: [s_delta kStageRows][row start, length,
  // slot: 3 x kStageRows][union bounds and scan: 32][x, y, z: 3 x
  // (stage_cap + 8)]
  uint32_t* s_delta = s_bins + kCountWaves * H * nbins;
  float* s_px = reinterpret_cast<float*>(s_delta + 4 * kStageRows + 32);
  float* s_py = s_px + stage_cap + 8;
  float* s_pz = s_py + stage_cap + 8;
  bool staged = false;
  int ux0 = 0, uy0 = 0, uz0 = 0, unyu = 0;
  if constexpr (kStage) {
    uint32_t* s_rgs = s_delta + kStageRows;
    uint32_t* s_rlen = s_rgs + kStageRows;
    uint32_t* s_roff = s_rlen + kStageRows;
    float* s_uni = reinterpret_cast<float*>(s_roff + kStageRows);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t i = h0 + threadIdx.x; i < h1; i += kCountThreads) {
      const float c[3] = {hv.x[i], hv.y[i], hv.z[i]};
      if (c[0] == c[0] && c[1] == c[1] && c[2] == c[2]) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          lo[d] = fminf(lo[d], c[d] - rbox);
          hi[d] = fmaxf(hi[d], c[d] + rbox);
        }
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        lo[d] = fminf(lo[d], __shfl_xor(lo[d], off, 64));
        hi[d] = fmaxf(hi[d], __shfl_xor(hi[d], off, 64));
      }
    // the waves' bounds, then every thread combines them (s_uni: 4 x 6)
    if (lane == 0)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        s_uni[wave * 6 + d] = lo[d];
        s_uni[wave * 6 + 3 + d] = hi[d];
      }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kCountWaves; ++w)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        lo[d] = fminf(lo[d], s_uni[w * 6 + d]);
        hi[d] = fmaxf(hi[d], s_uni[w * 6 + 3 + d]);
      }
    // union cell ranges; a member's cells lie inside them (cell_of is
    // monotone); no live halo: nothing to stage
    const bool any = lo[0] <= hi[0];
    ux0 = cell_of(lo[0], invx, gxm1);
    const int ux1 = cell_of(hi[0], invx, gxm1);
    uy0 = cell_of(lo[1], inv, gm1);
    const int uy1 = cell_of(hi[1], inv, gm1);
    uz0 = cell_of(lo[2], inv, gm1);
    const int uz1 = cell_of(hi[2], inv, gm1);
    unyu = uy1 - uy0 + 1;
    const int urows = unyu * (uz1 - uz0 + 1);
    if (any && urows <= kStageRows) {
      // each union row's x run: its first particle and length, then the
      // runs' slots in LDS by a prefix sum over the rows
      const int r = threadIdx.x;
      uint32_t gs = 0, len = 0;
      if (r < urows) {
        const int zq = r / unyu;
        const uint32_t kb = ((uint32_t)(uz0 + zq) * (uint32_t)G + (uint32_t)(uy0 + r - zq * unyu)) *
                            (uint32_t)GX;
        const uint32_t k0 = kb + (uint32_t)ux0, k1 = kb + (uint32_t)ux1;
        gs = k0 ? ix.cell_end[k0 - 1] : 0u;
        len = ix.cell_end[k1] - gs;
      }
      uint32_t total;
      const uint32_t incl = block_incl_sum(len, reinterpret_cast<uint32_t*>(s_uni + 24), &total);
      staged = total <= stage_cap;  // the same on every thread
      if (staged && r < urows) {
        s_rgs[r] = gs;
        s_rlen[r] = len;
        s_roff[r] = incl - len;
        s_delta[r] = incl - len - gs;  // LDS slot of particle i of row r: i + s_delta[r]
      }
      __syncthreads();
      if (staged) {
        for (int rr = wave; rr < urows; rr += kCountWaves) {
          const uint32_t g0 = s_rgs[rr], n0 = s_rlen[rr], o0 = s_roff[rr];
          for (uint32_t i = lane; i < n0; i += 64) {
            s_px[o0 + i] = ix.x[g0 + i];
            s_py[o0 + i] = ix.y[g0 + i];
            s_pz[o0 + i] = ix.z[g0 + i];
          }
        }
        // a slot past the end of the last row reads up to SEG - 1 beyond it
        if (threadIdx.x < 8) {
          s_px[total + threadIdx.x] = 0.0f;
          s_py[total + threadIdx.x] = 0.0f;
          s_pz[total + threadIdx.x] = 0.0f;
        }
      }
      __syncthreads();
    }
  }

  // the group's histogram starts at zero; each flush leaves it zero again
  if (kLds && !kInbox) {
    for (int i = gl; i < (int)nbins; i += GROUP) bins[i] = 0u;
  }
  for (uint64_t hb = h0 + (uint64_t)wave * H; hb < h1; hb += (uint64_t)kCountWaves * H) {
    const uint64_t hs = hb + grp;
    const bool active = hs < h1;
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    // counts are kept in the sorted halo order (rows of neighbouring halos
    // are neighbours in memory); they are put back in input order once, when
    // they are read out
    uint64_t row = hs;
    if (active) {
      cx = hv.x[hs];
      cy = hv.y[hs];
      cz = hv.z[hs];
      if (hv.slot) row = hv.slot[hs];
    }
    // A NaN centre makes every distance NaN: nothing is ever counted.  Its
    // box still passes every coordinate on the NaN axes, and the in-box
    // measurement counts those pair tests as the reference makes them.
    const bool live = active && (kInbox || ((cx == cx) && (cy == cy) && (cz == cz)));
    // Periodic images (src/objects/ArraySTR/MultiRangeCount.c:31-71,
    // src/objects/RTree/Mbr.c:284-313): an axis whose base box leaves
    // [0, 1000] is shifted by +1000 (low side wins) or -1000.
    float shift[3] = {0.0f, 0.0f, 0.0f};
    int oob = 0;
    {
      const float c[3] = {cx, cy, cz};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const bool below = c[d] - rbox < kBoxLo;
        const bool above = c[d] + rbox > kBoxHi;
        shift[d] = below ? kBoxHi : (above ? -kBoxHi : 0.0f);
        oob |= (below || above) ? 1 << d : 0;
      }
    }
    // images any halo of this wavefront has (scalar: skipping costs nothing)
    const uint32_t any_oob = (__ballot(live && (oob & 1)) ? 1u : 0u) |
                             (__ballot(live && (oob & 2)) ? 2u : 0u) |
                             (__ballot(live && (oob & 4)) ? 4u : 0u);
    for (uint32_t img = 0; img < 8; ++img) {
      if (img & ~any_oob) continue;
      const bool on = live && !(img & ~(uint32_t)oob);
      if (!__any(on)) continue;
      const float sx = (img & 1) ? cx + shift[0] : cx;
      const float sy = (img & 2) ? cy + shift[1] : cy;
      const float sz = (img & 4) ? cz + shift[2] : cz;
      const float lx = sx - rbox, hx = sx + rbox;
      const float ly = sy - rbox, hy = sy + rbox;
      const float lz = sz - rbox, hz = sz + rbox;
      // below `safe` a distance proves the box test
      float safe = 0.0f;
      if constexpr (kSphere) safe = safe_dist2(sx, sy, sz, rbox);
      int x0, x1, y0, y1, z0, z1;
      if constexpr (kSphere) {
        x0 = cell_of(lx, invx, gxm1), x1 = cell_of(hx, invx, gxm1);
        y0 = cell_of(ly, inv, gm1), y1 = cell_of(hy, inv, gm1);
        z0 = cell_of(lz, inv, gm1), z1 = cell_of(hz, inv, gm1);
      } else {
        x0 = cell_lo(lx, invx, gxm1), x1 = cell_hi(hx, invx, gxm1);
        y0 = cell_lo(ly, inv, gm1), y1 = cell_hi(hy, inv, gm1);
        z0 = cell_lo(lz, inv, gm1), z1 = cell_hi(hz, inv, gm1);
      }
      const int ny = y1 - y0 + 1;
      const int nrows = on ? ny * (z1 - z0 + 1) : 0;
      const int max_rows = (int)wave_max_of_groups<GROUP>((uint32_t)nrows);
      for (int rb = 0; rb < max_rows; rb += GROUP) {
        // this lane's row of the batch: start, length, and its segments
        const int r = rb + gl;
        uint32_t beg = 0, len = 0, sdelta = 0;
        if (r < nrows) {
          // r / ny: a reciprocal estimate is off by at most one here
          int zq = (int)(((float)r + 0.5f) * __builtin_amdgcn_rcpf((float)ny));
          zq -= zq * ny > r ? 1 : 0;
          zq += (zq + 1) * ny <= r ? 1 : 0;
          const int yy = y0 + (r - zq * ny), zz = z0 + zq;
          int xa = x0, xb = x1;
          bool keep = true;
          if constexpr (kSphere) {
            // the row's (y, z) cell rectangle, widened by 1e-3 for the
            // binning's rounding (edge cells reach to infinity); a particle
            // there with dist < r2max has |px - sx| < w
            const float ylo = yy > 0 ? (float)yy * cs - 1e-3f : -INFINITY;
            const float yhi = yy < gm1 ? (float)(yy + 1) * cs + 1e-3f : INFINITY;
            const float zlo = zz > 0 ? (float)zz * cs - 1e-3f : -INFINITY;
            const float zhi = zz < gm1 ? (float)(zz + 1) * cs + 1e-3f : INFINITY;
            const float dy = fmaxf(fmaxf(ylo - sy, sy - yhi), 0.0f);
            const float dz = fmaxf(fmaxf(zlo - sz, sz - zhi), 0.0f);
            const float rem = sh.r2max_hi - (dy * dy + dz * dz) * 0.99999f;
            keep = rem > 0.0f;
            if (keep) {
              const float w = __builtin_amdgcn_sqrtf(rem) * 1.00001f + 1e-3f;
              xa = max(xa, cell_of(sx - w, invx, gxm1));
              xb = min(xb, cell_of(sx + w, invx, gxm1));
            }
          }
          if (keep) {
            const uint32_t kb = ((uint32_t)zz * (uint32_t)G + (uint32_t)yy) * (uint32_t)GX;
            const uint32_t k0 = kb + (uint32_t)xa, k1 = kb + (uint32_t)xb;
            beg = k0 ? ix.cell_end[k0 - 1] : 0u;
            len = ix.cell_end[k1] - beg;
            // the base image's rows lie in the staged union
            if (kStage && staged && img == 0) sdelta = s_delta[(zz - uz0) * unyu + (yy - uy0)];
          }
        }
        const bool from_lds = kStage && staged && img == 0;
        const uint32_t nseg = (len + SEG - 1) / SEG;
        const uint32_t sincl = grp_scan_add<GROUP>(nseg, gl);
        const uint32_t sexcl = sincl - nseg;
        const uint32_t stotal = grp_last<GROUP>(sincl);
        // segment q of row j starts at particle abase_j + q * SEG and its row
        // ends at aend_j (abase wraps modulo 2^32; the sum does not)
        const uint32_t abase = beg - sexcl * (uint32_t)SEG;
        const uint32_t aend = beg + len;
        const uint32_t max_stotal = wave_max_of_groups<GROUP>(stotal);
        uint32_t carry = 0;
        // 64 and 16 lanes per halo (without LDS staging): the next round's
        // candidates are loaded before this round's are tested (its loads
        // wait out the tests; at 16 lanes the second set of registers leaves
        // 5 wavefronts per SIMD instead of 6, and still gains).  8 lanes keep
        // one round in flight.
        constexpr bool kAhead = GROUP == 64 || (GROUP == 16 && !kStage);
        // Round c0's candidates of this lane (never staged), as in
        // the round code below, for the round ahead.  Called by the whole
        // wave, rounds in order (carry).
        auto fetch_ahead = [&](uint32_t c0, float (&fx)[SEG], float (&fy)[SEG],
                               float (&fz)[SEG], uint32_t& slen) {
          mark[gl] = 0u;
          wave_sync();
          if (nseg > 0 && sexcl >= c0 && sexcl < c0 + GROUP) mark[sexcl - c0] = (uint32_t)gl + 1;
          wave_sync();
          const uint32_t m = grp_scan_max<GROUP>(mark[gl], gl);
          const uint32_t j = m ? m - 1 : carry;
          carry = grp_last<GROUP>(j);
          const uint32_t q = c0 + (uint32_t)gl;
          const uint32_t start = grp_lane<GROUP>(abase, j) + q * (uint32_t)SEG;
          const uint32_t rend = grp_lane<GROUP>(aend, j);
          slen = q < stotal ? min((uint32_t)SEG, rend - start) : 0u;
          wave_sync();  // the next call rewrites the marks
          const uint32_t boff = start * 4u;
#pragma unroll
          for (int u = 0; u < SEG; ++u) {
            fx[u] = buf_load(rx, boff + 4u * u);
            fy[u] = buf_load(ry, boff + 4u * u);
            fz[u] = buf_load(rz, boff + 4u * u);
          }
        };
        float nx[SEG], ny[SEG], nz[SEG];
        uint32_t nslen = 0;
        if constexpr (kAhead) {
          if (max_stotal > 0) fetch_ahead(0, nx, ny, nz, nslen);
        }
        for (uint32_t c0 = 0; c0 < max_stotal; c0 += GROUP) {
          float px[SEG], py[SEG], pz[SEG];
          uint32_t slen;
          if constexpr (kAhead) {
#pragma unroll
            for (int u = 0; u < SEG; ++u) {
              px[u] = nx[u];
              py[u] = ny[u];
              pz[u] = nz[u];
            }
            slen = nslen;
            if (c0 + GROUP < max_stotal) fetch_ahead(c0 + GROUP, nx, ny, nz, nslen);
          } else {
            // rows whose first segment falls in [c0, c0 + GROUP) mark its
            // slot; a max-scan over the marks gives every slot its row
            mark[gl] = 0u;
            wave_sync();
            if (nseg > 0 && sexcl >= c0 && sexcl < c0 + GROUP) mark[sexcl - c0] = (uint32_t)gl + 1;
            wave_sync();
            const uint32_t m = grp_scan_max<GROUP>(mark[gl], gl);
            const uint32_t j = m ? m - 1 : carry;
            // cross-lane reads run on every lane: one taken inside a branch
            // would read 0 from lanes the branch left inactive
            carry = grp_last<GROUP>(j);
            const uint32_t q = c0 + (uint32_t)gl;
            const uint32_t start = grp_lane<GROUP>(abase, j) + q * (uint32_t)SEG;
            const uint32_t rend = grp_lane<GROUP>(aend, j);
            slen = q < stotal ? min((uint32_t)SEG, rend - start) : 0u;
            if (from_lds) {
              // the staged row's slots; a slot past its group's segments
              // reads slot 0, and one past its row's end at most SEG - 1
              // beyond the last row (padded): values it never uses
              const uint32_t rsd = grp_lane<GROUP>(sdelta, j);
              const uint32_t li = q < stotal ? start + rsd : 0u;
#pragma unroll
              for (int u = 0; u < SEG; ++u) {
                px[u] = s_px[li + u];
                py[u] = s_py[li + u];
                pz[u] = s_pz[li + u];
              }
            } else {
              // a slot past its group's segments loads whatever its offset
              // names (bounds-checked by the buffer descriptor) and never
              // uses it
              const uint32_t boff = start * 4u;
#pragma unroll
              for (int u = 0; u < SEG; ++u) {
                px[u] = buf_load(rx, boff + 4u * u);
                py[u] = buf_load(ry, boff + 4u * u);
                pz[u] = buf_load(rz, boff + 4u * u);
              }
            }
          }
          if constexpr (kSphere) {
            // count pass, ascending shells: a counted pair has dist < r2max
            // (NaN fails it); the box test only decides for the few with
            // dist >= safe.  The shell-table reads of all SEG slots go out
            // together, then the bin increments.
            float dist[SEG];
            bool hit[SEG];
#pragma unroll
            for (int u = 0; u < SEG; ++u) {
              dist[u] = dpps_distance(px[u], py[u], pz[u], sx, sy, sz);
              // a pair outside the box gets distance +inf: the branch then
              // writes only a float, and the hit flags below stay lane masks
              if (dist[u] < sh.r2max && !(dist[u] < safe) &&
                  !(!(px[u] < lx) && !(hx < px[u]) && !(py[u] < ly) && !(hy < py[u]) &&
                    !(pz[u] < lz) && !(hz < pz[u])))
                dist[u] = INFINITY;
              hit[u] = ((uint32_t)u < slen) & (dist[u] < sh.r2max);
            }
            ShellEntry e[SEG];
#pragma unroll
            for (int u = 0; u < SEG; ++u) {
              // a slot without a hit reads entry 0 (its distance may be
              // anything); a hit's bucket is below nlut
              const uint32_t b =
                  __builtin_elementwise_sub_sat(__float_as_uint(dist[u]), sh.lut_base) >>
                  sh.lut_shift;
              e[u] = lut[hit[u] ? b : 0u];
            }
#pragma unroll
            for (int u = 0; u < SEG; ++u) {
              if (hit[u]) {
                int o = (int)e[u].off + (dist[u] < e[u].t ? 0 : 4);
                if constexpr (kShell == kShellTable)
                  while (!(dist[u] < r2[o >> 2])) o += 4;  // ends: dist < r2[nrads-1]
                if (kLds)
                  atomicAdd(reinterpret_cast<uint32_t*>(bins_bytes + o), 1u);
                else
                  atomicAdd(counts + (uint64_t)row * K + (o >> 2), 1u);
              }
            }
          } else {
#pragma unroll
            for (int u = 0; u < SEG; ++u) {
              // closed, NaN-permissive box (src/objects/RTree/Mbr.c:171-182)
              const bool in = (uint32_t)u < slen && !(px[u] < lx) && !(hx < px[u]) &&
                              !(py[u] < ly) && !(hy < py[u]) && !(pz[u] < lz) && !(hz < pz[u]);
              if constexpr (kInbox) {  // measurement pass: count the pair test only
                inbox += in ? 1 : 0;
              } else {  // any thresholds: first match by a plain scan
                if (in) {
                  const int o = shell_off<kShell>(dpps_distance(px[u], py[u], pz[u], sx, sy, sz),
                                                  r2, lut, sh);
                  if (o >= 0) {
                    if (kLds)
                      atomicAdd(reinterpret_cast<uint32_t*>(bins_bytes + o), 1u);
                    else
                      atomicAdd(counts + (uint64_t)row * K + (o >> 2), 1u);
                  }
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
      if (active) {
        uint32_t* dst = counts + (uint64_t)row * K;
        for (int k = gl; k < K; k += GROUP) {
          uint32_t v = 0;
          for (uint32_t c = 0; c < copies; ++c) {
            v += bins[c * ks + k];
            bins[c * ks + k] = 0u;
          }
          if (accumulate) dst[k] += v; else dst[k] = v;
        }
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

// Launch shape of the count kernel.
struct CountLaunch {
  uint32_t nblocks;
  size_t lds;
  int seg;
  uint32_t halos_per_block;
  uint32_t stage_cap;  // 0: no particle staging
};

// Copies of a halo's LDS histogram: 8 with 64 lanes per halo (as many as
// fit kMaxGroupBinBytes), else 1; and the words between copies.
int bin_copies(int nrads, int group) {
  if (nrads > kMaxLdsShells || group != 64) return 1;
  int c = 8;
  while (c > 1 && (size_t)kCountWaves * c * (size_t)(nrads | 1) * 4 > kMaxGroupBinBytes) c >>= 1;
  return c;
}

uint32_t bin_stride(int nrads, int group) {
  return bin_copies(nrads, group) > 1 ? (uint32_t)(nrads | 1) : (uint32_t)nrads;
}

template <int GROUP, int kShell, bool kLds, bool kInbox, bool kStage>
void launch_group(const IndexView& ix, const HaloView& hv, const Shells& sh, uint32_t* counts,
                  int accumulate, unsigned long long* inbox, const CountLaunch& L,
                  hipStream_t s) {
  const uint32_t copies = (uint32_t)bin_copies(sh.nrads, GROUP);
  const uint32_t ks = bin_stride(sh.nrads, GROUP);
  if (L.seg >= 8)
    count_group_kernel<GROUP, 8, kShell, kLds, kInbox, kStage>
        <<<L.nblocks, kCountThreads, L.lds, s>>>(ix, hv, sh, counts, L.halos_per_block,
                                                 accumulate, inbox, L.stage_cap, copies, ks);
  else
    count_group_kernel<GROUP, 4, kShell, kLds, kInbox, kStage>
        <<<L.nblocks, kCountThreads, L.lds, s>>>(ix, hv, sh, counts, L.halos_per_block,
                                                 accumulate, inbox, L.stage_cap, copies, ks);
}

template <int GROUP, int kShell>
void dispatch_flags(bool lds, bool in, const IndexView& ix, const HaloView& hv,
                    const Shells& sh, uint32_t* counts, int acc, unsigned long long* inbox,
                    const CountLaunch& L, hipStream_t s) {
  if (lds && !in) {
    if constexpr (kShell != kShellScan && GROUP != 64) {
      if (L.stage_cap) {
        launch_group<GROUP, kShell, true, false, true>(ix, hv, sh, counts, acc, inbox, L, s);
        return;
      }
    }
    launch_group<GROUP, kShell, true, false, false>(ix, hv, sh, counts, acc, inbox, L, s);
  } else if (lds && in) {
    launch_group<GROUP, kShell, true, true, false>(ix, hv, sh, counts, acc, inbox, L, s);
  } else if (!in) {
    launch_group<GROUP, kShell, false, false, false>(ix, hv, sh, counts, 1, inbox, L, s);
  } else {
    launch_group<GROUP, kShell, false, true, false>(ix, hv, sh, counts, 1, inbox, L, s);
  }
}

template <int GROUP>
void dispatch_group(int shell, bool lds, bool in, const IndexView& ix, const HaloView& hv,
                    const Shells& sh, uint32_t* counts, int acc, unsigned long long* inbox,
                    const CountLaunch& L, hipStream_t s) {
  if (shell == kShellOneStep)
    dispatch_flags<GROUP, kShellOneStep>(lds, in, ix, hv, sh, counts, acc, inbox, L, s);
  else if (shell == kShellTable)
    dispatch_flags<GROUP, kShellTable>(lds, in, ix, hv, sh, counts, acc, inbox, L, s);
  else
    dispatch_flags<GROUP, kShellScan>(lds, in, ix, hv, sh, counts, acc, inbox, L, s);
}

// c64[i] (= or +=) c32[i]: the chunk counts folded into the u64 totals.
__global__ void widen_counts_kernel(const uint32_t* __restrict__ c32, uint64_t* __restrict__ c64,
                                    uint64_t total, int store) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x)
    c64[i] = (store ? 0ull : c64[i]) + c32[i];
}

// inv[orig[i]] = i: for each input row, its position in sorted order.
__global__ void invert_rows_kernel(const uint32_t* __restrict__ orig, uint32_t* __restrict__ inv,
                                   uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    inv[orig[i]] = (uint32_t)i;
}

// dst[r * K + k] = c64[inv[r] * K + k] + c32[inv[r] * K + k] (either source
// may be absent; a u32 destination takes c32 alone): the rows back in input
// order.  A block writes kReadoutRows consecutive output rows, its lanes
// walking their kReadoutRows * K words in order, so the stores are
// contiguous and each source row is one contiguous read of K words.  Sources
// and output are touched once, so both go non-temporal (no L2 retention):
// 1.40 -> 1.33 ms for 1e7 x 40 u64 counts.
constexpr uint32_t kReadoutRows = 64;
constexpr int kReadoutUnroll = 4;

template <typename Out>
__global__ __launch_bounds__(256) void readout_counts_kernel(
    const uint32_t* __restrict__ c32, const uint64_t* __restrict__ c64,
    const uint32_t* __restrict__ inv, Out* __restrict__ dst, uint64_t n, uint32_t K) {
  // a thread's word o = row r * K + k advances by 256 words a step: r and k
  // are carried along instead of dividing every word's index by K
  const uint32_t step_r = 256u / K, step_k = 256u % K;
  const uint32_t r_first = threadIdx.x / K, k_first = threadIdx.x % K;
  for (uint64_t r0 = (uint64_t)blockIdx.x * kReadoutRows; r0 < n;
       r0 += (uint64_t)gridDim.x * kReadoutRows) {
    const uint32_t rows = (uint32_t)min((uint64_t)kReadoutRows, n - r0);
    const uint32_t m = rows * K;
    Out* d = dst + r0 * K;
    uint32_t r = r_first, k = k_first;
    for (uint32_t o0 = threadIdx.x; o0 < m; o0 += 256 * kReadoutUnroll) {
      Out v[kReadoutUnroll];
#pragma unroll
      for (int u = 0; u < kReadoutUnroll; ++u) {
        v[u] = 0;
        if (o0 + 256u * u < m) {
          const uint64_t src = (uint64_t)inv[r0 + r] * K + k;
          if constexpr (sizeof(Out) == 8)
            v[u] = (c64 ? __builtin_nontemporal_load(&c64[src]) : 0ull) +
                   (c32 ? __builtin_nontemporal_load(&c32[src]) : 0u);
          else
            v[u] = c32 ? __builtin_nontemporal_load(&c32[src]) : 0u;
        }
        r += step_r;
        k += step_k;
        if (k >= K) {
          k -= K;
          ++r;
        }
      }
#pragma unroll
      for (int u = 0; u < kReadoutUnroll; ++u)
        if (o0 + 256u * u < m) __builtin_nontemporal_store(v[u], &d[o0 + 256u * u]);
    }
  }
}

}  // namespace

hipError_t widen_counts(const uint32_t* c32, uint64_t* c64, uint64_t total, bool store,
                        hipStream_t s) {
  if (total == 0) return hipSuccess;
  uint64_t blocks = ceil_div(total, 256);
  if (blocks > 65536) blocks = 65536;
  widen_counts_kernel<<<(uint32_t)blocks, 256, 0, s>>>(c32, c64, total, store ? 1 : 0);
  return hipGetLastError();
}

hipError_t invert_rows(const uint32_t* orig, uint32_t* inv, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = ceil_div(n, 256);
  if (blocks > 65536) blocks = 65536;
  invert_rows_kernel<<<(uint32_t)blocks, 256, 0, s>>>(orig, inv, n);
  return hipGetLastError();
}

hipError_t readout_counts(const uint32_t* c32, const uint64_t* c64, const uint32_t* inv,
                          uint64_t* dst, uint64_t n, int K, hipStream_t s) {
  if (n == 0 || K <= 0) return hipSuccess;
  uint64_t blocks = ceil_div(n, kReadoutRows);
  if (blocks > 16384) blocks = 16384;
  readout_counts_kernel<uint64_t><<<(uint32_t)blocks, 256, 0, s>>>(c32, c64, inv, dst, n,
                                                                   (uint32_t)K);
  return hipGetLastError();
}

hipError_t readout_counts32(const uint32_t* c32, const uint32_t* inv, uint32_t* dst, uint64_t n,
                            int K, hipStream_t s) {
  if (n == 0 || K <= 0) return hipSuccess;
  uint64_t blocks = ceil_div(n, kReadoutRows);
  if (blocks > 16384) blocks = 16384;
  readout_counts_kernel<uint32_t><<<(uint32_t)blocks, 256, 0, s>>>(c32, nullptr, inv, dst, n,
                                                                   (uint32_t)K);
  return hipGetLastError();
}

constexpr size_t kNoLdsBytes = 4 * kCountWaves * 128;

int count_group_for(int nrads, int requested) {
  if (requested == 64 || requested == 16 || requested == 8) {
    if ((size_t)nrads * kCountWaves * (64 / requested) * 4 > kMaxGroupBinBytes &&
        nrads <= kMaxLdsShells)
      return 64;
    return requested;
  }
  return nrads <= 256 ? 16 : 64;
}

size_t count_stage_bytes(uint32_t stage_cap) {
  return stage_cap ? 4 * (4 * (size_t)kStageRows + 32 + 3 * ((size_t)stage_cap + 8)) : 0;
}

size_t count_lds_bytes(int nrads, int group) {
  if (nrads > kMaxLdsShells) return kNoLdsBytes;
  const size_t kpad = (size_t)((nrads + 3) & ~3);
  const size_t h = (size_t)(64 / group);
  const size_t bins = (size_t)bin_copies(nrads, group) * bin_stride(nrads, group);
  return 4 * (2 * (size_t)kLutSize + kpad + kCountWaves * 128 + kCountWaves * h * bins);
}

hipError_t count_halos(const IndexView& ix, const HaloView& hv, const Shells& sh,
                       uint32_t* counts, bool accumulate, unsigned long long* inbox,
                       int group, int seg, hipStream_t s, uint32_t stage_cap,
                       uint32_t stage_halos) {
  if (hv.n == 0 || sh.nrads <= 0) return hipSuccess;
  const bool lds = sh.nrads <= kMaxLdsShells;
  const bool measure = inbox != nullptr;
  if (!measure && !accumulate && (!lds || ix.n == 0)) {
    // the HBM-atomic path adds into the counts, and an empty chunk writes
    // nothing: either way a first chunk must start from zero
    hipError_t e = fill_u32(counts, 0u, hv.n * (uint64_t)sh.nrads, s);
    if (e != hipSuccess) return e;
  }
  if (ix.n == 0) return hipSuccess;
  if (ix.n > 0x3fffffffull) return hipErrorInvalidValue;  // 32-bit byte offsets
  const int g = lds ? count_group_for(sh.nrads, group) : 64;
  const int shell = !sh.sorted ? kShellScan : sh.one_step ? kShellOneStep : kShellTable;
  CountLaunch L;
  L.seg = seg;
  // staging: the count pass with ascending shells and LDS histograms, 8 or
  // 16 lanes per halo, within the LDS a workgroup can have
  L.stage_cap = (stage_cap && lds && !measure && shell != kShellScan && g != 64) ? stage_cap : 0u;
  L.halos_per_block = L.stage_cap && stage_halos ? stage_halos : kHalosPerBlock;
  L.lds = count_lds_bytes(sh.nrads, g) + count_stage_bytes(L.stage_cap);
  if (L.lds > kMaxCountLds) return hipErrorInvalidValue;
  uint64_t blocks = ceil_div(hv.n, L.halos_per_block);
  blocks = (blocks + 7) & ~7ull;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  L.nblocks = (uint32_t)blocks;
  const int acc = accumulate ? 1 : 0;
  if (g == 16)
    dispatch_group<16>(shell, lds, measure, ix, hv, sh, counts, acc, inbox, L, s);
  else if (g == 8)
    dispatch_group<8>(shell, lds, measure, ix, hv, sh, counts, acc, inbox, L, s);
  else
    dispatch_group<64>(shell, lds, measure, ix, hv, sh, counts, acc, inbox, L, s);
  return hipGetLastError();
}

}  // namespace sjcs
