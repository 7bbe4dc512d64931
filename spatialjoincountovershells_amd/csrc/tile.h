// tile.h -- the shared-candidate count of halo tiles.
//
// Halos whose Morton cells coincide (so their centres lie within about one
// unit of each other) form tiles.  One wavefront counts a whole tile: it walks
// the rows of the union of the tile's query boxes once, each lane holding one
// candidate particle, and tests every candidate against every halo of the
// tile.  The loads, cell-row walk and bookkeeping of a candidate are paid once
// per tile instead of once per halo, which is what dense, clustered halo sets
// need.  Pair by pair the test is the reference's, as in the per-halo kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace sjcs {

// Tile t is halos [start[t], start[t] + len[t]) of a HaloView.
struct TileView {
  const uint32_t* start;
  const uint32_t* len;
  uint32_t n;
  uint32_t max_len;  // longest tile: sizes the LDS histograms of a wavefront
};

// Halos per tile the count can take for nrads shells (its LDS histograms
// hold kTileBinWords counters per wavefront); 0 when tiles are not used.
constexpr int kTileMaxHalos = 64;
constexpr int kTileBinWords = 2048;
int tile_cap(int nrads);

// Workspace (u32 words) count_tiles needs for ntiles tiles.
uint64_t tile_work_words(uint32_t ntiles);

// counts[i*nrads + k] (= or +=) of every halo i of every tile, for ascending
// thresholds only (Shells::sorted).  With `inbox` not null, nothing is
// counted and inbox receives += the in-box pair tests.  `work` holds
// tile_work_words(tv.n) words.  No allocation.
hipError_t count_tiles(const IndexView& ix, const HaloView& hv, const Shells& sh,
                       const TileView& tv, uint32_t* counts, bool accumulate,
                       unsigned long long* inbox, uint32_t* work, hipStream_t s);

// dst[i] = src[perm[i]] for the halo arrays x, y, z and, when drow is not
// null, row.
hipError_t permute_halos(const float* sx, const float* sy, const float* sz, const uint32_t* srow,
                         float* dx, float* dy, float* dz, uint32_t* drow, const uint32_t* perm,
                         uint64_t n, hipStream_t s);

}  // namespace sjcs
