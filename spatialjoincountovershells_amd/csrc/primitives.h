// primitives.h -- host launchers for the device-wide building blocks:
// scans, the LSD radix sort and the cell table.  All launches go to the
// given stream and allocate nothing; callers pass the workspace.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "gadget.h"

namespace sjcs {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096 keys per block
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

// Cell sort (grids of at most kCsMaxCells cells along y and z,
// kCsMaxCellsX along x): tiles of the scatter passes, and rows of up to
// kCsRowCap particles of the last pass sorted in LDS.
constexpr int kCsThreads = 512;
// 8192 particles per block: each digit's run per tile is ~37 particles, so
// few cache lines are written in part (the scatter passes write 1.02x the
// particles' bytes to HBM; 4096-particle tiles wrote 1.24x and 1.34x); the
// staging takes 96 KB of LDS
constexpr int kCsItems = 16;
constexpr int kCsTile = kCsThreads * kCsItems;
constexpr int kCsRowThreads = 512;
constexpr int kCsRowItems = 8;
constexpr int kCsRowCap = kCsRowThreads * kCsRowItems;  // 4096 particles
constexpr int kCsMaxCells = 256;    // along y and z
constexpr int kCsMaxCellsX = 1024;  // along x

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Workspace (u32 words) a radix sort of n keys needs: per-tile digit
// histograms plus the scan partials over them.
inline uint64_t radix_workspace_words(uint64_t n) {
  uint64_t tiles = ceil_div(n ? n : 1, kSortTile);
  uint64_t hist = (uint64_t)kRadix * tiles;
  return hist + ceil_div(hist, kScanTile) + 64;
}

// Rows of the last cell-sort pass longer than kCsRowCap particles are cut
// into kCsTile tiles: at most n / (kCsRowCap + 1) such rows (and at most one
// per row of the grid), and at most one partial tile each.
inline uint64_t cs_long_rows_max(uint64_t n) {
  const uint64_t a = n / (kCsRowCap + 1), b = (uint64_t)kCsMaxCells * kCsMaxCells;
  return a < b ? a : b;
}
inline uint64_t cs_long_tiles_max(uint64_t n) { return ceil_div(n, kCsTile) + cs_long_rows_max(n); }

// Words of the cell sort's tile histograms: those of the z and y passes,
// later those of the long rows (kCsMaxCellsX digits per tile) in the same
// words.
inline uint64_t cs_hist_words(uint64_t n) {
  const uint64_t tiles1 = ceil_div(n ? n : 1, kCsTile);
  uint64_t hist = (uint64_t)kCsMaxCells * (2 * tiles1 + kCsMaxCells);
  const uint64_t hist3 = (uint64_t)kCsMaxCellsX * cs_long_tiles_max(n);
  if (hist3 > hist) hist = hist3;
  return (hist + 3) & ~3ull;
}
// Workspace (u32 words) of cell_sort_particles for n particles: the tile
// histograms, the long rows' records (4 words each) and their tiles' rows,
// slab starts and tiles, two counters and the scan partials.
inline uint64_t cell_sort_workspace_words(uint64_t n) {
  const uint64_t tiles2 = ceil_div(n ? n : 1, kCsTile) + kCsMaxCells;
  return cs_hist_words(n) + 4 * cs_long_rows_max(n) + cs_long_tiles_max(n) +
         2 * (kCsMaxCells + 1) + 4 + ceil_div((uint64_t)kCsMaxCells * tiles2, kScanTile) + 64;
}

// Workspace (u32 words) for a scan of n words.
inline uint64_t scan_workspace_words(uint64_t n) {
  return ceil_div(n ? n : 1, kScanTile) + 64;
}

// out[i] = in[0] + ... + in[i-1] (exclusive), in place allowed.
hipError_t exclusive_sum_u32(const uint32_t* in, uint32_t* out, uint64_t n,
                             uint32_t* ws, hipStream_t s);
// out[i] = max(in[0..i]) (inclusive), in place allowed.
hipError_t inclusive_max_u32(const uint32_t* in, uint32_t* out, uint64_t n,
                             uint32_t* ws, hipStream_t s);

hipError_t fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
hipError_t fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s);

// Stable LSD radix sort of (key, payload[npay]) by the low `key_bits` bits
// of key.  Buffers come in pairs [0]/[1]; data starts in [0]; returns in
// *result which pair holds the sorted data.  npay is 3 or 4.
hipError_t radix_sort_pairs(uint32_t* keys[2], uint32_t* pay[2][4], int npay,
                            uint64_t n, int key_bits, uint32_t* ws,
                            hipStream_t s, int* result);

// Particles of a chunk for the cell sort: packed FP32 xyz, or (uv set) a
// compressed file's uint16 triplets, decoded with gd where the first pass
// reads them; compressed files of one tile per side only (gd.one_tile).
struct ParticleSource {
  const float* xyz = nullptr;
  const uint16_t* uv = nullptr;
  GadgetDecode gd{};
};

// Particle chunk (n < 2^32) -> SoA coordinates in cell-key order in x[0],
// y[0], z[0] (x[1], y[1], z[1] are scratch), and the cell table:
// cell_end[c] = number of particles in cells <= c.  Needs g.g <=
// kCsMaxCells and g.gx <= kCsMaxCellsX.  The order of the particles inside
// one cell is unspecified.
hipError_t cell_sort_particles(const ParticleSource& src, uint64_t n, Grid g, float* x[2],
                               float* y[2], float* z[2], uint32_t* cell_end, uint32_t* ws,
                               hipStream_t s);

// Particle chunk (packed xyz) -> cell keys + SoA coordinates.
hipError_t particle_keys(const float* xyz, uint64_t n, Grid g, uint32_t* key,
                         float* x, float* y, float* z, hipStream_t s);

// Halo centres (3 FP32 every stride_bytes) -> Morton keys (10 bits per
// axis; the top key_bits of the 30, for a coarser order in fewer sort
// passes) + SoA + row.
hipError_t halo_keys(const float* pts, size_t stride_bytes, uint64_t n, uint32_t* key, float* x,
                     float* y, float* z, uint32_t* row, hipStream_t s, int key_bits = 30);

// *flag = 1 when the sorted keys hold a run of at least `run` equal keys,
// else 0.
hipError_t has_key_runs(const uint32_t* keys, uint64_t n, uint32_t run, uint32_t* flag,
                        hipStream_t s);

// A coarse description of how points fill the box: every step-th point
// (at most kOccSample of them) is counted in kOccCells^3 cells (coordinates
// px[i*stride], py[i*stride], pz[i*stride], clamped into the box), and
// hist[b] (kOccBins words) receives the number of cells holding between
// 2^(b-1) and 2^b - 1 of them (b = 0: empty cells).  ws holds
// occupancy_workspace_words() words.
constexpr int kOccCells = 64;
constexpr int kOccBins = 33;
constexpr uint64_t kOccSample = 1u << 22;
inline uint64_t occupancy_workspace_words() {
  return (uint64_t)kOccCells * kOccCells * kOccCells;
}
hipError_t occupancy_histogram(const float* px, const float* py, const float* pz, uint32_t stride,
                               uint64_t n, uint32_t* ws, uint32_t* hist, hipStream_t s);

// Cell table from keys sorted ascending: cell_end[c] = #keys <= c.
hipError_t build_cell_end(const uint32_t* sorted_keys, uint64_t n,
                          uint32_t* cell_end, uint32_t ncells, uint32_t* ws,
                          hipStream_t s);

}  // namespace sjcs
