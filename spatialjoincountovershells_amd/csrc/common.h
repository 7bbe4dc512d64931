// common.h -- shared definitions of the SJCS HIP library (gfx950).
//
// Numerics contract (the reference's, restated):
//   * every distance is ((dx*dx + dy*dy) + dz*dz) in FP32, each op rounded,
//     no FMA (include/objects/RTree/Distance.h:11-13, DPPS imm 0x77).  The
//     library is compiled with -ffp-contract=off so the compiler cannot
//     fuse them, and FP32 denormals are kept (no flush-to-zero).
//   * the box test is !(p < lo) && !(hi < p) per axis, so a NaN passes
//     (src/objects/RTree/Mbr.c:171-182).
//   * a pair counts in the first shell k with dist < r2[k]
//     (src/objects/ArraySTR/MultiRangeCount.c:107-113).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sjcs {

constexpr float kBoxLo = 0.0f;
constexpr float kBoxHi = 1000.0f;

// Cell grid over the periodic box: G cells along y and z, GX along x (a
// multiple of G: finer x cells cut a query's rows closer to its sphere), key
// = (cz*G + cy)*GX + cx, so the particles of one (cy, cz) row are contiguous
// in key order and any run of x cells of a row is one contiguous index range.
struct Grid {
  float inv_cell;    // G / 1000
  float cell;        // 1000 / G
  float inv_cell_x;  // GX / 1000
  int g;             // cells along y and z
  int gx;            // cells along x
  uint32_t ncells;   // g * g * gx
};

// Monotone (non-decreasing) map from a raw FP32 coordinate to a cell, clamped
// -- never wrapped -- to [0, g-1].  Because it is monotone, lo <= p <= hi
// implies cell(lo) <= cell(p) <= cell(hi): scanning the cells of a query
// box's corners, then re-testing the box per particle, finds exactly the
// particles the reference's tree walk finds.  NaN maps to cell 0 (a NaN
// particle's distance is NaN and is never counted).
__device__ __forceinline__ int cell_of(float v, float inv_cell, int gm1) {
  float t = v * inv_cell;
  t = fmaxf(t, 0.0f);
  t = fminf(t, (float)gm1);
  return (int)t;
}

// Cell range of a query box edge.  A NaN edge (NaN centre or NaN radius)
// passes every box test on that axis, so it opens the range fully.
__device__ __forceinline__ int cell_lo(float lo, float inv_cell, int gm1) {
  return (lo != lo) ? 0 : cell_of(lo, inv_cell, gm1);
}
__device__ __forceinline__ int cell_hi(float hi, float inv_cell, int gm1) {
  return (hi != hi) ? gm1 : cell_of(hi, inv_cell, gm1);
}

__device__ __forceinline__ uint32_t cell_key(float x, float y, float z,
                                             const Grid& gr) {
  int gm1 = gr.g - 1;
  uint32_t cx = (uint32_t)cell_of(x, gr.inv_cell_x, gr.gx - 1);
  uint32_t cy = (uint32_t)cell_of(y, gr.inv_cell, gm1);
  uint32_t cz = (uint32_t)cell_of(z, gr.inv_cell, gm1);
  return (cz * (uint32_t)gr.g + cy) * (uint32_t)gr.gx + cx;
}

// Shell lookup tables for the first-match rule.  When r2 is non-decreasing,
// NaN-free and sign-bit-free, the FP32 bit pattern of a non-negative distance
// is monotone in its value, so a table indexed by (bits(d) - bits(r2[0])) >>
// shift gives a lower bound k on the first matching shell.  Each entry also
// carries r2[k]; when no bucket holds two thresholds (`one_step`), the shell
// is k + (d >= r2[k]), else a short upward scan finishes it.  The base is 0
// (bucket = bits >> shift) whenever that table is fine enough.
struct ShellEntry {
  uint32_t off;  // 4 * k, k = first shell that can match a distance of this bucket
  float t;       // r2[k]
};

struct Shells {
  const float* r2;          // device, nrads
  const ShellEntry* lut;    // device, nlut (sorted case only)
  int nrads;
  int sorted;               // 1: lut path valid; 0: plain first-match scan
  int one_step;             // 1: every bucket holds at most one threshold
  uint32_t lut_base;        // bits of r2[0]
  int lut_shift;
  int nlut;
  float rbox;               // (float)sqrt((double)r2[nrads-1]), from the host
  float r2max;              // r2[nrads-1]
  float r2max_hi;           // r2max * (1 + 1e-5): margin of the sphere cut
};

constexpr int kLutSize = 1024;
constexpr int kMaxLdsShells = 1024;  // larger K counts straight into HBM

// Device buffers of one indexed particle chunk.
struct IndexView {
  const float* x;
  const float* y;
  const float* z;
  const uint32_t* cell_end;  // ncells: number of particles with key <= c
  Grid grid;
  uint64_t n;
};

// Halos in a spatially coherent order; halo i's counts are row slot[i] of
// the counts (row i when slot is null).
struct HaloView {
  const float* x;
  const float* y;
  const float* z;
  const uint32_t* slot;
  uint64_t n;
};

}  // namespace sjcs
