// gadget.h -- decode of compressed (Gadget) particle files on the device.
//
// Particle n of a file has id = id_start + n.  The id names a grid point
// through GadId's index arithmetic (src/Gadget/GadId.c:70-86, 112-137):
//   m = id - 1,  IDt = m / Ng3D,  IDg = m - IDt*Ng3D     (C division: toward 0)
//   IDg -> (a, b, c) by Ng2D, Ng1D;  IDt -> (at, bt, ct) by Nt2D, Nt1D
//   ip = (at*Ng1D + a, bt*Ng1D + b, ct*Ng1D + c)
// and the uint16 triplet displaces it (BufVector3D.c:168-201, the branch the
// reference builds):
//   pos[d] = (float)((double)(cell_interval*ip[d] + disp_min) + unitsep*uv[d])
// with cell_interval*ip[d] + disp_min one FP32 FMA (gcc contracts it under
// -std=gnu99 for an FMA-capable -march) or two FP32 roundings, and the
// double product exact (a 24-bit significand times a 16-bit integer).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/sjcs.h"

namespace sjcs {

// Launch constants of one file's decode.
struct GadgetDecode {
  int64_t id_start;
  int64_t ng1d, ng2d, ng3d;  // grid points per tile: side, plane, volume
  int64_t nt1d, nt2d;        // tiles: per side, per plane
  double inv_ng1d, inv_ng2d, inv_ng3d, inv_nt1d, inv_nt2d;  // 1/x rounded
  float cell_interval;
  float disp_min;
  double unitsep;
  int fused;
  int one_tile;  // one tile per side and every id - 1 in [0, Ng3D)
};

// Integer division in double arithmetic.  Every id, quotient and remainder
// here is an integer of magnitude below 2^50, so each is exact in a double:
// q0 = trunc(a * inv), inv = 1/d rounded, is within 1/4 of a / d (two
// roundings of 2^-53 each on a quotient below 2^50) and so off by at most
// one; r = fma(-q0, d, a) is exact (q0*d < 2^51), and its sign fixes q0.
// Cheaper on the vector ALU than 64-bit integer division or conversions.
__device__ __forceinline__ void divmod_nonneg(double a, double d, double inv, double& q,
                                              double& r) {
  q = trunc(a * inv);
  r = fma(-q, d, a);
  // selects, not branches: a branch per particle makes the compiler wait
  // for each particle's loads before it issues the next particle's
  const double s = r < 0.0 ? -1.0 : (r >= d ? 1.0 : 0.0);
  q += s;
  r = fma(-s, d, r);
}

// C's a / d and a % d: quotient truncated toward zero, remainder with the
// sign of a.
__device__ __forceinline__ void divmod_c(double a, double d, double inv, double& q, double& r) {
  const bool neg = a < 0.0;
  divmod_nonneg(fabs(a), d, inv, q, r);
  q = neg ? -q : q;
  r = neg ? -r : r;
}

// Grid point of particle i of the file (its id is id_start + i), per axis
// (GadId.c:70-86, 112-137).  kOneTile: one tile per side and
// 0 <= id - 1 < Ng3D for the whole file, so the tile number is 0 and id - 1
// is the number within the tile.
template <bool kOneTile>
__device__ __forceinline__ void gadget_grid(const GadgetDecode& g, uint64_t i, int ip[3]) {
  const double ng1 = (double)g.ng1d, ng2 = (double)g.ng2d, ng3 = (double)g.ng3d;
  const double m = (double)(g.id_start - 1) + (double)i;
  double ga, gb, gc, ta = 0.0, tb = 0.0, tc = 0.0, r;
  if (kOneTile) {
    divmod_nonneg(m, ng2, g.inv_ng2d, ga, r);
    divmod_nonneg(r, ng1, g.inv_ng1d, gb, gc);
  } else {
    double idt, idg;
    divmod_c(m, ng3, g.inv_ng3d, idt, idg);
    divmod_c(idg, ng2, g.inv_ng2d, ga, r);
    divmod_c(r, ng1, g.inv_ng1d, gb, gc);
    divmod_c(idt, (double)g.nt2d, g.inv_nt2d, ta, r);
    divmod_c(r, (double)g.nt1d, g.inv_nt1d, tb, tc);
  }
  ip[0] = (int)fma(ta, ng1, ga);
  ip[1] = (int)fma(tb, ng1, gb);
  ip[2] = (int)fma(tc, ng1, gc);
}

// One coordinate: grid point ip of the axis displaced by the uint16 d.
__device__ __forceinline__ float gadget_coord(const GadgetDecode& g, int ip, uint32_t d) {
  const float f = (float)ip;
  // the library is built with -ffp-contract=off: the unfused form is two
  // roundings, the fused one a single FMA
  const float s = g.fused ? __builtin_fmaf(g.cell_interval, f, g.disp_min)
                          : g.cell_interval * f + g.disp_min;
  return (float)((double)s + g.unitsep * (double)d);
}

// Position of particle i of the file from its uint16 triplet d.  Callers
// load the triplets of all their particles before decoding any, so that the
// loads are in flight together.
template <bool kOneTile>
__device__ __forceinline__ void gadget_point(const GadgetDecode& g, uint64_t i,
                                             const uint32_t d[3], float p[3]) {
  int ip[3];
  gadget_grid<kOneTile>(g, i, ip);
#pragma unroll
  for (int k = 0; k < 3; ++k) p[k] = gadget_coord(g, ip[k], d[k]);
}

// Fills d for the n particles of a file decoded with p.  Fails (false, err
// set) unless the reference's integer arithmetic stays exact for every id of
// the file: 1 <= Ng1D, Nt1D <= 46340 (Ng2D and Nt2D are int products there),
// |id - 1| < 2^50, and every grid coordinate within +-(2^31 - 1).
bool gadget_prepare(const sjcs_gadget_decode& p, uint64_t n, GadgetDecode* d, std::string* err);

// xyz[3i + k] = coordinate k of particle i, for i < n.  No allocation.
hipError_t decode_gadget(const uint16_t* uv, uint64_t n, const GadgetDecode& d, float* xyz,
                         hipStream_t s);

}  // namespace sjcs
