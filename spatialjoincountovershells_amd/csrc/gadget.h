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

// Fills d for the n particles of a file decoded with p.  Fails (false, err
// set) unless the reference's integer arithmetic stays exact for every id of
// the file: 1 <= Ng1D, Nt1D <= 46340 (Ng2D and Nt2D are int products there),
// |id - 1| < 2^50, and every grid coordinate within +-(2^31 - 1).
bool gadget_prepare(const sjcs_gadget_decode& p, uint64_t n, GadgetDecode* d, std::string* err);

// xyz[3i + k] = coordinate k of particle i, for i < n.  No allocation.
hipError_t decode_gadget(const uint16_t* uv, uint64_t n, const GadgetDecode& d, float* xyz,
                         hipStream_t s);

}  // namespace sjcs
