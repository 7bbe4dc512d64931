// gadget.hip -- compressed (Gadget) particle decode: the header-derived
// parameters (host) and the per-particle decode (device).
#include <limits.h>
#include <math.h>

#include <algorithm>

#include "gadget.h"

namespace sjcs {

static_assert(sizeof(sjcs_gadget_header) == 256, "the Gadget header is 256 bytes");

namespace {

constexpr int kDecodeThreads = 256;
constexpr uint64_t kDecodeMaxBlocks = 1u << 16;

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
  if (r < 0.0) {
    q -= 1.0;
    r += d;
  } else if (r >= d) {
    q += 1.0;
    r -= d;
  }
}

// C's a / d and a % d: quotient truncated toward zero, remainder with the
// sign of a.
__device__ __forceinline__ void divmod_c(double a, double d, double inv, double& q, double& r) {
  const bool neg = a < 0.0;
  divmod_nonneg(fabs(a), d, inv, q, r);
  if (neg) {
    q = -q;
    r = -r;
  }
}

// kOneTile: one tile per side and 0 <= id - 1 < Ng3D for the whole file, so
// the tile number is 0 and id - 1 is the number within the tile.
template <bool kOneTile>
__global__ __launch_bounds__(kDecodeThreads) void gadget_decode_kernel(
    const uint16_t* __restrict__ uv, uint64_t n, GadgetDecode g, float* __restrict__ xyz) {
  const uint64_t stride = (uint64_t)gridDim.x * kDecodeThreads;
  const double ng1 = (double)g.ng1d, ng2 = (double)g.ng2d, ng3 = (double)g.ng3d;
  const double base = (double)(g.id_start - 1);
  for (uint64_t i = (uint64_t)blockIdx.x * kDecodeThreads + threadIdx.x; i < n; i += stride) {
    // grid point of id = id_start + i (GadId.c:70-86, 112-137)
    const double m = base + (double)i;
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
    const int ip[3] = {(int)fma(ta, ng1, ga), (int)fma(tb, ng1, gb), (int)fma(tc, ng1, gc)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float f = (float)ip[k];
      // the library is built with -ffp-contract=off: the unfused form is
      // two roundings, the fused one a single FMA
      const float s = g.fused ? __builtin_fmaf(g.cell_interval, f, g.disp_min)
                              : g.cell_interval * f + g.disp_min;
      const double v = (double)s + g.unitsep * (double)uv[3 * i + k];
      xyz[3 * i + k] = (float)v;
    }
  }
}

}  // namespace

bool gadget_prepare(const sjcs_gadget_decode& p, uint64_t n, GadgetDecode* d, std::string* err) {
  constexpr int64_t kSideMax = 46340;  // largest side whose square fits an int
  constexpr int64_t kIdMax = int64_t(1) << 50;
  const int64_t ng1d = p.ng1d, nt1d = p.nt1d;
  if (n && (ng1d < 1 || ng1d > kSideMax)) {
    *err = "compressed particles: grid points per side " + std::to_string(ng1d) +
           " outside [1, 46340]";
    return false;
  }
  if (n && (nt1d < 1 || nt1d > kSideMax)) {
    *err = "compressed particles: tiles per side " + std::to_string(nt1d) + " outside [1, 46340]";
    return false;
  }
  const int64_t g1 = std::max<int64_t>(ng1d, 1), t1 = std::max<int64_t>(nt1d, 1);
  d->id_start = p.id_start;
  d->ng1d = g1;
  d->ng2d = g1 * g1;
  d->ng3d = g1 * g1 * g1;
  d->nt1d = t1;
  d->nt2d = t1 * t1;
  d->inv_ng1d = 1.0 / (double)d->ng1d;
  d->inv_ng2d = 1.0 / (double)d->ng2d;
  d->inv_ng3d = 1.0 / (double)d->ng3d;
  d->inv_nt1d = 1.0 / (double)d->nt1d;
  d->inv_nt2d = 1.0 / (double)d->nt2d;
  d->cell_interval = p.cell_interval;
  d->disp_min = p.disp_min;
  d->unitsep = p.unitsep;
  d->fused = p.fused ? 1 : 0;
  d->one_tile = 0;
  if (n == 0) return true;
  // ids id_start .. id_start + n - 1, as m = id - 1
  if (n > (uint64_t)kIdMax || p.id_start < 2 - kIdMax || p.id_start > kIdMax) {
    *err = "compressed particles: ids beyond 2^50";
    return false;
  }
  const int64_t m_first = p.id_start - 1, m_last = p.id_start - 1 + (int64_t)(n - 1);
  if (m_last >= kIdMax) {
    *err = "compressed particles: ids beyond 2^50";
    return false;
  }
  // every tile index is at most |IDt| in magnitude, every grid coordinate at
  // most |IDt| * Ng1D + Ng1D - 1: keep them inside an int, as GadId holds them
  const int64_t idt_max = std::max(std::llabs(m_first / d->ng3d), std::llabs(m_last / d->ng3d));
  if (idt_max > ((int64_t)INT_MAX - (g1 - 1)) / g1) {
    *err = "compressed particles: ids give grid points beyond 2^31";
    return false;
  }
  d->one_tile = (t1 == 1 && m_first >= 0 && m_last < d->ng3d) ? 1 : 0;
  return true;
}

hipError_t decode_gadget(const uint16_t* uv, uint64_t n, const GadgetDecode& d, float* xyz,
                         hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint64_t blocks = std::min<uint64_t>((n + kDecodeThreads - 1) / kDecodeThreads,
                                             kDecodeMaxBlocks);
  if (d.one_tile)
    hipLaunchKernelGGL(gadget_decode_kernel<true>, dim3((unsigned)blocks), dim3(kDecodeThreads), 0,
                       s, uv, n, d, xyz);
  else
    hipLaunchKernelGGL(gadget_decode_kernel<false>, dim3((unsigned)blocks), dim3(kDecodeThreads),
                       0, s, uv, n, d, xyz);
  return hipGetLastError();
}

}  // namespace sjcs

extern "C" int sjcs_gadget_decode_params(const sjcs_gadget_header* h, int ntile,
                                         sjcs_gadget_decode* out) {
  if (!h || !out || ntile < 1) return SJCS_ERR_INVALID;
  // GadgetHeader.c:7-24
  const long long total =
      (long long)h->npartTotal[1] +
      (long long)((unsigned long long)(long long)h->npartTotal[2] << 32);
  if (total < 0) return SJCS_ERR_INVALID;
  const int at1d = (int)pow((double)total + 0.5, 1 / 3.);
  out->id_start = h->id_start;
  out->ng1d = at1d / ntile;
  out->nt1d = ntile;
  out->fused = 1;
  out->cell_interval = (float)(h->BoxSize / (double)at1d);
  out->disp_min = h->disp_min;
  const float span = h->disp_max - h->disp_min;  // an FP32 difference
  out->unitsep = (double)span / pow(2, h->pos_bits);
  return SJCS_OK;
}
