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

template <bool kOneTile>
__global__ __launch_bounds__(kDecodeThreads) void gadget_decode_kernel(
    const uint16_t* __restrict__ uv, uint64_t n, GadgetDecode g, float* __restrict__ xyz) {
  const uint64_t stride = (uint64_t)gridDim.x * kDecodeThreads;
  for (uint64_t i = (uint64_t)blockIdx.x * kDecodeThreads + threadIdx.x; i < n; i += stride) {
    const uint32_t d[3] = {uv[3 * i], uv[3 * i + 1], uv[3 * i + 2]};
    float p[3];
    gadget_point<kOneTile>(g, i, d, p);
    xyz[3 * i] = p[0];
    xyz[3 * i + 1] = p[1];
    xyz[3 * i + 2] = p[2];
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
