// primitives.hip -- device-wide scans, the cell sort that replaces the
// reference's STR bulk load, and the cell table.
//
// The reference builds its index with a parallel LSD radix sort over FP32
// coordinate keys, one axis after another, then packs runs of 10 into tree
// nodes (src/objects/ArraySTR/Build.c:845-963).  Here the index is a cell
// grid: the particles in cell order plus a table of where every cell ends.
// Grids of up to 256 x 256 x 1024 cells use a three-pass MSD sort whose
// digits are single coordinates' cells (cz over the chunk, cy inside each
// cz slab, cx inside each (cz, cy) row); each pass counts per tile, scans,
// and scatters each tile through LDS so every digit's run is written out
// contiguously.  Finer grids use one stable LSD radix sort of (cell key, x,
// y, z) by 8-bit digits (wave ballots give the stable rank).

#include "primitives.h"

namespace sjcs {
namespace {

struct OpSum {
  __device__ static __forceinline__ uint32_t apply(uint32_t a, uint32_t b) { return a + b; }
};
struct OpMax {
  __device__ static __forceinline__ uint32_t apply(uint32_t a, uint32_t b) { return a > b ? a : b; }
};

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

template <class Op>
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v = Op::apply(v, t);
  }
  return v;
}

template <class Op>
__device__ __forceinline__ uint32_t wave_reduce(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = Op::apply(v, __shfl_xor(v, off, 64));
  return v;
}

// Exclusive block scan; both sum and max have identity 0 on u32.
template <class Op, int NT>
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  uint32_t inc = wave_incl_scan<Op>(v, lane);
  uint32_t exc = __shfl_up(inc, 1, 64);
  if (lane == 0) exc = 0;
  if (lane == 63) s_tmp[w] = inc;
  __syncthreads();
  uint32_t prefix = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) {
    uint32_t t = s_tmp[i];
    if (i < w) prefix = Op::apply(prefix, t);
    tot = Op::apply(tot, t);
  }
  __syncthreads();
  *total = tot;
  return Op::apply(prefix, exc);
}

// ---------------------------------------------------------------- scans ----

template <class Op>
__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(
    const uint32_t* __restrict__ in, uint64_t n, uint32_t* __restrict__ partials) {
  __shared__ uint32_t s_tmp[kScanThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint32_t acc = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    uint64_t i = base + (uint64_t)j * kScanThreads + threadIdx.x;
    if (i < n) acc = Op::apply(acc, in[i]);
  }
  acc = wave_reduce<Op>(acc);
  if (lane_id() == 0) s_tmp[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t r = 0;
    for (int i = 0; i < kScanThreads / 64; ++i) r = Op::apply(r, s_tmp[i]);
    partials[blockIdx.x] = r;
  }
}

// Exclusive scan of the per-tile partials, in place, by one block.
template <class Op>
__global__ __launch_bounds__(1024) void scan_partials_kernel(uint32_t* p, uint32_t np) {
  __shared__ uint32_t s_tmp[16];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < np; base += 1024) {
    uint32_t i = base + threadIdx.x;
    uint32_t v = i < np ? p[i] : 0;
    uint32_t tot;
    uint32_t exc = block_excl_scan<Op, 1024>(v, s_tmp, &tot);
    if (i < np) p[i] = Op::apply(carry, exc);
    carry = Op::apply(carry, tot);
  }
}

template <class Op, bool kExclusive>
__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(
    const uint32_t* in, uint32_t* out, uint64_t n, const uint32_t* __restrict__ partials) {
  __shared__ uint32_t s_tmp[kScanThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  if (base + kScanItems <= n) {
    const uint4* src = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
    for (int j = 0; j < kScanItems / 4; ++j) {
      uint4 q = src[j];
      v[4 * j] = q.x; v[4 * j + 1] = q.y; v[4 * j + 2] = q.z; v[4 * j + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) v[j] = (base + j < n) ? in[base + j] : 0u;
  }
  uint32_t acc = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) acc = Op::apply(acc, v[j]);
  uint32_t tot;
  uint32_t run = Op::apply(partials[blockIdx.x], block_excl_scan<Op, kScanThreads>(acc, s_tmp, &tot));
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    uint32_t x = v[j];
    if (kExclusive) {
      v[j] = run;
      run = Op::apply(run, x);
    } else {
      run = Op::apply(run, x);
      v[j] = run;
    }
  }
  if (base + kScanItems <= n) {
    uint4* dst = reinterpret_cast<uint4*>(out + base);
#pragma unroll
    for (int j = 0; j < kScanItems / 4; ++j)
      dst[j] = make_uint4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
  } else {
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
      if (base + j < n) out[base + j] = v[j];
  }
}

// A short array scanned by one block in chunks of 1024 (in place allowed:
// a chunk's reads precede its writes across the scan's barrier).
template <class Op, bool kExclusive>
__global__ __launch_bounds__(1024) void scan_single_kernel(const uint32_t* in, uint32_t* out,
                                                           uint32_t n) {
  __shared__ uint32_t s_tmp[16];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < n ? in[i] : 0u;
    uint32_t tot;
    const uint32_t exc = block_excl_scan<Op, 1024>(v, s_tmp, &tot);
    if (i < n) out[i] = Op::apply(carry, kExclusive ? exc : Op::apply(exc, v));
    carry = Op::apply(carry, tot);
  }
}

// Arrays up to this length take the one-launch scan (a serving request's
// radix histograms, small tile lists): one launch instead of three.
constexpr uint64_t kScanSingleMax = 8192;

template <class Op, bool kExclusive>
hipError_t scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* ws, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (n <= kScanSingleMax) {
    scan_single_kernel<Op, kExclusive><<<1, 1024, 0, s>>>(in, out, (uint32_t)n);
    return hipGetLastError();
  }
  uint64_t tiles = ceil_div(n, kScanTile);
  if (tiles > 0xffffffffull) return hipErrorInvalidValue;
  scan_reduce_kernel<Op><<<(uint32_t)tiles, kScanThreads, 0, s>>>(in, n, ws);
  scan_partials_kernel<Op><<<1, 1024, 0, s>>>(ws, (uint32_t)tiles);
  scan_apply_kernel<Op, kExclusive><<<(uint32_t)tiles, kScanThreads, 0, s>>>(in, out, n, ws);
  return hipGetLastError();
}

// ----------------------------------------------------------------- fills ----

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fill_u64_kernel(uint64_t* p, uint64_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

uint32_t fill_blocks(uint64_t n) {
  uint64_t b = ceil_div(n, 256);
  return (uint32_t)(b < 65536 ? (b ? b : 1) : 65536);
}

// ------------------------------------------------------------ radix sort ----

// Mask of the lanes of this wave whose `digit` equals this lane's (among
// lanes with `valid`); a lane that is not valid gets 0.
__device__ __forceinline__ uint64_t match_digit(uint32_t digit, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < kRadixBits; ++b) {
    bool bit = (digit >> b) & 1u;
    uint64_t bb = __ballot(bit);
    peers &= bit ? bb : ~bb;
  }
  return valid ? peers : 0ull;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << lane_id()) - 1ull;
}

// Per-tile digit histogram, digit-major: hist[d * ntiles + tile].
__global__ __launch_bounds__(kSortThreads) void radix_upsweep_kernel(
    const uint32_t* __restrict__ keys, uint64_t n, int shift,
    uint32_t* __restrict__ hist, uint32_t ntiles) {
  __shared__ uint32_t s_h[kRadix];
  const int tid = threadIdx.x;
  s_h[tid] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int j = 0; j < kSortItems; ++j) {
    uint64_t i = base + (uint64_t)j * kSortThreads + tid;
    bool valid = i < n;
    uint32_t d = valid ? (keys[i] >> shift) & (kRadix - 1) : 0u;
    uint64_t peers = match_digit(d, valid);
    // one LDS add per distinct digit of the wave: sorted or clumped input
    // would otherwise pile every lane onto one counter
    if (valid && (peers & lanemask_lt()) == 0)
      atomicAdd(&s_h[d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  hist[(uint64_t)tid * ntiles + blockIdx.x] = s_h[tid];
}

// One LSD pass over a tile of kSortTile keys (and NPAY payload words per
// key).  Each wavefront ranks its own kSortTile / 4 consecutive items in
// order (digit peers by ballots, a per-wave running count per digit in
// LDS), so the ranking needs no block barrier per item; one scan over the
// waves' digit counts then gives every item its stable position in the
// tile.  The tile goes out through LDS one array at a time (keys, then each
// payload), each staged slot carrying its output index, so the block needs
// 36 KB of LDS and several blocks share a CU.
template <int NPAY>
__global__ __launch_bounds__(kSortThreads) void radix_downsweep_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ p0i,
    const uint32_t* __restrict__ p1i, const uint32_t* __restrict__ p2i,
    const uint32_t* __restrict__ p3i, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ p0o, uint32_t* __restrict__ p1o,
    uint32_t* __restrict__ p2o, uint32_t* __restrict__ p3o, uint64_t n,
    int shift, const uint32_t* __restrict__ hist, uint32_t ntiles) {
  static_assert(kSortThreads == kRadix, "one thread per digit");
  constexpr int kWaves = kSortThreads / 64;
  constexpr int kWaveItems = kSortTile / kWaves;
  __shared__ uint32_t s_buf[kSortTile];   // staged keys, then each payload
  __shared__ uint32_t s_dst[kSortTile];   // output index of each staged slot
  __shared__ uint32_t s_cnt[kWaves][kRadix];
  __shared__ uint32_t s_goff[kRadix];
  __shared__ uint32_t s_tmp[kWaves];

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
  const uint32_t count = (uint32_t)((n - base) < (uint64_t)kSortTile ? (n - base) : kSortTile);
  const uint32_t* pin[4] = {p0i, p1i, p2i, p3i};
  uint32_t* pout[4] = {p0o, p1o, p2o, p3o};

#pragma unroll
  for (int w = 0; w < kWaves; ++w) s_cnt[w][tid] = 0;

  // item j of this lane: tile index wave * kWaveItems + j * 64 + lane
  uint32_t key[kSortItems], rank[kSortItems];
  uint32_t pay[NPAY][kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t t = (uint32_t)(wave * kWaveItems + j * 64 + lane);
    const bool valid = t < count;
    key[j] = valid ? kin[base + t] : 0u;
#pragma unroll
    for (int p = 0; p < NPAY; ++p) pay[p][j] = valid ? pin[p][base + t] : 0u;
  }
  __syncthreads();

  // rank among this wave's earlier items of the same digit
  uint32_t* wc = s_cnt[wave];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const bool valid = (uint32_t)(wave * kWaveItems + j * 64 + lane) < count;
    const uint32_t d = (key[j] >> shift) & (kRadix - 1);
    const uint64_t peers = match_digit(d, valid);
    const uint32_t before = (uint32_t)__popcll(peers & lanemask_lt());
    const uint32_t b = wc[d];
    rank[j] = b + before;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (valid && before == 0) wc[d] = b + (uint32_t)__popcll(peers);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();

  // digit tid: its items in earlier waves, its start in the tile, its
  // global start (scanned histogram) minus that
  uint32_t run = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const uint32_t c = s_cnt[w][tid];
    s_cnt[w][tid] = run;
    run += c;
  }
  uint32_t total;
  const uint32_t lstart = block_excl_scan<OpSum, kSortThreads>(run, s_tmp, &total);
#pragma unroll
  for (int w = 0; w < kWaves; ++w) s_cnt[w][tid] += lstart;
  s_goff[tid] = hist[(uint64_t)tid * ntiles + blockIdx.x] - lstart;
  __syncthreads();

  // tile slot of every item; stage the keys with their output indices
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    if ((uint32_t)(wave * kWaveItems + j * 64 + lane) < count) {
      const uint32_t d = (key[j] >> shift) & (kRadix - 1);
      const uint32_t lp = s_cnt[wave][d] + rank[j];
      rank[j] = lp;
      s_buf[lp] = key[j];
      s_dst[lp] = s_goff[d] + lp;
    }
  }
  __syncthreads();
  // consecutive threads write consecutive slots of a digit's run
  for (uint32_t l = tid; l < count; l += kSortThreads) kout[s_dst[l]] = s_buf[l];
#pragma unroll
  for (int p = 0; p < NPAY; ++p) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j)
      if ((uint32_t)(wave * kWaveItems + j * 64 + lane) < count) s_buf[rank[j]] = pay[p][j];
    __syncthreads();
    for (uint32_t l = tid; l < count; l += kSortThreads) pout[p][s_dst[l]] = s_buf[l];
  }
}

// ------------------------------------------------------------- cell sort ----
//
// For grids of at most kCsMaxCells cells per axis, particles are put in cell
// order (cz, cy, cx) by three MSD passes whose digits are single coordinates'
// cell indices, recomputed from the coordinate each time (no key array):
//   1. by cz over the whole chunk, packed xyz in, SoA out: per-tile digit
//      histograms, one exclusive scan over them (digit-major), a scatter;
//   2. by cy inside every cz slab: tiles never straddle a slab and the
//      histograms are laid out [slab][cy][tile of the slab], so again one
//      exclusive scan gives every (tile, digit) its output offset;
//   3. by cx inside every (cz, cy) row: one block per row counts, scans,
//      writes the row's entries of the cell table and writes the row out in
//      cx order.
// Ranks inside a tile come from LDS atomics, so the order of the particles
// of one cell is not kept; nothing downstream depends on it, every count
// being a sum over whole cells.

// Blocks go round-robin over the 8 XCDs; with a grid of 8 * ceil(ntiles / 8)
// blocks this gives each XCD one contiguous range of tiles, so the partial
// cache lines where one tile's digit run meets the next tile's are written
// through the same L2.  Returns a value >= ntiles for a block with no tile.
__device__ __forceinline__ uint32_t xcd_tile() {
  const uint32_t per = gridDim.x >> 3;
  return (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
}

// Items of a tile, and where its digit-d count lives: hist[hbase + d * hstride].
struct TileRange {
  uint64_t begin;
  uint32_t len;
  uint64_t hbase;
  uint32_t hstride;
};

__device__ __forceinline__ bool tile_range_chunk(uint32_t t, uint32_t ntiles, uint64_t n,
                                                 TileRange& r) {
  if (t >= ntiles) return false;
  r.begin = (uint64_t)t * kCsTile;
  const uint64_t left = n - r.begin;
  r.len = left < (uint64_t)kCsTile ? (uint32_t)left : (uint32_t)kCsTile;
  r.hbase = t;
  r.hstride = ntiles;
  return true;
}

// Tile t of the slab-major tiling: slab s starts at ss[s] and owns tiles
// [tb[s], tb[s+1]); tb[G] is the number of tiles.
__device__ __forceinline__ bool tile_range_slab(uint32_t t, const uint32_t* ss,
                                                const uint32_t* tb, int G, TileRange& r) {
  if (t >= tb[G]) return false;
  int lo = 0, hi = G - 1;  // the last slab s with tb[s] <= t; then t < tb[s + 1]
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tb[mid] <= t) lo = mid; else hi = mid - 1;
  }
  const uint32_t i = t - tb[lo];
  const uint32_t ts = tb[lo + 1] - tb[lo];
  r.begin = (uint64_t)ss[lo] + (uint64_t)i * kCsTile;
  const uint64_t left = (uint64_t)ss[lo + 1] - r.begin;
  r.len = left < (uint64_t)kCsTile ? (uint32_t)left : (uint32_t)kCsTile;
  r.hbase = (uint64_t)tb[lo] * (uint64_t)G + i;
  r.hstride = ts;
  return true;
}

// LDS histogram count of digit d for the lanes with `valid`.  The lanes that
// hold the digit of the wave's first valid lane add in one atomic: clustered
// particles put many lanes of a wave on one digit, and LDS atomics on one
// address serialise.  Called by the whole wave.
__device__ __forceinline__ void lds_count(uint32_t* cnt, uint32_t d, bool valid) {
  const uint64_t vm = __ballot(valid);
  if (!vm) return;
  const int first = __builtin_ctzll(vm);
  const uint32_t d0 = (uint32_t)__builtin_amdgcn_readlane((int)d, first);
  const uint64_t peers = __ballot(valid && d == d0);
  const bool lead = lane_id() == first;
  if (valid && (d != d0 || lead)) atomicAdd(&cnt[d], lead ? (uint32_t)__popcll(peers) : 1u);
}

// Long row k of the last pass: lrows[k] = (row cz * G + cy, first tile,
// first particle, length); tile t belongs to long row trow[t].  The row's
// tiles are kCsTile particles (the last one shorter); their cx histograms
// are tile-major: hist[(first tile + i) * GX + d].
struct LongTile {
  uint32_t row, tile0, start, len, i, ts;
};

__device__ __forceinline__ LongTile long_tile(uint32_t t, const uint4* __restrict__ lrows,
                                              const uint32_t* __restrict__ trow) {
  const uint4 L = lrows[trow[t]];
  return LongTile{L.x, L.y, L.z, L.w, t - L.y, (L.w + kCsTile - 1) / kCsTile};
}

__device__ __forceinline__ TileRange long_range(const LongTile& lt, int GX) {
  TileRange r;
  r.begin = (uint64_t)lt.start + (uint64_t)lt.i * kCsTile;
  const uint32_t left = lt.len - lt.i * (uint32_t)kCsTile;
  r.len = left < (uint32_t)kCsTile ? left : (uint32_t)kCsTile;
  r.hbase = (uint64_t)(lt.tile0 + lt.i) * (uint64_t)GX;
  r.hstride = 1;
  return r;
}

// Tiles [0, ntiles) of a persistent grid of a multiple of 8 blocks: XCD x
// takes the contiguous range [ntiles * x / 8, ntiles * (x + 1) / 8), its
// blocks in turn, so neighbouring tiles meet in one L2.
__device__ __forceinline__ void xcd_range(uint32_t ntiles, uint32_t& t, uint32_t& end,
                                          uint32_t& step) {
  const uint32_t x = blockIdx.x & 7u;
  step = gridDim.x >> 3;
  t = (uint32_t)((uint64_t)ntiles * x / 8) + (blockIdx.x >> 3);
  end = (uint32_t)((uint64_t)ntiles * (x + 1) / 8);
}

// Particle sources of the sort passes.  A pass loads what its items need
// first (load / load1<AXIS>: the raw words of particle i) and turns them into
// coordinates afterwards (finish / finish1<AXIS>), so that every item's loads
// are in flight before any of them is waited for.
struct Packed3 {
  float x, y, z;
};

struct SrcPacked {  // packed FP32 xyz (the file format)
  const float* __restrict__ xyz;
  using Raw = Packed3;
  using Raw1 = float;
  __device__ Raw load(uint64_t i) const { return reinterpret_cast<const Packed3*>(xyz)[i]; }
  __device__ void finish(uint64_t, const Raw& w, float& x, float& y, float& z) const {
    x = w.x;
    y = w.y;
    z = w.z;
  }
  template <int AXIS>
  __device__ Raw1 load1(uint64_t i) const {
    return xyz[3 * i + AXIS];
  }
  template <int AXIS>
  __device__ float finish1(uint64_t, Raw1 w) const {
    return w;
  }
};

struct SrcSoA {  // the sort's SoA buffers
  const float* __restrict__ x;
  const float* __restrict__ y;
  const float* __restrict__ z;
  using Raw = Packed3;
  using Raw1 = float;
  __device__ Raw load(uint64_t i) const { return Packed3{x[i], y[i], z[i]}; }
  __device__ void finish(uint64_t, const Raw& w, float& px, float& py, float& pz) const {
    px = w.x;
    py = w.y;
    pz = w.z;
  }
  template <int AXIS>
  __device__ Raw1 load1(uint64_t i) const {
    return AXIS == 0 ? x[i] : (AXIS == 1 ? y[i] : z[i]);
  }
  template <int AXIS>
  __device__ float finish1(uint64_t, Raw1 w) const {
    return w;
  }
};

// A compressed file's uint16 triplets, decoded where they are read (the
// decode is exact and the same wherever it runs).  Files of one tile per
// side only: the general decode holds too many registers per item for the
// scatter to keep all of a thread's items in flight.
struct SrcGadget {
  const uint16_t* __restrict__ uv;
  GadgetDecode g;
  struct Raw {
    uint32_t d[3];
  };
  using Raw1 = uint32_t;
  __device__ Raw load(uint64_t i) const {
    return Raw{{uv[3 * i], uv[3 * i + 1], uv[3 * i + 2]}};
  }
  __device__ void finish(uint64_t i, const Raw& w, float& x, float& y, float& z) const {
    float p[3];
    gadget_point<true>(g, i, w.d, p);
    x = p[0];
    y = p[1];
    z = p[2];
  }
  template <int AXIS>
  __device__ Raw1 load1(uint64_t i) const {
    return uv[3 * i + AXIS];
  }
  template <int AXIS>
  __device__ float finish1(uint64_t i, Raw1 w) const {
    int ip[3];
    gadget_grid<true>(g, i, ip);
    return gadget_coord(g, ip[AXIS], w);
  }
};

// Histogram of tile r's cells of coordinate AXIS of src (ndig <= kDig cells,
// 1/inv wide), written to hist[r.hbase + d * r.hstride].
template <int kDig, int AXIS, class Src>
__device__ __forceinline__ void hist_tile(const Src& src, float inv, int ndig, const TileRange& r,
                                          uint32_t* s_h, uint32_t* __restrict__ hist) {
  const int tid = threadIdx.x;
  for (int i = tid; i < kDig; i += kCsThreads) s_h[i] = 0;
  __syncthreads();
  typename Src::Raw1 w[kCsItems];
  if (r.len == (uint32_t)kCsTile) {  // full tile: no guard between the loads
#pragma unroll
    for (int j = 0; j < kCsItems; ++j) w[j] = src.template load1<AXIS>(r.begin + j * kCsThreads + tid);
  } else {
#pragma unroll
    for (int j = 0; j < kCsItems; ++j) {
      const uint32_t i = (uint32_t)(j * kCsThreads + tid);
      w[j] = typename Src::Raw1{};
      if (i < r.len) w[j] = src.template load1<AXIS>(r.begin + i);
    }
  }
  float v[kCsItems];
#pragma unroll
  for (int j = 0; j < kCsItems; ++j)
    v[j] = src.template finish1<AXIS>(r.begin + (uint32_t)(j * kCsThreads + tid), w[j]);
#pragma unroll
  for (int j = 0; j < kCsItems; ++j) {
    const bool ok = (uint32_t)(j * kCsThreads + tid) < r.len;
    lds_count(s_h, ok ? (uint32_t)cell_of(v[j], inv, ndig - 1) : 0u, ok);
  }
  __syncthreads();
  for (int d = tid; d < ndig; d += kCsThreads) hist[r.hbase + (uint64_t)d * r.hstride] = s_h[d];
}

// Per-tile histogram of the cells of coordinate AXIS (passes 1, 2).
template <int AXIS, bool SLAB, class Src>
__global__ __launch_bounds__(kCsThreads) void cs_hist_kernel(
    Src src, uint64_t n, float inv, int G, uint32_t ntiles, const uint32_t* __restrict__ ss,
    const uint32_t* __restrict__ tb, uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_h[kCsMaxCells];
  TileRange r;
  const uint32_t t = xcd_tile();
  if (!(SLAB ? tile_range_slab(t, ss, tb, G, r) : tile_range_chunk(t, ntiles, n, r))) return;
  hist_tile<kCsMaxCells, AXIS>(src, inv, G, r, s_h, hist);
}

// Scatter of tile r by the cell of coordinate AXIS (ndig <= kDig cells, 1/inv
// wide).  Without `lt` the digit runs go to the offsets of the scanned
// histograms hist[r.hbase + d * r.hstride].  With `lt` (a long row of pass
// 3) the offsets come from the row's tile-major counts: the row's start,
// plus the row's particles of lower digits, plus the digit's particles in
// the row's earlier tiles; the row's first tile also writes the row's cell
// table.  Input packed xyz (AOS) or SoA; output SoA.
template <int AXIS, int kDig, class Src>
__device__ __forceinline__ void scatter_tile(
    const Src& src,
    float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz, float inv, int ndig,
    const TileRange& r, const uint32_t* __restrict__ hist, const LongTile* lt,
    uint32_t* __restrict__ cell_end, float (*s_c)[kCsTile], uint32_t* s_cnt, uint32_t* s_off,
    uint32_t* s_tmp) {
  static_assert(kDig <= kCsThreads || kDig % kCsThreads == 0, "digits per thread");
  constexpr int Q = kDig > kCsThreads ? kDig / kCsThreads : 1;
  const int dm1 = ndig - 1;
  const int tid = threadIdx.x;
  for (int i = tid; i < kDig; i += kCsThreads) s_cnt[i] = 0;

  typename Src::Raw w[kCsItems];
  if (r.len == (uint32_t)kCsTile) {  // full tile: no guard between the loads
#pragma unroll
    for (int j = 0; j < kCsItems; ++j) w[j] = src.load(r.begin + j * kCsThreads + tid);
  } else {
#pragma unroll
    for (int j = 0; j < kCsItems; ++j) {
      const uint32_t i = (uint32_t)(j * kCsThreads + tid);
      w[j] = typename Src::Raw{};
      if (i < r.len) w[j] = src.load(r.begin + i);
    }
  }
  float px[kCsItems], py[kCsItems], pz[kCsItems];
#pragma unroll
  for (int j = 0; j < kCsItems; ++j)
    src.finish(r.begin + (uint32_t)(j * kCsThreads + tid), w[j], px[j], py[j], pz[j]);
  __syncthreads();

  uint32_t rk[kCsItems];  // digit << 16 | slot among the tile's items of that digit
#pragma unroll
  for (int j = 0; j < kCsItems; ++j) {
    rk[j] = 0;
    if ((uint32_t)(j * kCsThreads + tid) < r.len) {
      const float v = AXIS == 0 ? px[j] : (AXIS == 1 ? py[j] : pz[j]);
      const uint32_t d = (uint32_t)cell_of(v, inv, dm1);
      rk[j] = (d << 16) | atomicAdd(&s_cnt[d], 1u);
    }
  }
  __syncthreads();
  uint32_t c[Q], sum = 0;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int d = tid * Q + q;
    c[q] = d < kDig ? s_cnt[d] : 0u;
    sum += c[q];
  }
  uint32_t total;
  uint32_t ls = block_excl_scan<OpSum, kCsThreads>(sum, s_tmp, &total);
  if (lt) {
    // the digit's count in the row's earlier tiles (pre) and in all of them
    // (tot); consecutive threads read consecutive digits of each tile
    uint32_t pre[Q], tot[Q], tsum = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int d = tid * Q + q;
      pre[q] = tot[q] = 0u;
      if (d < ndig) {
        const uint32_t* hc = hist + (uint64_t)lt->tile0 * (uint64_t)ndig + (uint64_t)d;
        uint32_t i = 0;
#pragma unroll 4
        for (; i < lt->i; ++i) pre[q] += hc[(uint64_t)i * ndig];
        tot[q] = pre[q];
#pragma unroll 4
        for (; i < lt->ts; ++i) tot[q] += hc[(uint64_t)i * ndig];
      }
      tsum += tot[q];
    }
    uint32_t rtotal;
    uint32_t ds = lt->start + block_excl_scan<OpSum, kCsThreads>(tsum, s_tmp, &rtotal);
    uint32_t* ce = cell_end + (uint64_t)lt->row * (uint64_t)ndig;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int d = tid * Q + q;
      if (d < ndig) {
        s_off[d] = ds + pre[q] - ls;
        s_cnt[d] = ls;
        if (lt->i == 0) ce[d] = ds + tot[q];
      }
      ds += tot[q];
      ls += c[q];
    }
  } else {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int d = tid * Q + q;
      if (d < ndig) {
        // out index = offset of the tile's digit run + (slot - local start);
        // u32 arithmetic wraps back into range
        s_off[d] = hist[r.hbase + (uint64_t)d * r.hstride] - ls;
        s_cnt[d] = ls;
      }
      ls += c[q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kCsItems; ++j) {
    if ((uint32_t)(j * kCsThreads + tid) < r.len) {
      const uint32_t lp = s_cnt[rk[j] >> 16] + (rk[j] & 0xffffu);
      s_c[0][lp] = px[j];
      s_c[1][lp] = py[j];
      s_c[2][lp] = pz[j];
    }
  }
  __syncthreads();
  for (uint32_t l = tid; l < r.len; l += kCsThreads) {
    const float x = s_c[0][l], y = s_c[1][l], z = s_c[2][l];
    const float v = AXIS == 0 ? x : (AXIS == 1 ? y : z);
    const uint32_t gp = s_off[cell_of(v, inv, dm1)] + l;
    ox[gp] = x;
    oy[gp] = y;
    oz[gp] = z;
  }
}

// Scatter of one tile of passes 1 (by cz, chunk tiles) and 2 (by cy, slab
// tiles).
template <int AXIS, bool SLAB, class Src>
__global__ __launch_bounds__(kCsThreads) void cs_scatter_kernel(
    Src src,
    float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz, uint64_t n, Grid g,
    uint32_t ntiles, const uint32_t* __restrict__ ss, const uint32_t* __restrict__ tb,
    const uint32_t* __restrict__ hist) {
  __shared__ float s_c[3][kCsTile];
  __shared__ uint32_t s_cnt[kCsMaxCells];
  __shared__ uint32_t s_off[kCsMaxCells];
  __shared__ uint32_t s_tmp[kCsThreads / 64];
  TileRange r;
  const uint32_t t = xcd_tile();
  if (!(SLAB ? tile_range_slab(t, ss, tb, g.g, r) : tile_range_chunk(t, ntiles, n, r))) return;
  scatter_tile<AXIS, kCsMaxCells>(src, ox, oy, oz, g.inv_cell, g.g, r, hist, nullptr, nullptr, s_c,
                                  s_cnt, s_off, s_tmp);
}

// Slab starts and the slab-major tiling of pass 2, from pass 1's scanned
// histograms; clears the long-row counters of pass 3.  One block of
// kCsMaxCells threads.
__global__ __launch_bounds__(kCsMaxCells) void cs_slab_tiles_kernel(
    const uint32_t* __restrict__ hist1, uint32_t tiles1, uint64_t n, int G,
    uint32_t* __restrict__ ss, uint32_t* __restrict__ tb, uint32_t* __restrict__ lctr) {
  __shared__ uint32_t s_tmp[kCsMaxCells / 64];
  const int tid = threadIdx.x;
  const uint32_t st = tid < G ? hist1[(uint64_t)tid * tiles1] : 0u;
  const uint32_t en = tid + 1 < G ? hist1[(uint64_t)(tid + 1) * tiles1] : (uint32_t)n;
  const uint32_t nt = tid < G ? (uint32_t)(((uint64_t)(en - st) + kCsTile - 1) / kCsTile) : 0u;
  uint32_t total;
  const uint32_t excl = block_excl_scan<OpSum, kCsMaxCells>(nt, s_tmp, &total);
  if (tid < G) {
    ss[tid] = st;
    tb[tid] = excl;
  }
  if (tid == 0) {
    ss[G] = (uint32_t)n;
    tb[G] = total;
  }
  if (tid < 2) lctr[tid] = 0u;
}

// First particle of row rw = cz * G + cy after pass 2 (rw = G*G: n).
__device__ __forceinline__ uint32_t row_begin(uint32_t rw, int G, uint64_t n,
                                              const uint32_t* ss, const uint32_t* tb,
                                              const uint32_t* hist2) {
  if (rw >= (uint32_t)(G * G)) return (uint32_t)n;
  const uint32_t sl = rw / (uint32_t)G, d = rw - sl * (uint32_t)G;
  const uint32_t ts = tb[sl + 1] - tb[sl];
  return ts ? hist2[(uint64_t)tb[sl] * (uint64_t)G + (uint64_t)d * ts] : ss[sl];
}

// Exclusive scan of the row's cx counts s_cnt[0 .. kCsMaxCellsX) with
// kCsRowThreads threads, kCsMaxCellsX / kCsRowThreads consecutive digits
// each.  Writes the row's cell table entries (start + inclusive count) and
// leaves in s_cnt[d] the digit's first slot.
__device__ __forceinline__ void cs_row_scan(uint32_t* s_cnt, uint32_t* s_tmp, uint32_t* ce,
                                            uint32_t start, int gx) {
  constexpr int Q = kCsMaxCellsX / kCsRowThreads;
  const int tid = threadIdx.x;
  uint32_t c[Q], sum = 0;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    c[q] = s_cnt[tid * Q + q];
    sum += c[q];
  }
  uint32_t total;
  uint32_t run = block_excl_scan<OpSum, kCsRowThreads>(sum, s_tmp, &total);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int d = tid * Q + q;
    if (d < gx) {
      ce[d] = start + run + c[q];
      s_cnt[d] = run;
    }
    run += c[q];
  }
}

// Pass 3: every row in cx order, and the cell table.  Rows of up to
// kCsRowCap particles are sorted in LDS and written out contiguously, one
// block each.  A longer row is entered in the long-row list (lctr[0] counts
// tiles, lctr[1] rows) for the tiled kernels below.
__global__ __launch_bounds__(kCsRowThreads) void cs_row_kernel(
    const float* __restrict__ ix, const float* __restrict__ iy, const float* __restrict__ iz,
    float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz,
    uint32_t* __restrict__ cell_end, uint64_t n, Grid g, const uint32_t* __restrict__ ss,
    const uint32_t* __restrict__ tb, const uint32_t* __restrict__ hist2, uint32_t nrows,
    uint32_t* __restrict__ lctr, uint4* __restrict__ lrows, uint32_t* __restrict__ trow) {
  __shared__ float s_c[3][kCsRowCap];
  __shared__ uint32_t s_cnt[kCsMaxCellsX];
  __shared__ uint32_t s_tmp[kCsRowThreads / 64];
  const uint32_t rw = xcd_tile();
  if (rw >= nrows) return;
  const int G = g.g, GX = g.gx, gxm1 = GX - 1;
  const float invx = g.inv_cell_x;
  const int tid = threadIdx.x;
  const uint32_t start = row_begin(rw, G, n, ss, tb, hist2);
  const uint32_t len = row_begin(rw + 1, G, n, ss, tb, hist2) - start;
  if (len > (uint32_t)kCsRowCap) {
    const uint32_t nt = (len + kCsTile - 1) / kCsTile;
    if (tid == 0) {
      const uint32_t first = atomicAdd(&lctr[0], nt);
      const uint32_t k = atomicAdd(&lctr[1], 1u);
      lrows[k] = make_uint4(rw, first, start, len);
      s_tmp[0] = first;
      s_tmp[1] = k;
    }
    __syncthreads();
    const uint32_t first = s_tmp[0], k = s_tmp[1];
    for (uint32_t i = tid; i < nt; i += kCsRowThreads) trow[first + i] = k;
    return;
  }
  for (int i = tid; i < kCsMaxCellsX; i += kCsRowThreads) s_cnt[i] = 0;
  float px[kCsRowItems], py[kCsRowItems], pz[kCsRowItems];
#pragma unroll
  for (int j = 0; j < kCsRowItems; ++j) {
    const uint32_t i = (uint32_t)(j * kCsRowThreads + tid);
    px[j] = py[j] = pz[j] = 0.0f;
    if (i < len) {
      px[j] = ix[start + i];
      py[j] = iy[start + i];
      pz[j] = iz[start + i];
    }
  }
  __syncthreads();
  uint32_t* ce = cell_end + (uint64_t)rw * (uint64_t)GX;
  uint32_t rk[kCsRowItems];  // digit << 16 | slot among the row's items of that digit
#pragma unroll
  for (int j = 0; j < kCsRowItems; ++j) {
    rk[j] = 0;
    if ((uint32_t)(j * kCsRowThreads + tid) < len) {
      const uint32_t d = (uint32_t)cell_of(px[j], invx, gxm1);
      rk[j] = (d << 16) | atomicAdd(&s_cnt[d], 1u);
    }
  }
  __syncthreads();
  cs_row_scan(s_cnt, s_tmp, ce, start, GX);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kCsRowItems; ++j) {
    if ((uint32_t)(j * kCsRowThreads + tid) < len) {
      const uint32_t lp = s_cnt[rk[j] >> 16] + (rk[j] & 0xffffu);
      s_c[0][lp] = px[j];
      s_c[1][lp] = py[j];
      s_c[2][lp] = pz[j];
    }
  }
  __syncthreads();
  for (uint32_t l = tid; l < len; l += kCsRowThreads) {
    ox[start + l] = s_c[0][l];
    oy[start + l] = s_c[1][l];
    oz[start + l] = s_c[2][l];
  }
}

// Pass 3, long rows: per-tile cx histograms of every long row's tiles
// (persistent grid over the lctr[0] tiles).
__global__ __launch_bounds__(kCsThreads) void cs_long_hist_kernel(
    const float* __restrict__ x, Grid g, const uint32_t* __restrict__ lctr,
    const uint4* __restrict__ lrows, const uint32_t* __restrict__ trow,
    uint32_t* __restrict__ hist3) {
  __shared__ uint32_t s_h[kCsMaxCellsX];
  uint32_t t, end, step;
  xcd_range(lctr[0], t, end, step);
  for (; t < end; t += step) {
    const TileRange r = long_range(long_tile(t, lrows, trow), g.gx);
    __syncthreads();  // the previous tile's readers of s_h are done
    hist_tile<kCsMaxCellsX, 0>(SrcSoA{x, x, x}, g.inv_cell_x, g.gx, r, s_h, hist3);
  }
}

// Pass 3, long rows: every long row's tiles scattered by cx, and the long
// rows' cell table (persistent grid over the lctr[0] tiles).
__global__ __launch_bounds__(kCsThreads) void cs_long_scatter_kernel(
    const float* __restrict__ ix, const float* __restrict__ iy, const float* __restrict__ iz,
    float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz, Grid g,
    const uint32_t* __restrict__ lctr, const uint4* __restrict__ lrows,
    const uint32_t* __restrict__ trow, const uint32_t* __restrict__ hist3,
    uint32_t* __restrict__ cell_end) {
  __shared__ float s_c[3][kCsTile];
  __shared__ uint32_t s_cnt[kCsMaxCellsX];
  __shared__ uint32_t s_off[kCsMaxCellsX];
  __shared__ uint32_t s_tmp[kCsThreads / 64];
  uint32_t t, end, step;
  xcd_range(lctr[0], t, end, step);
  for (; t < end; t += step) {
    const LongTile lt = long_tile(t, lrows, trow);
    const TileRange r = long_range(lt, g.gx);
    __syncthreads();  // the previous tile's readers of the LDS are done
    scatter_tile<0, kCsMaxCellsX>(SrcSoA{ix, iy, iz}, ox, oy, oz, g.inv_cell_x, g.gx, r, hist3, &lt,
                                         cell_end, s_c, s_cnt, s_off, s_tmp);
  }
}

// -------------------------------------------------------------- key make ----

__global__ void particle_keys_kernel(const float* __restrict__ xyz, uint64_t n, Grid g,
                                     uint32_t* __restrict__ key, float* __restrict__ x,
                                     float* __restrict__ y, float* __restrict__ z) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    float px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
    key[i] = cell_key(px, py, pz, g);
    x[i] = px;
    y[i] = py;
    z[i] = pz;
  }
}

// Spread the low 10 bits of v to every third bit.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ void halo_keys_kernel(const char* __restrict__ pts, uint64_t stride, uint64_t n,
                                 int shift, uint32_t* __restrict__ key, float* __restrict__ x,
                                 float* __restrict__ y, float* __restrict__ z,
                                 uint32_t* __restrict__ row) {
  const float inv = 1024.0f / 1000.0f;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const float* c = reinterpret_cast<const float*>(pts + i * stride);
    float hx = c[0], hy = c[1], hz = c[2];
    uint32_t cx = (uint32_t)cell_of(hx, inv, 1023);
    uint32_t cy = (uint32_t)cell_of(hy, inv, 1023);
    uint32_t cz = (uint32_t)cell_of(hz, inv, 1023);
    key[i] = (spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2)) >> shift;
    x[i] = hx;
    y[i] = hy;
    z[i] = hz;
    row[i] = (uint32_t)i;
  }
}

// cell_end pass 1: the last index of every run of equal keys marks its cell.
__global__ void mark_cell_ends_kernel(const uint32_t* __restrict__ k, uint64_t n,
                                      uint32_t* __restrict__ end_of) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = k[i];
    if (i + 1 == n || k[i + 1] != c) end_of[c] = (uint32_t)(i + 1);
  }
}

uint32_t grid_blocks(uint64_t n) {
  uint64_t b = ceil_div(n, 256);
  return (uint32_t)(b < 262144 ? (b ? b : 1) : 262144);
}

}  // namespace

hipError_t exclusive_sum_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* ws,
                             hipStream_t s) {
  return scan_u32<OpSum, true>(in, out, n, ws, s);
}

hipError_t inclusive_max_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* ws,
                             hipStream_t s) {
  return scan_u32<OpMax, false>(in, out, n, ws, s);
}

hipError_t fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  fill_u32_kernel<<<fill_blocks(n), 256, 0, s>>>(p, v, n);
  return hipGetLastError();
}

hipError_t fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  fill_u64_kernel<<<fill_blocks(n), 256, 0, s>>>(p, v, n);
  return hipGetLastError();
}

hipError_t radix_sort_pairs(uint32_t* keys[2], uint32_t* pay[2][4], int npay, uint64_t n,
                            int key_bits, uint32_t* ws, hipStream_t s, int* result) {
  *result = 0;
  if (n <= 1 || key_bits <= 0) return hipSuccess;
  if (npay != 3 && npay != 4) return hipErrorInvalidValue;
  const uint64_t tiles = ceil_div(n, kSortTile);
  if (tiles > 0xffffffffull / kRadix) return hipErrorInvalidValue;
  uint32_t* hist = ws;
  uint32_t* scan_ws = ws + (uint64_t)kRadix * tiles;
  int cur = 0;
  for (int shift = 0; shift < key_bits; shift += kRadixBits) {
    const int nxt = cur ^ 1;
    radix_upsweep_kernel<<<(uint32_t)tiles, kSortThreads, 0, s>>>(keys[cur], n, shift, hist,
                                                                 (uint32_t)tiles);
    hipError_t e = exclusive_sum_u32(hist, hist, (uint64_t)kRadix * tiles, scan_ws, s);
    if (e != hipSuccess) return e;
    if (npay == 3) {
      radix_downsweep_kernel<3><<<(uint32_t)tiles, kSortThreads, 0, s>>>(
          keys[cur], pay[cur][0], pay[cur][1], pay[cur][2], nullptr, keys[nxt], pay[nxt][0],
          pay[nxt][1], pay[nxt][2], nullptr, n, shift, hist, (uint32_t)tiles);
    } else {
      radix_downsweep_kernel<4><<<(uint32_t)tiles, kSortThreads, 0, s>>>(
          keys[cur], pay[cur][0], pay[cur][1], pay[cur][2], pay[cur][3], keys[nxt],
          pay[nxt][0], pay[nxt][1], pay[nxt][2], pay[nxt][3], n, shift, hist,
          (uint32_t)tiles);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    cur = nxt;
  }
  *result = cur;
  return hipSuccess;
}

hipError_t cell_sort_particles(const ParticleSource& src, uint64_t n, Grid g, float* x[2],
                               float* y[2], float* z[2], uint32_t* cell_end, uint32_t* ws,
                               hipStream_t s) {
  const int G = g.g;
  if (G > kCsMaxCells || G < 1 || g.gx > kCsMaxCellsX || g.gx < 1) return hipErrorInvalidValue;
  if (n > 0xffffffffull || (src.uv && !src.gd.one_tile)) return hipErrorInvalidValue;
  if (n == 0) return fill_u32(cell_end, 0u, g.ncells, s);
  const uint64_t tiles1 = ceil_div(n, kCsTile);
  const uint64_t tiles2 = tiles1 + (uint64_t)G;  // each slab adds at most one partial tile
  if (tiles2 * (uint64_t)G > 0xffffffffull) return hipErrorInvalidValue;
  // the long rows' histograms (pass 3) reuse the words of passes 1 and 2
  uint32_t* hist1 = ws;
  uint32_t* hist2 = hist1 + (uint64_t)kCsMaxCells * tiles1;
  uint32_t* hist3 = ws;
  uint4* lrows = reinterpret_cast<uint4*>(ws + cs_hist_words(n));
  uint32_t* trow = reinterpret_cast<uint32_t*>(lrows + cs_long_rows_max(n));
  uint32_t* ss = trow + cs_long_tiles_max(n);
  uint32_t* tb = ss + kCsMaxCells + 1;
  uint32_t* lctr = tb + kCsMaxCells + 1;
  uint32_t* scan_ws = lctr + 4;
  const uint32_t nt1 = (uint32_t)tiles1;
  const uint32_t nb1 = (uint32_t)(8 * ceil_div(tiles1, 8));
  const uint32_t nb2 = (uint32_t)(8 * ceil_div(tiles2, 8));
  const uint32_t nrows = (uint32_t)(G * G);
  const uint32_t nb3 = (uint32_t)(8 * ceil_div(nrows, 8));
  // pass 1: by cz, packed xyz (or compressed triplets decoded as they are
  // read) -> buffers 0
  hipError_t e = hipSuccess;
  auto pass1 = [&](auto in) {
    cs_hist_kernel<2, false><<<nb1, kCsThreads, 0, s>>>(in, n, g.inv_cell, G, nt1, nullptr,
                                                         nullptr, hist1);
    e = exclusive_sum_u32(hist1, hist1, (uint64_t)G * tiles1, scan_ws, s);
    if (e == hipSuccess)
      cs_scatter_kernel<2, false><<<nb1, kCsThreads, 0, s>>>(in, x[0], y[0], z[0], n, g, nt1,
                                                              nullptr, nullptr, hist1);
  };
  if (!src.uv)
    pass1(SrcPacked{src.xyz});
  else
    pass1(SrcGadget{src.uv, src.gd});
  if (e != hipSuccess) return e;
  cs_slab_tiles_kernel<<<1, kCsMaxCells, 0, s>>>(hist1, nt1, n, G, ss, tb, lctr);
  // pass 2: by cy inside each slab, buffers 0 -> 1
  cs_hist_kernel<1, true><<<nb2, kCsThreads, 0, s>>>(SrcSoA{x[0], y[0], z[0]}, n, g.inv_cell, G,
                                                      0u, ss, tb, hist2);
  e = exclusive_sum_u32(hist2, hist2, (uint64_t)G * tiles2, scan_ws, s);
  if (e != hipSuccess) return e;
  cs_scatter_kernel<1, true><<<nb2, kCsThreads, 0, s>>>(SrcSoA{x[0], y[0], z[0]}, x[1], y[1], z[1],
                                                         n, g, 0u, ss, tb, hist2);
  // pass 3: by cx inside each row, buffers 1 -> 0, and the cell table; rows
  // of more than kCsRowCap particles in kCsTile tiles, like passes 1 and 2
  cs_row_kernel<<<nb3, kCsRowThreads, 0, s>>>(x[1], y[1], z[1], x[0], y[0], z[0], cell_end, n, g,
                                              ss, tb, hist2, nrows, lctr, lrows, trow);
  const uint64_t long_rows = cs_long_rows_max(n) < nrows ? cs_long_rows_max(n) : nrows;
  if (long_rows > 0) {
    // persistent grids: the scatter's LDS allows one block per CU
    const uint64_t per_xcd = ceil_div(cs_long_tiles_max(n), 8);
    const uint32_t nbh = (uint32_t)(8 * (per_xcd < 64 ? per_xcd : 64));
    const uint32_t nbc = (uint32_t)(8 * (per_xcd < 32 ? per_xcd : 32));
    cs_long_hist_kernel<<<nbh, kCsThreads, 0, s>>>(x[1], g, lctr, lrows, trow, hist3);
    cs_long_scatter_kernel<<<nbc, kCsThreads, 0, s>>>(x[1], y[1], z[1], x[0], y[0], z[0], g, lctr,
                                                       lrows, trow, hist3, cell_end);
  }
  return hipGetLastError();
}

hipError_t particle_keys(const float* xyz, uint64_t n, Grid g, uint32_t* key, float* x,
                         float* y, float* z, hipStream_t s) {
  if (n == 0) return hipSuccess;
  particle_keys_kernel<<<grid_blocks(n), 256, 0, s>>>(xyz, n, g, key, x, y, z);
  return hipGetLastError();
}

hipError_t halo_keys(const float* pts, size_t stride_bytes, uint64_t n, uint32_t* key, float* x,
                     float* y, float* z, uint32_t* row, hipStream_t s, int key_bits) {
  if (n == 0) return hipSuccess;
  if (key_bits < 1 || key_bits > 30) return hipErrorInvalidValue;
  halo_keys_kernel<<<grid_blocks(n), 256, 0, s>>>(reinterpret_cast<const char*>(pts),
                                                  (uint64_t)stride_bytes, n, 30 - key_bits, key,
                                                  x, y, z, row);
  return hipGetLastError();
}

// flag = 1 when keys[i] == keys[i + run - 1] for some i (a run of at least
// `run` equal keys, the keys being sorted).
__global__ void key_runs_kernel(const uint32_t* __restrict__ k, uint64_t n, uint32_t run,
                                uint32_t* __restrict__ flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i + run - 1 < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (k[i] == k[i + run - 1]) *flag = 1u;
}

hipError_t has_key_runs(const uint32_t* keys, uint64_t n, uint32_t run, uint32_t* flag,
                        hipStream_t s) {
  hipError_t e = fill_u32(flag, 0u, 1, s);
  if (e != hipSuccess || run < 1 || n < run) return e;
  key_runs_kernel<<<grid_blocks(n), 256, 0, s>>>(keys, n, run, flag);
  return hipGetLastError();
}

hipError_t build_cell_end(const uint32_t* sorted_keys, uint64_t n, uint32_t* cell_end,
                          uint32_t ncells, uint32_t* ws, hipStream_t s) {
  hipError_t e = fill_u32(cell_end, 0u, ncells, s);
  if (e != hipSuccess || n == 0) return e;
  mark_cell_ends_kernel<<<grid_blocks(n), 256, 0, s>>>(sorted_keys, n, cell_end);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return inclusive_max_u32(cell_end, cell_end, ncells, ws, s);
}


namespace {

// Points j * step (j < m) counted into their cells; `stride` is the element
// stride of one point times step, so point j * step is element j * stride.
__global__ void occupancy_count_kernel(const float* __restrict__ px, const float* __restrict__ py,
                                       const float* __restrict__ pz, uint64_t stride, uint64_t m,
                                       uint32_t* __restrict__ cells) {
  const float inv = (float)kOccCells / 1000.0f;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = j * stride;  // element index: point j * step of the input
    const uint32_t cx = (uint32_t)cell_of(px[i], inv, kOccCells - 1);
    const uint32_t cy = (uint32_t)cell_of(py[i], inv, kOccCells - 1);
    const uint32_t cz = (uint32_t)cell_of(pz[i], inv, kOccCells - 1);
    atomicAdd(cells + (cz * kOccCells + cy) * kOccCells + cx, 1u);
  }
}

__global__ void occupancy_hist_kernel(const uint32_t* __restrict__ cells, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kOccBins];
  if (threadIdx.x < kOccBins) h[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t ncell = (uint32_t)kOccCells * kOccCells * kOccCells;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x) {
    const uint32_t v = cells[c];
    atomicAdd(&h[v ? 32 - __builtin_clz(v) : 0], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kOccBins && h[threadIdx.x]) atomicAdd(hist + threadIdx.x, h[threadIdx.x]);
}

}  // namespace

hipError_t occupancy_histogram(const float* px, const float* py, const float* pz, uint32_t stride,
                               uint64_t n, uint32_t* ws, uint32_t* hist, hipStream_t s) {
  hipError_t e = fill_u32(ws, 0u, occupancy_workspace_words(), s);
  if (e == hipSuccess) e = fill_u32(hist, 0u, kOccBins, s);
  if (e != hipSuccess) return e;
  if (n) {
    // every step-th point, at most kOccSample of them: a description, not
    // a count (clustered points make the cell atomics contend)
    const uint64_t step = ceil_div(n, kOccSample);
    const uint64_t m = ceil_div(n, step);
    uint64_t blocks = ceil_div(m, 256);
    if (blocks > 16384) blocks = 16384;
    occupancy_count_kernel<<<(uint32_t)blocks, 256, 0, s>>>(px, py, pz, (uint64_t)stride * step, m,
                                                             ws);
  }
  occupancy_hist_kernel<<<256, 256, 0, s>>>(ws, hist);
  return hipGetLastError();
}

}  // namespace sjcs
