// engine.hip -- the C-ABI (include/sjcs.h): device state, chunk pipeline,
// and the ArraySTR-shaped index API.
//
// Per particle chunk (one reference particle file, src/astr_rcount/main.c:
// 246-262): upload -> cell keys -> LSD radix sort -> cell table -> count of
// every halo, added to u64 counts that stay on the device until the caller
// asks for them.  There is no CPU fallback: every entry point that computes
// needs a HIP device and fails with SJCS_ERR_DEVICE without one.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sjcs.h"
#include "common.h"
#include "count.h"
#include "gadget.h"
#include "primitives.h"
#include "rtree.h"
#include "tile.h"

using namespace sjcs;

namespace {

thread_local std::string g_create_error;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  template <class T>
  T* as() const { return static_cast<T*>(p); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// Grow-only device buffer.
hipError_t ensure(DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return hipSuccess;
  b.release();
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return e;
  }
  b.bytes = bytes;
  return hipSuccess;
}

// Grow-only device buffer that grows by half at least.
hipError_t ensure_grow(DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && bytes) return hipSuccess;
  return ensure(b, std::max(bytes, b.bytes + b.bytes / 2));
}

ParticleSource packed(const float* xyz) {
  ParticleSource s;
  s.xyz = xyz;
  return s;
}

struct ShellState {
  int nrads = 0;
  std::vector<float> r2;
  DevBuf d_r2, d_lut;
  Shells view{};
};

// Halos in Morton order (30-bit keys of a 1024^3 grid): a halo's count row
// is its position in that order, fixed for the halo set.  Tiles are runs of
// halos whose keys agree above their low 3*level bits -- level 0: one Morton
// cell (halos within about a unit, runs of at least kTileMin); level L: a
// cell of 2^L grid cells per side (runs of at least 2).  For the count the
// tiled runs go first, in order, then every other halo ("working order",
// slot[i] = count row of working halo i); runs longer than tile_cap are cut
// into tiles.
constexpr uint32_t kTileMin = 4;
constexpr int kTileLevelMax = 6;

struct HaloState {
  uint64_t n = 0;
  DevBuf bx, by, bz, brow, bkey;  // Morton order: centres, input rows, keys
  DevBuf inv;                      // input row -> count row
  bool runs0 = false;              // some kTileMin halos share a Morton key
  std::vector<uint32_t> hkeys;     // host copy of bkey, fetched for tiling
  int level = -1;                  // tile level of the working order
  bool permuted = false;           // working order differs from Morton order
  DevBuf x, y, z, slot;            // working order (when permuted)
  uint64_t n_tiled = 0;
  std::vector<uint32_t> run_start, run_len;  // runs, in working positions
  DevBuf tstart, tlen;  // tiles for tile_cap == tcap
  DevBuf twork;         // count_tiles workspace
  uint32_t ntiles = 0;
  uint32_t tmax = 0;  // longest tile
  int tcap = 0;
  void clear_tiles() {
    permuted = false;
    n_tiled = 0;
    run_start.clear();
    run_len.clear();
    ntiles = 0;
    tmax = 0;
    tcap = 0;
  }
  void release() {
    for (DevBuf* d : {&bx, &by, &bz, &brow, &bkey, &inv, &x, &y, &z, &slot, &tstart, &tlen, &twork})
      d->release();
    hkeys.clear();
    runs0 = false;
    level = -1;
    clear_tiles();
  }
};

// A built particle index: sorted SoA coordinates and the cell table, plus
// the sort's ping-pong buffers.
struct IndexBuffers {
  DevBuf key[2], x[2], y[2], z[2];
  DevBuf cell_end;
  Grid grid{};
  uint64_t n = 0;
  int sel = 0;
  void release() {
    for (int i = 0; i < 2; ++i) {
      key[i].release();
      x[i].release();
      y[i].release();
      z[i].release();
    }
    cell_end.release();
  }
  IndexView view() const {
    IndexView v;
    v.x = x[sel].as<float>();
    v.y = y[sel].as<float>();
    v.z = z[sel].as<float>();
    v.cell_end = cell_end.as<uint32_t>();
    v.grid = grid;
    v.n = n;
    return v;
  }
};

int key_bits_for(uint32_t ncells) {
  int b = 0;
  while (b < 32 && (1ull << b) < (uint64_t)ncells) ++b;
  return b;
}

// Cell edge along y and z: 0.9 of the query half-width by default (a query
// box then spans about 3 rows per axis, and its sphere cuts most of each
// row), coarsened until there are at most ~2 cells per particle; about 8
// particles per cell (at most 256 cells per axis) when the radius is not
// known.  At most 1024 cells per
// axis.  Along x the cells are `x_split` times finer (4 by default), within
// 1024 cells and ~4 cells per particle.
Grid choose_grid(uint64_t n, float rbox, int64_t cell_milli, int64_t x_split) {
  double s;
  if (cell_milli > 0)
    s = (double)cell_milli / 1000.0;
  else if (rbox > 0.0f && isfinite(rbox))
    s = (double)rbox * 0.9;
  else
    s = 1000.0 / std::min(256.0, std::max(1.0, cbrt((double)n / 8.0)));
  double gd = ceil(1000.0 / s);
  if (!(gd >= 1.0)) gd = 1.0;
  if (gd > 1024.0) gd = 1024.0;
  const double cells_cap = (double)std::max<uint64_t>(4096, 2 * n);
  if (cell_milli <= 0 && gd > cbrt(cells_cap)) gd = floor(cbrt(cells_cap));
  if (gd < 1.0) gd = 1.0;
  double k = (double)(x_split > 0 ? x_split : 4);
  const double kcap = floor(2.0 * cells_cap / (gd * gd * gd));
  if (x_split <= 0 && k > kcap) k = std::max(1.0, kcap);
  double gxd = std::min(1024.0, gd * k);
  if (gxd < gd) gxd = gd;
  Grid g;
  g.g = (int)gd;
  g.gx = (int)gxd;
  g.inv_cell = (float)g.g / 1000.0f;
  g.cell = 1000.0f / (float)g.g;
  g.inv_cell_x = (float)g.gx / 1000.0f;
  g.ncells = (uint32_t)g.g * (uint32_t)g.g * (uint32_t)g.gx;
  return g;
}

uint32_t float_bits(float f) {
  uint32_t b;
  memcpy(&b, &f, 4);
  return b;
}

float bits_float(uint32_t b) {
  float f;
  memcpy(&f, &b, 4);
  return f;
}

}  // namespace

// Device scratch of sjcs_index_count_noalloc, for callers that cannot
// allocate device memory: a fixed array in the library's code object, one
// copy on every device the library runs on, held until the process ends.
// Host code never passes `scratch` itself to a kernel or a copy (on the host
// it names a stand-in, not device memory): it uses the device address
// hipGetSymbolAddress gives, looked up once per device (scratch_address).
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be current.
hipError_t scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t h = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (h != hipSuccess) return h;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}
}  // namespace

struct sjcs_engine {
  int device = 0;
  void* scratch = nullptr;  // device address of `scratch` on this device
  hipStream_t stream = nullptr;   // counts, halo sorts, the index API
  hipStream_t bstream = nullptr;  // chunk uploads and cell sorts
  std::string err;

  bool have_shells = false;
  ShellState shells;
  bool have_halos = false;
  HaloState halos;

  // Counts, rows in sorted halo order: u32 running counts of the chunks
  // since the last fold, and u64 totals they are folded into before any
  // count could reach 2^32 (a particle counts at most once per image).
  DevBuf counts;    // u32 [nhalos][K]
  DevBuf counts64;  // u64 [nhalos][K]
  DevBuf ucounts;   // input order, for read-out: u64, or u32 (readout_u32)
  bool counts_fresh = true;  // the u32 counts hold nothing: next chunk stores
  bool c64_live = false;     // counts64 holds totals
  uint64_t pending = 0;      // bound on any u32 count
  uint64_t counts_gen = 1;   // changes whenever the totals may change
  uint64_t readout_gen = 0;  // totals `ucounts` holds (input row order)
  bool readout_u32 = false;  // ... as u32 (only while !c64_live)


  // Chunk i uses slot i % 2: its upload (bstream) overlaps the count of
  // chunk i - 1 (stream).  Its cell sort waits for that count ("overlap" 0,
  // the default: sort and count kernels would compete for the same CUs) or
  // only for the count that last read the slot ("overlap" 1).
  IndexBuffers ibs[2];
  DevBuf cin[2];              // packed chunk staged on the device (host uploads)
  DevBuf cuv[2];              // compressed triplets of a chunk, before the decode
  void* cpin[2] = {nullptr, nullptr};  // pinned host staging of a chunk
  size_t cpin_bytes[2] = {0, 0};
  int slot = 0;               // slot of the next chunk
  int last = -1;              // slot of the last chunk
  hipEvent_t built[2] = {nullptr, nullptr};     // slot's sort done
  hipEvent_t counted[2] = {nullptr, nullptr};   // slot's count done
  hipEvent_t uploaded[2] = {nullptr, nullptr};  // slot's pinned buffer read
  // every read of the slot's device staging (cin / cuv) queued so far is
  // done: the pieces of one chunk alternate between slots but all read the
  // staging of the slot the chunk was uploaded to, and in graph mode their
  // sorts run on `stream`, not in upload order on `bstream`
  hipEvent_t stage_read[2] = {nullptr, nullptr};
  DevBuf in;  // halo / query staging on the device
  DevBuf ws;  // sort / scan workspace
  DevBuf inbox;
  void* pinned = nullptr;  // pinned staging of halos / queries, and of count downloads
  size_t pinned_bytes = 0;
  hipEvent_t dl[2] = {nullptr, nullptr};  // a download piece reached page-locked memory

  int64_t cell_size_milli = 0;
  int64_t kernel = 0;
  int64_t lanes_per_halo = 0;
  int64_t segment = 0;
  int64_t index_sort = 0;
  int64_t x_split = 0;
  int64_t overlap = 0;
  int64_t fold_limit = 0;  // 0: fold before a u32 count could pass 2^32 - 1
  int64_t piece_max = 0;   // 0: pieces as large as 32-bit offsets and counts allow
  int64_t graph = 0;       // 1: each piece's sort + count replayed from a graph
  int64_t tiles = 1;       // 1: halo tiles share their candidates (ascending shells)
  int64_t tile_level = 0;  // halos tiled together: one Morton cell of 2^level grid cells
  int64_t tile_halos = 0;  // most halos per tile (0: as many as the LDS histograms hold)
  int64_t stage_particles = 0;  // particles a count workgroup stages in LDS (0: none)
  int64_t stage_halos = 0;      // halos per count workgroup when staging (0: 64)
  int64_t autotune_cache = 1;  // 1: sjcs_autotune reuses the setting found for a workload
  bool autotune_hit = false;   // the last sjcs_autotune took its setting from the cache

  // Captured sort + count of a piece, per slot and store/accumulate; valid
  // while input, size and `epoch` match.  `epoch` changes whenever a pointer
  // or a value captured in a graph may have changed.
  struct ChunkGraph {
    hipGraphExec_t exec = nullptr;
    const void* in = nullptr;     // the piece's particles (packed xyz or triplets)
    GadgetDecode gd{};            // triplets: the decode's constants
    uint64_t n = 0;
    uint64_t epoch = 0;
  };
  ChunkGraph graphs[2][2];
  uint64_t epoch = 1;

  // last chunk: upload start/end, sort start/end, count start/end
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  bool timed = false;
  // last halo preparation (keys, Morton sort, tiling) and last read-out
  // of the counts into input order: start/end each
  hipEvent_t hev[2] = {nullptr, nullptr};
  hipEvent_t rev[2] = {nullptr, nullptr};
  bool halo_timed = false, readout_timed = false;
};

struct sjcs_index {
  sjcs_engine* e = nullptr;
  uint64_t cap = 0;
  uint64_t n = 0;
  bool built = false;
  DevBuf aos;
  IndexBuffers ib;
};

namespace {

int fail(sjcs_engine* e, int code, const std::string& msg) {
  if (e) e->err = msg;
  return code;
}

int hip_fail(sjcs_engine* e, hipError_t h, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(h);
  return fail(e, h == hipErrorOutOfMemory ? SJCS_ERR_NOMEM : SJCS_ERR_DEVICE, m);
}

#define HIPCALL(e, expr, what)                          \
  do {                                                  \
    hipError_t h_ = (expr);                             \
    if (h_ != hipSuccess) return hip_fail((e), h_, (what)); \
  } while (0)

int use_device(sjcs_engine* e) {
  HIPCALL(e, hipSetDevice(e->device), "hipSetDevice");
  return SJCS_OK;
}

// Waits for all work of the engine (both streams).
int quiesce(sjcs_engine* e) {
  HIPCALL(e, hipStreamSynchronize(e->bstream), "hipStreamSynchronize");
  HIPCALL(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
  return SJCS_OK;
}

// Pinned host staging of at least `bytes`; waits for all work first so a
// previous upload from it has finished.
int ensure_pinned(sjcs_engine* e, size_t bytes) {
  int rc = quiesce(e);
  if (rc != SJCS_OK) return rc;
  if (e->pinned_bytes >= bytes) return SJCS_OK;
  if (e->pinned) (void)hipHostFree(e->pinned);
  e->pinned = nullptr;
  e->pinned_bytes = 0;
  HIPCALL(e, hipHostMalloc(&e->pinned, bytes, hipHostMallocDefault), "hipHostMalloc");
  e->pinned_bytes = bytes;
  return SJCS_OK;
}

// Host threads moving one page-locked piece into the caller's memory.
void parallel_copy(char* dst, const char* src, size_t bytes) {
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t nt = std::min<size_t>(std::max(hw, 1u), 16u);
  const size_t part = ((bytes + nt - 1) / nt + 4095) & ~(size_t)4095;
  std::vector<std::thread> ts;
  for (size_t t = 1; t < nt && t * part < bytes; ++t)
    ts.emplace_back([=] { memcpy(dst + t * part, src + t * part, std::min(part, bytes - t * part)); });
  memcpy(dst, src, std::min(part, bytes));
  for (std::thread& t : ts) t.join();
}

// Device memory -> the caller's (pageable) host memory.  A large copy goes
// in pieces through two page-locked buffers: piece i + 1 crosses the link
// while host threads move piece i into place (a copy straight into pageable
// memory runs at a fraction of the link's rate).  The engine is idle when
// this is called.
constexpr size_t kDownloadPiece = 64ull << 20;
int download(sjcs_engine* e, void* dst, const void* d_src, size_t bytes) {
  if (bytes <= kDownloadPiece) {
    HIPCALL(e, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, e->stream),
            "hipMemcpy(counts)");
    HIPCALL(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    return SJCS_OK;
  }
  int rc = ensure_pinned(e, 2 * kDownloadPiece);
  if (rc != SJCS_OK) return rc;
  char* pin = static_cast<char*>(e->pinned);
  const char* src = static_cast<const char*>(d_src);
  char* out = static_cast<char*>(dst);
  const size_t npieces = (bytes + kDownloadPiece - 1) / kDownloadPiece;
  for (size_t i = 0; i <= npieces; ++i) {
    if (i < npieces) {
      // buffer i % 2 held piece i - 2, moved out in the previous iteration
      const size_t off = i * kDownloadPiece, len = std::min(kDownloadPiece, bytes - off);
      HIPCALL(e, hipMemcpyAsync(pin + (i % 2) * kDownloadPiece, src + off, len,
                                hipMemcpyDeviceToHost, e->stream),
              "hipMemcpy(counts)");
      HIPCALL(e, hipEventRecord(e->dl[i % 2], e->stream), "hipEventRecord");
    }
    if (i >= 1) {
      const size_t j = i - 1, off = j * kDownloadPiece;
      HIPCALL(e, hipEventSynchronize(e->dl[j % 2]), "hipEventSynchronize");
      parallel_copy(out + off, pin + (j % 2) * kDownloadPiece,
                    std::min(kDownloadPiece, bytes - off));
    }
  }
  return SJCS_OK;
}

int ensure_ws(sjcs_engine* e, uint64_t words) {
  if (e->ws.bytes < words * 4) ++e->epoch;
  HIPCALL(e, ensure(e->ws, words * 4), "hipMalloc(workspace)");
  return SJCS_OK;
}

// Shell view (device pointers left null) and first-match table of nrads
// squared thresholds, host only: the table's entries go to lut[0, *nlut),
// at most kLutSize of them (one entry when the thresholds are not sorted).
void shell_table_into(const float* r2, int nrads, Shells* out, ShellEntry* lut, int* nlut_out) {
  bool sorted = true;
  for (int k = 0; k < nrads && sorted; ++k) {
    uint32_t b = float_bits(r2[k]);
    if (b >= 0x7f800001u) sorted = false;  // NaN or sign bit set
    if (k > 0 && !(r2[k - 1] <= r2[k])) sorted = false;
  }
  Shells v{};
  v.nrads = nrads;
  v.sorted = sorted ? 1 : 0;
  v.r2max = r2[nrads - 1];
  v.r2max_hi = r2[nrads - 1] * (1.0f + 1e-5f);
  // src/objects/ArraySTR/MultiRangeCount.c:19: float radius = sqrt(r2[K-1])
  v.rbox = (float)sqrt((double)r2[nrads - 1]);
  lut[0] = ShellEntry{0u, 0.0f};
  *nlut_out = 1;
  if (sorted) {
    // prefer a table based at bits 0 (one shift finds the bucket); fall back
    // to one based at r2[0] when buckets that coarse hold two thresholds
    const uint32_t top = float_bits(r2[nrads - 1]);
    for (int attempt = 0; attempt < 2; ++attempt) {
      const uint32_t base = attempt == 0 ? 0u : float_bits(r2[0]);
      const uint32_t range = top - base;
      int shift = 0;
      while ((range >> shift) >= (uint32_t)kLutSize) ++shift;
      const int nlut = (int)(range >> shift) + 1;
      bool one_step = true;
      for (int b = 0; b < nlut; ++b) {
        // smallest and largest distance of bucket b (bucket 0 also takes
        // every distance below the base; the last ends below r2max)
        const float dmin = b == 0 ? 0.0f : bits_float(base + ((uint32_t)b << shift));
        const uint32_t hib = b + 1 < nlut ? base + ((uint32_t)(b + 1) << shift) - 1u
                                          : (top > base ? top - 1u : top);
        const float dmax = bits_float(hib);
        int k = 0;
        while (k < nrads - 1 && !(dmin < r2[k])) ++k;
        lut[b].off = 4u * (uint32_t)k;
        lut[b].t = r2[k];
        // k + 1 is the answer for any d >= r2[k] of this bucket only if the
        // next threshold lies above the bucket
        if (k + 1 < nrads && !(dmax < r2[k + 1])) one_step = false;
      }
      if (!one_step && attempt == 0) continue;
      v.lut_base = base;
      v.lut_shift = shift;
      v.nlut = nlut;
      v.one_step = one_step ? 1 : 0;
      *nlut_out = nlut;
      break;
    }
  }
  *out = v;
}

void shell_table(const float* r2, int nrads, Shells* out, std::vector<ShellEntry>* out_lut) {
  std::vector<ShellEntry> lut(kLutSize);
  int nlut = 1;
  shell_table_into(r2, nrads, out, lut.data(), &nlut);
  lut.resize(nlut);
  out_lut->swap(lut);
}

// Validates and uploads shell thresholds; builds the first-match table.
int make_shells(sjcs_engine* e, ShellState& st, const float* r2, int nrads) {
  if (!r2 || nrads <= 0) return fail(e, SJCS_ERR_INVALID, "nrads must be >= 1");
  st.nrads = nrads;
  st.r2.assign(r2, r2 + nrads);
  Shells v{};
  std::vector<ShellEntry> lut;
  shell_table(r2, nrads, &v, &lut);
  HIPCALL(e, ensure(st.d_r2, sizeof(float) * nrads), "hipMalloc(shells)");
  HIPCALL(e, ensure(st.d_lut, sizeof(ShellEntry) * lut.size()), "hipMalloc(shell table)");
  HIPCALL(e, hipMemcpyAsync(st.d_r2.p, r2, sizeof(float) * nrads, hipMemcpyHostToDevice, e->stream),
          "hipMemcpy(shells)");
  HIPCALL(e, hipMemcpyAsync(st.d_lut.p, lut.data(), sizeof(ShellEntry) * lut.size(),
                            hipMemcpyHostToDevice, e->stream),
          "hipMemcpy(shell table)");
  HIPCALL(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
  v.r2 = st.d_r2.as<float>();
  v.lut = st.d_lut.as<ShellEntry>();
  st.view = v;
  return SJCS_OK;
}

// Working order of the halos for tile level `level` (see HaloState):
// runs found on the host from the Morton keys, the tiled runs moved to the
// front on the device.  Leaves the Morton order as working order when no
// run qualifies.  Callers make sure no count in flight reads the halos.
int apply_tiles(sjcs_engine* e, HaloState& hs, int level) {
  hs.clear_tiles();
  hs.level = level;
  const uint64_t n = hs.n;
  const uint32_t minrun = level == 0 ? kTileMin : 2;
  if (n < minrun || (level == 0 && !hs.runs0)) return SJCS_OK;
  if (hs.hkeys.size() != n) {
    hs.hkeys.resize(n);
    HIPCALL(e, hipMemcpy(hs.hkeys.data(), hs.bkey.p, 4 * n, hipMemcpyDeviceToHost), "halo keys");
  }
  const int shift = 3 * level;
  const std::vector<uint32_t>& keys = hs.hkeys;
  std::vector<uint32_t> rs, rl;
  uint64_t ntiled = 0;
  // runs of one Morton key first (halos at the same spot are evaluated once
  // per tile); at a level above 0, runs of the coarser cell among the halos
  // left between them
  for (uint64_t i = 0; i < n;) {
    uint64_t j = i + 1;
    while (j < n && keys[j] == keys[i]) ++j;
    if (j - i >= kTileMin) {
      rs.push_back((uint32_t)i);
      rl.push_back((uint32_t)(j - i));
      ntiled += j - i;
      i = j;
      continue;
    }
    if (level == 0) {
      i = j;
      continue;
    }
    // a run of the level's cell, stopping before the next run of one key
    uint64_t k = i;
    while (k < n && (keys[k] >> shift) == (keys[i] >> shift)) {
      uint64_t m = k + 1;
      while (m < n && keys[m] == keys[k]) ++m;
      if (m - k >= kTileMin) break;
      k = m;
    }
    if (k - i >= minrun) {
      rs.push_back((uint32_t)i);
      rl.push_back((uint32_t)(k - i));
      ntiled += k - i;
    }
    i = k > i ? k : j;
  }
  if (!ntiled) return SJCS_OK;
  std::vector<uint32_t> perm;
  perm.reserve(n);
  std::vector<uint32_t> nstart(rs.size());
  for (size_t r = 0; r < rs.size(); ++r) {
    nstart[r] = (uint32_t)perm.size();
    for (uint32_t k = 0; k < rl[r]; ++k) perm.push_back(rs[r] + k);
  }
  for (uint64_t i = 0, r = 0; i < n; ++i) {
    if (r < rs.size() && i == rs[r]) {
      i += rl[r] - 1;
      ++r;
      continue;
    }
    perm.push_back((uint32_t)i);
  }
  const size_t b = 4 * n;
  HIPCALL(e, ensure(hs.x, b), "hipMalloc(halos)");
  HIPCALL(e, ensure(hs.y, b), "hipMalloc(halos)");
  HIPCALL(e, ensure(hs.z, b), "hipMalloc(halos)");
  HIPCALL(e, ensure(hs.slot, b), "hipMalloc(halos)");
  // slot[i] = perm[i]: a halo's count row is its Morton position
  HIPCALL(e, hipMemcpy(hs.slot.p, perm.data(), b, hipMemcpyHostToDevice), "hipMemcpy(halos)");
  HIPCALL(e, permute_halos(hs.bx.as<float>(), hs.by.as<float>(), hs.bz.as<float>(), nullptr,
                           hs.x.as<float>(), hs.y.as<float>(), hs.z.as<float>(), nullptr,
                           hs.slot.as<uint32_t>(), n, e->stream),
          "halo tiling");
  HIPCALL(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
  hs.permuted = true;
  hs.n_tiled = ntiled;
  hs.run_start.swap(nstart);
  hs.run_len.swap(rl);
  return SJCS_OK;
}

// Halo centres from a device AoS (3 FP32 every `stride` bytes) -> Morton
// order (count rows), input rows, the read-out's inverse; with `tiles`, the
// working order for the engine's tile level.
int sort_halos(sjcs_engine* e, HaloState& hs, const float* d_xyz, size_t stride, uint64_t n,
               bool tiles, bool timed) {
  if (n > 0xffffffffull) return fail(e, SJCS_ERR_INVALID, "at most 2^32-1 halos");
  const size_t b = sizeof(uint32_t) * (n ? n : 1);
  DevBuf key[2], x[2], y[2], z[2], row[2], inv;
  int rc = SJCS_OK;
  hipError_t h = hipSuccess;
  for (int i = 0; i < 2 && h == hipSuccess; ++i) {
    h = ensure(key[i], b);
    if (h == hipSuccess) h = ensure(x[i], b);
    if (h == hipSuccess) h = ensure(y[i], b);
    if (h == hipSuccess) h = ensure(z[i], b);
    if (h == hipSuccess) h = ensure(row[i], b);
  }
  if (h == hipSuccess) h = ensure(inv, b);
  if (h == hipSuccess) rc = ensure_ws(e, radix_workspace_words(n));
  if (h == hipSuccess && rc == SJCS_OK && timed) h = hipEventRecord(e->hev[0], e->stream);
  if (h == hipSuccess && rc == SJCS_OK)
    h = halo_keys(d_xyz, stride, n, key[0].as<uint32_t>(), x[0].as<float>(), y[0].as<float>(),
                  z[0].as<float>(), row[0].as<uint32_t>(), e->stream);
  int sel = 0;
  if (h == hipSuccess && rc == SJCS_OK) {
    uint32_t* keys[2] = {key[0].as<uint32_t>(), key[1].as<uint32_t>()};
    uint32_t* pay[2][4] = {
        {x[0].as<uint32_t>(), y[0].as<uint32_t>(), z[0].as<uint32_t>(), row[0].as<uint32_t>()},
        {x[1].as<uint32_t>(), y[1].as<uint32_t>(), z[1].as<uint32_t>(), row[1].as<uint32_t>()}};
    h = radix_sort_pairs(keys, pay, 4, n, 30, e->ws.as<uint32_t>(), e->stream, &sel);
  }
  // the read-out gathers each input row from its count row
  if (h == hipSuccess && rc == SJCS_OK)
    h = invert_rows(row[sel].as<uint32_t>(), inv.as<uint32_t>(), n, e->stream);
  // the timed part ends with the device's sort (the host's tiling below is
  // not device time)
  if (h == hipSuccess && rc == SJCS_OK && timed) h = hipEventRecord(e->hev[1], e->stream);
  // the host looks for runs of one Morton key only when the device saw one
  uint32_t runs = 0;
  if (h == hipSuccess && rc == SJCS_OK && tiles && n >= kTileMin) {
    h = ensure(e->inbox, 16);
    if (h == hipSuccess)
      h = has_key_runs(key[sel].as<uint32_t>(), n, kTileMin, e->inbox.as<uint32_t>(), e->stream);
    if (h == hipSuccess)
      h = hipMemcpyAsync(&runs, e->inbox.p, 4, hipMemcpyDeviceToHost, e->stream);
  }
  if (h == hipSuccess) h = hipStreamSynchronize(e->stream);
  if (h != hipSuccess && rc == SJCS_OK) rc = hip_fail(e, h, "halo sort");
  if (rc == SJCS_OK) {
    hs.release();
    hs.n = n;
    hs.runs0 = runs != 0;
    hs.bx = x[sel];
    hs.by = y[sel];
    hs.bz = z[sel];
    hs.brow = row[sel];
    hs.bkey = key[sel];
    hs.inv = inv;
    x[sel] = y[sel] = z[sel] = row[sel] = key[sel] = inv = DevBuf();
  }
  if (rc == SJCS_OK && tiles) rc = apply_tiles(e, hs, (int)e->tile_level);
  if (timed) e->halo_timed = rc == SJCS_OK;
  // the sort's other buffers go after the timed part (a free waits for the
  // device)
  for (int i = 0; i < 2; ++i) {
    key[i].release();
    x[i].release();
    y[i].release();
    z[i].release();
    row[i].release();
  }
  inv.release();
  return rc;
}

// Packs strided host points into pinned memory and uploads them to `dst`.
int upload_points(sjcs_engine* e, DevBuf& dst, const float* pts, size_t stride, uint64_t n) {
  const size_t bytes = 12 * (size_t)n;
  HIPCALL(e, ensure(dst, bytes), "hipMalloc(points)");
  if (n == 0) return SJCS_OK;
  int rc = ensure_pinned(e, bytes);
  if (rc != SJCS_OK) return rc;
  float* p = static_cast<float*>(e->pinned);
  if (stride == 12) {
    memcpy(p, pts, bytes);
  } else {
    const char* src = reinterpret_cast<const char*>(pts);
    for (uint64_t i = 0; i < n; ++i) memcpy(p + 3 * i, src + i * stride, 12);
  }
  HIPCALL(e, hipMemcpyAsync(dst.p, p, bytes, hipMemcpyHostToDevice, e->stream), "hipMemcpy(points)");
  return SJCS_OK;
}

bool sort_by_axis(const sjcs_engine* e, const Grid& g) {
  return g.g <= kCsMaxCells && g.gx <= kCsMaxCellsX && e->index_sort != 1;
}

// Sizes ib and the workspace for n particles on grid g; *grew tells whether
// anything was reallocated (after all work of the engine was done).
int reserve_index(sjcs_engine* e, IndexBuffers& ib, uint64_t n, Grid g, bool* grew) {
  *grew = false;
  if (n > 0xffffffffull) return fail(e, SJCS_ERR_INVALID, "at most 2^32-1 particles per chunk");
  const size_t b = sizeof(uint32_t) * (n ? n : 1);
  // grids of at most kCsMaxCells cells per axis use the per-axis cell sort;
  // finer ones (or "index_sort" 1) the radix sort of packed cell keys
  const bool by_axis = sort_by_axis(e, g);
  const uint64_t wsw = by_axis ? cell_sort_workspace_words(n)
                               : std::max(radix_workspace_words(n), scan_workspace_words(g.ncells));
  const size_t ceb = sizeof(uint32_t) * (size_t)g.ncells;
  bool grow = e->ws.bytes < wsw * 4 || ib.cell_end.bytes < ceb;
  for (int i = 0; i < 2; ++i)
    grow = grow || (!by_axis && ib.key[i].bytes < b) || ib.x[i].bytes < b ||
           ib.y[i].bytes < b || ib.z[i].bytes < b;
  if (grow) {
    int rc = quiesce(e);
    if (rc != SJCS_OK) return rc;
    *grew = true;
    ++e->epoch;
  }
  for (int i = 0; i < 2; ++i) {
    if (!by_axis) HIPCALL(e, ensure(ib.key[i], b), "hipMalloc(index)");
    HIPCALL(e, ensure(ib.x[i], b), "hipMalloc(index)");
    HIPCALL(e, ensure(ib.y[i], b), "hipMalloc(index)");
    HIPCALL(e, ensure(ib.z[i], b), "hipMalloc(index)");
  }
  HIPCALL(e, ensure(ib.cell_end, ceb), "hipMalloc(cell table)");
  int rc = ensure_ws(e, wsw);
  if (rc != SJCS_OK) return rc;
  ib.grid = g;
  ib.n = n;
  return SJCS_OK;
}

// Enqueues the index build of the n particles of src into ib (sized by
// reserve_index) on stream st; launches only.  Compressed triplets need the
// per-axis sort (the packed-key sort reads decoded positions) and a file of
// one tile per side.
int enqueue_index(sjcs_engine* e, IndexBuffers& ib, const ParticleSource& src, uint64_t n, Grid g,
                  hipStream_t st) {
  const bool by_axis = sort_by_axis(e, g);
  if (src.uv && (!by_axis || !src.gd.one_tile))
    return fail(e, SJCS_ERR_STATE, "compressed particles reach the index build undecoded");
  if (by_axis) {
    float* xs[2] = {ib.x[0].as<float>(), ib.x[1].as<float>()};
    float* ys[2] = {ib.y[0].as<float>(), ib.y[1].as<float>()};
    float* zs[2] = {ib.z[0].as<float>(), ib.z[1].as<float>()};
    HIPCALL(e, cell_sort_particles(src, n, g, xs, ys, zs, ib.cell_end.as<uint32_t>(),
                                   e->ws.as<uint32_t>(), st),
            "cell sort");
    ib.sel = 0;
    return SJCS_OK;
  }
  HIPCALL(e, particle_keys(src.xyz, n, g, ib.key[0].as<uint32_t>(), ib.x[0].as<float>(),
                           ib.y[0].as<float>(), ib.z[0].as<float>(), st),
          "cell keys");
  uint32_t* keys[2] = {ib.key[0].as<uint32_t>(), ib.key[1].as<uint32_t>()};
  uint32_t* pay[2][4] = {
      {ib.x[0].as<uint32_t>(), ib.y[0].as<uint32_t>(), ib.z[0].as<uint32_t>(), nullptr},
      {ib.x[1].as<uint32_t>(), ib.y[1].as<uint32_t>(), ib.z[1].as<uint32_t>(), nullptr}};
  int sel = 0;
  HIPCALL(e, radix_sort_pairs(keys, pay, 3, n, key_bits_for(g.ncells), e->ws.as<uint32_t>(), st,
                              &sel),
          "radix sort");
  ib.sel = sel;
  HIPCALL(e, build_cell_end(ib.key[sel].as<uint32_t>(), n, ib.cell_end.as<uint32_t>(), g.ncells,
                            e->ws.as<uint32_t>(), st),
          "cell table");
  return SJCS_OK;
}

// Builds the cell index of n particles at d_xyz (device, packed) into ib,
// on stream st.  Buffers that must grow are reallocated only after all work
// of the engine is done.
int build_index(sjcs_engine* e, IndexBuffers& ib, const ParticleSource& src, uint64_t n, Grid g,
                hipStream_t st) {
  bool grew = false;
  int rc = reserve_index(e, ib, n, g, &grew);
  if (rc != SJCS_OK) return rc;
  return enqueue_index(e, ib, src, n, g, st);
}

// The halos in working order (see HaloState).
HaloView halo_view(const HaloState& h) {
  HaloView v;
  v.x = (h.permuted ? h.x : h.bx).as<float>();
  v.y = (h.permuted ? h.y : h.by).as<float>();
  v.z = (h.permuted ? h.z : h.bz).as<float>();
  v.slot = h.permuted ? h.slot.as<uint32_t>() : nullptr;
  v.n = h.n;
  return v;
}

int prepare_counts(sjcs_engine* e) {
  ++e->counts_gen;
  e->counts_fresh = true;
  e->c64_live = false;
  e->pending = 0;
  if (!e->have_shells || !e->have_halos) return SJCS_OK;
  const size_t bytes = sizeof(uint32_t) * (size_t)e->halos.n * (size_t)e->shells.nrads;
  if (e->counts.bytes < bytes) {
    int rc = quiesce(e);
    if (rc != SJCS_OK) return rc;
    ++e->epoch;
  }
  HIPCALL(e, ensure(e->counts, bytes), "hipMalloc(counts)");
  return SJCS_OK;
}

// Lanes per halo of the count kernel: "kernel" 1 forces the plain one
// halo per wavefront kernel, else "lanes_per_halo" (0 = automatic).
// Automatic: 64 lanes when a box holds thousands of particles (the row walk
// then dominates), else 16.
int lanes(const sjcs_engine* e, uint64_t n, float rbox) {
  if (e->kernel == 1) return 64;
  if (e->lanes_per_halo) return (int)e->lanes_per_halo;
  const double side = 2.0 * (double)rbox / 1000.0;
  const double per_box = (double)n * side * side * side;
  return !(per_box <= 4096.0) ? 64 : 16;
}
int seg_len(const sjcs_engine* e) { return e->segment ? (int)e->segment : 4; }

// Most images of one halo a particle can lie in: 1 unless the query box is
// about as wide as the periodic box (or the radius is not finite).
uint64_t images_max(const sjcs_engine* e) {
  const float rb = e->shells.view.rbox;
  return rb < 499.0f ? 1u : 8u;
}

// Folds the u32 running counts into the u64 totals (stream).
int fold_counts(sjcs_engine* e) {
  const uint64_t total = e->halos.n * (uint64_t)e->shells.nrads;
  if (e->counts64.bytes < total * sizeof(uint64_t)) {
    int rc = quiesce(e);
    if (rc != SJCS_OK) return rc;
    HIPCALL(e, ensure(e->counts64, total * sizeof(uint64_t)), "hipMalloc(counts)");
  }
  HIPCALL(e, widen_counts(e->counts.as<uint32_t>(), e->counts64.as<uint64_t>(), total,
                          !e->c64_live, e->stream),
          "fold counts");
  e->c64_live = true;
  e->counts_fresh = true;
  e->pending = 0;
  return SJCS_OK;
}

// Tiles of the current halos for the current shells (host runs cut into
// pieces of at most tile_cap halos, uploaded once); false when the count
// runs without tiles.
int ensure_tiles(sjcs_engine* e, HaloState& hs, const Shells& sh, bool* use) {
  *use = false;
  int cap = tile_cap(sh.nrads);
  if (e->tile_halos > 0 && e->tile_halos < cap) cap = (int)e->tile_halos;
  if (!e->tiles || !sh.sorted || hs.n_tiled == 0 || cap < 2) return SJCS_OK;
  *use = true;
  if (hs.tcap == cap) return SJCS_OK;
  std::vector<uint32_t> ts, tl;
  for (size_t r = 0; r < hs.run_start.size(); ++r) {
    const uint32_t len = hs.run_len[r];
    const uint32_t parts = (len + cap - 1) / cap;
    for (uint32_t p = 0, at = 0; p < parts; ++p) {
      const uint32_t m = len / parts + (p < len % parts ? 1 : 0);
      ts.push_back(hs.run_start[r] + at);
      tl.push_back(m);
      at += m;
    }
  }
  int rc = quiesce(e);  // the tile table is read by counts in flight
  if (rc != SJCS_OK) return rc;
  ++e->epoch;
  HIPCALL(e, ensure(hs.tstart, 4 * ts.size()), "hipMalloc(tiles)");
  HIPCALL(e, ensure(hs.tlen, 4 * tl.size()), "hipMalloc(tiles)");
  HIPCALL(e, ensure(hs.twork, 4 * tile_work_words((uint32_t)ts.size())), "hipMalloc(tiles)");
  HIPCALL(e, hipMemcpy(hs.tstart.p, ts.data(), 4 * ts.size(), hipMemcpyHostToDevice), "hipMemcpy");
  HIPCALL(e, hipMemcpy(hs.tlen.p, tl.data(), 4 * tl.size(), hipMemcpyHostToDevice), "hipMemcpy");
  hs.ntiles = (uint32_t)ts.size();
  hs.tmax = tl.empty() ? 0u : *std::max_element(tl.begin(), tl.end());
  hs.tcap = cap;
  return SJCS_OK;
}

// Launches the count of the current halos against index `ix` on stream st:
// tiles (when `use_tiles`) and then the remaining halos one by one.
hipError_t count_chunk(sjcs_engine* e, const IndexView& ix, const HaloState& hs, const Shells& sh,
                       bool use_tiles, uint32_t* counts, bool accumulate,
                       unsigned long long* inbox, hipStream_t st) {
  HaloView hv = halo_view(hs);
  const int lanes_n = lanes(e, ix.n, sh.rbox);
  const uint32_t scap = (uint32_t)e->stage_particles, shal = (uint32_t)e->stage_halos;
  if (!use_tiles)
    return count_halos(ix, hv, sh, counts, accumulate, inbox, lanes_n, seg_len(e), st, scap, shal);
  TileView tv;
  tv.start = hs.tstart.as<uint32_t>();
  tv.len = hs.tlen.as<uint32_t>();
  tv.n = hs.ntiles;
  tv.max_len = hs.tmax;
  hipError_t h = count_tiles(ix, hv, sh, tv, counts, accumulate, inbox, hs.twork.as<uint32_t>(), st);
  if (h != hipSuccess) return h;
  // the other halos, after the tiled runs; slot[] names their count rows
  const uint64_t nt = hs.n_tiled;
  hv.x += nt;
  hv.y += nt;
  hv.z += nt;
  hv.slot += nt;
  hv.n -= nt;
  return count_halos(ix, hv, sh, counts, accumulate, inbox, lanes_n, seg_len(e), st, scap, shal);
}

int piece_done(sjcs_engine* e, uint64_t n);

// Records stage_read[s] on stream st when p points into slot s's device
// staging (cin or cuv): work queued on st so far is the last reader of it.
int note_staging_read(sjcs_engine* e, const void* p, hipStream_t st) {
  const char* c = static_cast<const char*>(p);
  for (int s = 0; s < 2; ++s) {
    const DevBuf* bufs[2] = {&e->cin[s], &e->cuv[s]};
    for (const DevBuf* b : bufs) {
      const char* lo = b->as<char>();
      if (lo && c >= lo && c < lo + b->bytes) {
        HIPCALL(e, hipEventRecord(e->stage_read[s], st), "hipEventRecord");
        return SJCS_OK;
      }
    }
  }
  return SJCS_OK;
}

// Graph mode: sort and count of a piece on `stream` as one captured graph,
// captured on first use and replayed while nothing it holds has changed.
// The sort waits for the slot's upload; it follows the previous count in
// stream order.
int graph_piece(sjcs_engine* e, const ParticleSource& src, uint64_t n, const Grid& g) {
  const int sl = e->slot;
  bool grew = false, use_tiles = false;
  int rc = reserve_index(e, e->ibs[sl], n, g, &grew);
  if (rc == SJCS_OK) rc = ensure_tiles(e, e->halos, e->shells.view, &use_tiles);
  if (rc != SJCS_OK) return rc;
  HIPCALL(e, hipStreamWaitEvent(e->stream, e->uploaded[sl], 0), "hipStreamWaitEvent");
  HIPCALL(e, hipEventRecord(e->ev[2], e->stream), "hipEventRecord");
  HIPCALL(e, hipEventRecord(e->ev[3], e->stream), "hipEventRecord");
  HIPCALL(e, hipEventRecord(e->ev[4], e->stream), "hipEventRecord");
  sjcs_engine::ChunkGraph& cg = e->graphs[sl][e->counts_fresh ? 0 : 1];
  const void* in = src.uv ? static_cast<const void*>(src.uv) : static_cast<const void*>(src.xyz);
  if (!cg.exec || cg.in != in || cg.n != n || cg.epoch != e->epoch ||
      (src.uv && memcmp(&cg.gd, &src.gd, sizeof(GadgetDecode)) != 0)) {
    if (cg.exec) (void)hipGraphExecDestroy(cg.exec);
    cg.exec = nullptr;
    hipGraph_t graph = nullptr;
    HIPCALL(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal),
            "hipStreamBeginCapture");
    int rq = enqueue_index(e, e->ibs[sl], src, n, g, e->stream);
    hipError_t hc = hipSuccess;
    if (rq == SJCS_OK)
      hc = count_chunk(e, e->ibs[sl].view(), e->halos, e->shells.view, use_tiles,
                       e->counts.as<uint32_t>(), !e->counts_fresh,
                       nullptr, e->stream);
    const hipError_t he = hipStreamEndCapture(e->stream, &graph);
    if (rq != SJCS_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rq;
    }
    if (hc != hipSuccess || he != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return hip_fail(e, hc != hipSuccess ? hc : he, "graph capture");
    }
    const hipError_t hi = hipGraphInstantiate(&cg.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (hi != hipSuccess) {
      cg.exec = nullptr;
      return hip_fail(e, hi, "hipGraphInstantiate");
    }
    cg.in = in;
    cg.gd = src.uv ? src.gd : GadgetDecode{};
    cg.n = n;
    cg.epoch = e->epoch;
  }
  HIPCALL(e, hipGraphLaunch(cg.exec, e->stream), "hipGraphLaunch");
  rc = note_staging_read(e, in, e->stream);
  if (rc != SJCS_OK) return rc;
  HIPCALL(e, hipEventRecord(e->ev[5], e->stream), "hipEventRecord");
  HIPCALL(e, hipEventRecord(e->built[sl], e->stream), "hipEventRecord");
  HIPCALL(e, hipEventRecord(e->counted[sl], e->stream), "hipEventRecord");
  return piece_done(e, n);
}

// Sorts one piece into the current slot (bstream) and counts it (stream).
int add_piece(sjcs_engine* e, ParticleSource src, uint64_t n) {
  const int sl = e->slot;
  const uint64_t img = images_max(e);
  const uint64_t limit = e->fold_limit > 0 ? (uint64_t)e->fold_limit : 0xffffffffull;
  if (!e->counts_fresh && e->pending + img * n > limit) {
    int rc = fold_counts(e);
    if (rc != SJCS_OK) return rc;
  }
  const Grid g = choose_grid(n, e->shells.view.rbox, e->cell_size_milli, e->x_split);
  if (src.uv && (!sort_by_axis(e, g) || !src.gd.one_tile)) {
    // the packed-key sort, and the cell sort for files of several tiles per
    // side, need the positions first: decode into the slot's staging once
    // the sort that last read it is done
    const int sl = e->slot;
    if (e->cin[sl].bytes < 12 * (size_t)n) {
      HIPCALL(e, hipEventSynchronize(e->counted[sl]), "hipEventSynchronize");
      HIPCALL(e, hipEventSynchronize(e->stage_read[sl]), "hipEventSynchronize");
    }
    HIPCALL(e, ensure(e->cin[sl], 12 * (size_t)n), "hipMalloc(points)");
    HIPCALL(e, hipStreamWaitEvent(e->bstream, e->counted[sl], 0), "hipStreamWaitEvent");
    HIPCALL(e, hipStreamWaitEvent(e->bstream, e->stage_read[sl], 0), "hipStreamWaitEvent");
    HIPCALL(e, decode_gadget(src.uv, n, src.gd, e->cin[sl].as<float>(), e->bstream),
            "decode compressed particles");
    int rn = note_staging_read(e, src.uv, e->bstream);
    if (rn != SJCS_OK) return rn;
    if (e->graph) HIPCALL(e, hipEventRecord(e->uploaded[sl], e->bstream), "hipEventRecord");
    src = ParticleSource{};
    src.xyz = e->cin[sl].as<float>();
  }
  if (e->graph) return graph_piece(e, src, n, g);
  bool use_tiles = false;
  int rt = ensure_tiles(e, e->halos, e->shells.view, &use_tiles);
  if (rt != SJCS_OK) return rt;
  HIPCALL(e, hipStreamWaitEvent(e->bstream, e->counted[sl], 0), "hipStreamWaitEvent");
  if (!e->overlap && e->last >= 0)
    HIPCALL(e, hipStreamWaitEvent(e->bstream, e->counted[e->last], 0), "hipStreamWaitEvent");
  HIPCALL(e, hipEventRecord(e->ev[2], e->bstream), "hipEventRecord");
  int rc = build_index(e, e->ibs[sl], src, n, g, e->bstream);
  if (rc == SJCS_OK)
    rc = note_staging_read(e, src.uv ? static_cast<const void*>(src.uv)
                                     : static_cast<const void*>(src.xyz),
                           e->bstream);
  if (rc != SJCS_OK) return rc;
  HIPCALL(e, hipEventRecord(e->ev[3], e->bstream), "hipEventRecord");
  HIPCALL(e, hipEventRecord(e->built[sl], e->bstream), "hipEventRecord");
  HIPCALL(e, hipStreamWaitEvent(e->stream, e->built[sl], 0), "hipStreamWaitEvent");
  HIPCALL(e, hipEventRecord(e->ev[4], e->stream), "hipEventRecord");
  HIPCALL(e, count_chunk(e, e->ibs[sl].view(), e->halos, e->shells.view, use_tiles,
                         e->counts.as<uint32_t>(),
                         !e->counts_fresh, nullptr, e->stream),
          "count kernel");
  HIPCALL(e, hipEventRecord(e->ev[5], e->stream), "hipEventRecord");
  HIPCALL(e, hipEventRecord(e->counted[sl], e->stream), "hipEventRecord");
  return piece_done(e, n);
}

// Bookkeeping after a piece was enqueued in slot e->slot.
int piece_done(sjcs_engine* e, uint64_t n) {
  const int sl = e->slot;
  e->last = sl;
  e->slot = sl ^ 1;
  e->counts_fresh = false;
  ++e->counts_gen;
  e->pending += images_max(e) * n;
  e->timed = true;
  return SJCS_OK;
}

// A chunk, in pieces small enough for 32-bit particle offsets and for
// per-piece counts below 2^32.
int add_chunk(sjcs_engine* e, const ParticleSource& src, uint64_t n) {
  uint64_t cap = std::min<uint64_t>(0x3fffffffull, 0xffffffffull / images_max(e));
  if (e->piece_max > 0) cap = std::min<uint64_t>(cap, (uint64_t)e->piece_max);
  uint64_t done = 0;
  do {
    const uint64_t m = std::min(cap, n - done);
    ParticleSource piece = src;
    if (src.uv) {
      piece.uv = src.uv + 3 * done;
      piece.gd.id_start = src.gd.id_start + (int64_t)done;
    } else {
      piece.xyz = src.xyz + 3 * done;
    }
    int rc = add_piece(e, piece, m);
    if (rc != SJCS_OK) return rc;
    done += m;
  } while (done < n);
  return SJCS_OK;
}

int check_ready(sjcs_engine* e) {
  if (!e) return SJCS_ERR_INVALID;
  if (!e->have_shells) return fail(e, SJCS_ERR_STATE, "no shells set (sjcs_set_shells)");
  if (!e->have_halos) return fail(e, SJCS_ERR_STATE, "no halos set (sjcs_set_halos)");
  return use_device(e);
}

// Stages `bytes` of host data of the next chunk in its slot's pinned buffer
// and queues their upload into the slot's device buffer dst on bstream, after
// the slot's last sort (graph mode: its last graph, on the count stream) has
// read the slot's device staging.  Records ev[0] before the upload.
int stage_chunk(sjcs_engine* e, const void* src, size_t bytes, DevBuf& dst) {
  const int sl = e->slot;
  const size_t pin = bytes ? bytes : 16;
  // the slot's pinned buffer is free once its last upload is done; its
  // device staging once the sort that read it is done
  HIPCALL(e, hipEventSynchronize(e->uploaded[sl]), "hipEventSynchronize");
  if (e->cpin_bytes[sl] < pin) {
    if (e->cpin[sl]) (void)hipHostFree(e->cpin[sl]);
    e->cpin[sl] = nullptr;
    e->cpin_bytes[sl] = 0;
    HIPCALL(e, hipHostMalloc(&e->cpin[sl], pin, hipHostMallocDefault), "hipHostMalloc");
    e->cpin_bytes[sl] = pin;
  }
  if (dst.bytes < bytes) {
    HIPCALL(e, hipEventSynchronize(e->counted[sl]), "hipEventSynchronize");
    HIPCALL(e, hipEventSynchronize(e->stage_read[sl]), "hipEventSynchronize");
  }
  HIPCALL(e, ensure(dst, bytes), "hipMalloc(points)");
  if (bytes) memcpy(e->cpin[sl], src, bytes);
  HIPCALL(e, hipStreamWaitEvent(e->bstream, e->counted[sl], 0), "hipStreamWaitEvent");
  HIPCALL(e, hipStreamWaitEvent(e->bstream, e->stage_read[sl], 0), "hipStreamWaitEvent");
  HIPCALL(e, hipEventRecord(e->ev[0], e->bstream), "hipEventRecord");
  if (bytes)
    HIPCALL(e, hipMemcpyAsync(dst.p, e->cpin[sl], bytes, hipMemcpyHostToDevice, e->bstream),
            "hipMemcpy(points)");
  return SJCS_OK;
}

}  // namespace

extern "C" {

int sjcs_engine_create(int device, sjcs_engine** out) {
  if (!out) return SJCS_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  hipError_t h = hipGetDeviceCount(&ndev);
  if (h != hipSuccess || ndev == 0) {
    g_create_error = std::string("no HIP device: ") + hipGetErrorString(h);
    return SJCS_ERR_DEVICE;
  }
  if (device < 0 || device >= ndev) {
    g_create_error = "device index out of range";
    return SJCS_ERR_INVALID;
  }
  sjcs_engine* e = new sjcs_engine();
  e->device = device;
  h = hipSetDevice(device);
  if (h == hipSuccess) h = scratch_address(device, &e->scratch);
  if (h == hipSuccess) h = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (h == hipSuccess) h = hipStreamCreateWithFlags(&e->bstream, hipStreamNonBlocking);
  for (int i = 0; i < 6 && h == hipSuccess; ++i) h = hipEventCreate(&e->ev[i]);
  for (int i = 0; i < 2 && h == hipSuccess; ++i) {
    h = hipEventCreate(&e->hev[i]);
    if (h == hipSuccess) h = hipEventCreate(&e->rev[i]);
  }
  for (int i = 0; i < 2 && h == hipSuccess; ++i) {
    h = hipEventCreateWithFlags(&e->built[i], hipEventDisableTiming);
    if (h == hipSuccess) h = hipEventCreateWithFlags(&e->counted[i], hipEventDisableTiming);
    if (h == hipSuccess) h = hipEventCreateWithFlags(&e->uploaded[i], hipEventDisableTiming);
    if (h == hipSuccess) h = hipEventCreateWithFlags(&e->stage_read[i], hipEventDisableTiming);
    if (h == hipSuccess) h = hipEventCreateWithFlags(&e->dl[i], hipEventDisableTiming);
    // recorded once so that the first waits on them return at once
    if (h == hipSuccess) h = hipEventRecord(e->built[i], e->stream);
    if (h == hipSuccess) h = hipEventRecord(e->counted[i], e->stream);
    if (h == hipSuccess) h = hipEventRecord(e->uploaded[i], e->stream);
    if (h == hipSuccess) h = hipEventRecord(e->stage_read[i], e->stream);
  }
  // the pair-test total (u64) and the halo sort's run flag (u32); clearing
  // it is the process's first kernel of this library, which loads the
  // library's code onto the device here rather than in the first halo sort
  if (h == hipSuccess) h = ensure(e->inbox, 16);
  if (h == hipSuccess) h = fill_u32(e->inbox.as<uint32_t>(), 0u, 4, e->stream);
  if (h == hipSuccess) h = hipStreamSynchronize(e->stream);
  if (h != hipSuccess) {
    g_create_error = std::string("engine setup: ") + hipGetErrorString(h);
    sjcs_engine_destroy(e);
    return SJCS_ERR_DEVICE;
  }
  *out = e;
  return SJCS_OK;
}

void sjcs_engine_destroy(sjcs_engine* e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->bstream) (void)hipStreamSynchronize(e->bstream);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  e->shells.d_r2.release();
  e->shells.d_lut.release();
  e->halos.release();
  e->counts.release();
  e->counts64.release();
  e->ucounts.release();
  for (int i = 0; i < 2; ++i) {
    for (int j = 0; j < 2; ++j)
      if (e->graphs[i][j].exec) (void)hipGraphExecDestroy(e->graphs[i][j].exec);
    e->ibs[i].release();
    e->cin[i].release();
    e->cuv[i].release();
    if (e->cpin[i]) (void)hipHostFree(e->cpin[i]);
    if (e->built[i]) (void)hipEventDestroy(e->built[i]);
    if (e->counted[i]) (void)hipEventDestroy(e->counted[i]);
    if (e->uploaded[i]) (void)hipEventDestroy(e->uploaded[i]);
    if (e->stage_read[i]) (void)hipEventDestroy(e->stage_read[i]);
    if (e->dl[i]) (void)hipEventDestroy(e->dl[i]);
  }
  e->in.release();
  e->ws.release();
  e->inbox.release();
  if (e->pinned) (void)hipHostFree(e->pinned);
  for (int i = 0; i < 6; ++i)
    if (e->ev[i]) (void)hipEventDestroy(e->ev[i]);
  for (int i = 0; i < 2; ++i) {
    if (e->hev[i]) (void)hipEventDestroy(e->hev[i]);
    if (e->rev[i]) (void)hipEventDestroy(e->rev[i]);
  }
  if (e->bstream) (void)hipStreamDestroy(e->bstream);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

const char* sjcs_last_error(const sjcs_engine* e) {
  return e ? e->err.c_str() : g_create_error.c_str();
}

int sjcs_set_shells(sjcs_engine* e, const float* radiuses_sq, int nrads) {
  if (!e) return SJCS_ERR_INVALID;
  int rc = use_device(e);
  if (rc == SJCS_OK) rc = quiesce(e);  // in-flight counts read the shells
  if (rc != SJCS_OK) return rc;
  ++e->epoch;
  rc = make_shells(e, e->shells, radiuses_sq, nrads);
  if (rc != SJCS_OK) return rc;
  e->have_shells = true;
  return prepare_counts(e);
}

int sjcs_set_halos(sjcs_engine* e, const float* points, size_t stride_bytes, uint64_t nhalos) {
  if (!e) return SJCS_ERR_INVALID;
  if ((!points && nhalos) || stride_bytes < 12)
    return fail(e, SJCS_ERR_INVALID, "halo points: null pointer or stride < 12");
  int rc = use_device(e);
  if (rc == SJCS_OK) rc = quiesce(e);  // in-flight counts read the halos
  if (rc != SJCS_OK) return rc;
  ++e->epoch;
  // the records go up as they are (no host-side gather or page-locked
  // staging); the key kernel reads the centres at their stride
  // (the last record is read up to its centre only)
  const size_t bytes = nhalos ? stride_bytes * (size_t)(nhalos - 1) + 12 : 0;
  HIPCALL(e, ensure(e->in, bytes), "hipMalloc(halos)");
  if (nhalos)
    HIPCALL(e, hipMemcpy(e->in.p, points, bytes, hipMemcpyHostToDevice), "hipMemcpy(halos)");
  rc = sort_halos(e, e->halos, e->in.as<float>(), stride_bytes, nhalos, true, true);
  if (rc != SJCS_OK) return rc;
  e->have_halos = true;
  return prepare_counts(e);
}

int sjcs_set_halos_device(sjcs_engine* e, const float* d_xyz, uint64_t nhalos) {
  if (!e) return SJCS_ERR_INVALID;
  if (!d_xyz && nhalos) return fail(e, SJCS_ERR_INVALID, "halo points: null pointer");
  int rc = use_device(e);
  if (rc == SJCS_OK) rc = quiesce(e);  // in-flight counts read the halos
  if (rc != SJCS_OK) return rc;
  ++e->epoch;
  rc = sort_halos(e, e->halos, d_xyz, 12, nhalos, true, true);
  if (rc != SJCS_OK) return rc;
  e->have_halos = true;
  return prepare_counts(e);
}

int sjcs_reset_counts(sjcs_engine* e) {
  if (!e) return SJCS_ERR_INVALID;
  return prepare_counts(e);
}

int sjcs_add_particles(sjcs_engine* e, const float* xyz, uint64_t n) {
  int rc = check_ready(e);
  if (rc != SJCS_OK) return rc;
  if (!xyz && n) return fail(e, SJCS_ERR_INVALID, "particles: null pointer");
  const int sl = e->slot;
  const size_t bytes = 12 * (size_t)n;
  rc = stage_chunk(e, xyz, bytes, e->cin[sl]);
  if (rc != SJCS_OK) return rc;
  HIPCALL(e, hipEventRecord(e->ev[1], e->bstream), "hipEventRecord");
  HIPCALL(e, hipEventRecord(e->uploaded[sl], e->bstream), "hipEventRecord");
  return add_chunk(e, packed(e->cin[sl].as<float>()), n);
}

int sjcs_add_compressed_particles(sjcs_engine* e, const uint16_t* uv, uint64_t n,
                                  const sjcs_gadget_decode* p) {
  int rc = check_ready(e);
  if (rc != SJCS_OK) return rc;
  if ((!uv && n) || !p) return fail(e, SJCS_ERR_INVALID, "compressed particles: null pointer");
  GadgetDecode d;
  std::string err;
  if (!gadget_prepare(*p, n, &d, &err)) return fail(e, SJCS_ERR_INVALID, err);
  const int sl = e->slot;
  // the triplets go to the device as they are; the sort's first pass decodes
  // them as it reads them (the slot's staging is rewritten only after the
  // count that followed its last sort)
  rc = stage_chunk(e, uv, 6 * (size_t)n, e->cuv[sl]);
  if (rc != SJCS_OK) return rc;
  HIPCALL(e, hipEventRecord(e->ev[1], e->bstream), "hipEventRecord");
  HIPCALL(e, hipEventRecord(e->uploaded[sl], e->bstream), "hipEventRecord");
  ParticleSource src;
  src.uv = e->cuv[sl].as<uint16_t>();
  src.gd = d;
  return add_chunk(e, src, n);
}

int sjcs_add_compressed_particles_device(sjcs_engine* e, const uint16_t* d_uv, uint64_t n,
                                         const sjcs_gadget_decode* p) {
  int rc = check_ready(e);
  if (rc != SJCS_OK) return rc;
  if ((!d_uv && n) || !p) return fail(e, SJCS_ERR_INVALID, "compressed particles: null pointer");
  GadgetDecode d;
  std::string err;
  if (!gadget_prepare(*p, n, &d, &err)) return fail(e, SJCS_ERR_INVALID, err);
  // decoded inside the sort's first pass: no separate decode phase
  HIPCALL(e, hipEventRecord(e->ev[0], e->bstream), "hipEventRecord");
  HIPCALL(e, hipEventRecord(e->ev[1], e->bstream), "hipEventRecord");
  ParticleSource src;
  src.uv = d_uv;
  src.gd = d;
  return add_chunk(e, src, n);
}

int sjcs_decode_compressed_device(sjcs_engine* e, const uint16_t* d_uv, uint64_t n,
                                  const sjcs_gadget_decode* p, float* d_xyz) {
  if (!e) return SJCS_ERR_INVALID;
  if ((!d_uv || !d_xyz) && n) return fail(e, SJCS_ERR_INVALID, "decode: null pointer");
  if (!p) return fail(e, SJCS_ERR_INVALID, "decode: null parameters");
  int rc = use_device(e);
  if (rc != SJCS_OK) return rc;
  GadgetDecode d;
  std::string err;
  if (!gadget_prepare(*p, n, &d, &err)) return fail(e, SJCS_ERR_INVALID, err);
  HIPCALL(e, decode_gadget(d_uv, n, d, d_xyz, e->stream), "decode compressed particles");
  HIPCALL(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
  return SJCS_OK;
}

int sjcs_add_particles_device(sjcs_engine* e, const float* d_xyz, uint64_t n) {
  int rc = check_ready(e);
  if (rc != SJCS_OK) return rc;
  if (!d_xyz && n) return fail(e, SJCS_ERR_INVALID, "particles: null pointer");
  HIPCALL(e, hipEventRecord(e->ev[0], e->bstream), "hipEventRecord");
  HIPCALL(e, hipEventRecord(e->ev[1], e->bstream), "hipEventRecord");
  return add_chunk(e, packed(d_xyz), n);
}

int sjcs_sync(sjcs_engine* e) {
  if (!e) return SJCS_ERR_INVALID;
  int rc = use_device(e);
  if (rc != SJCS_OK) return rc;
  return quiesce(e);
}

int sjcs_get_counts(sjcs_engine* e, uint64_t* counts) {
  int rc = check_ready(e);
  if (rc != SJCS_OK) return rc;
  return sjcs_get_counts_range(e, 0, e->halos.n, counts);
}

// Rows [first, first + nhalos) of the totals in input order, u64 or (while
// the totals are the u32 running counts alone) u32.
int get_counts_rows(sjcs_engine* e, uint64_t first, uint64_t nhalos, void* counts, bool u32) {
  int rc = check_ready(e);
  if (rc != SJCS_OK) return rc;
  if (first > e->halos.n || nhalos > e->halos.n - first)
    return fail(e, SJCS_ERR_INVALID, "counts: halo range out of bounds");
  const size_t K = (size_t)e->shells.nrads;
  const size_t n = (size_t)nhalos * K;
  const size_t w = u32 ? sizeof(uint32_t) : sizeof(uint64_t);
  if (!counts && n) return fail(e, SJCS_ERR_INVALID, "counts: null pointer");
  rc = quiesce(e);
  if (rc != SJCS_OK) return rc;
  if (u32 && e->c64_live)
    return fail(e, SJCS_ERR_STATE,
                "counts: the totals are held in 64 bits; use sjcs_get_counts_range");
  if (n == 0) return SJCS_OK;
  if (e->counts_fresh && !e->c64_live) {
    memset(counts, 0, n * w);
    return SJCS_OK;
  }
  // every row back in input order once per state of the totals (and width);
  // ranges are then plain copies
  if (e->readout_gen != e->counts_gen || e->readout_u32 != u32) {
    const size_t all = (size_t)e->halos.n * K;
    HIPCALL(e, ensure(e->ucounts, all * w), "hipMalloc(counts)");
    HIPCALL(e, hipEventRecord(e->rev[0], e->stream), "hipEventRecord");
    if (u32)
      HIPCALL(e, readout_counts32(e->counts.as<uint32_t>(), e->halos.inv.as<uint32_t>(),
                                  e->ucounts.as<uint32_t>(), e->halos.n, e->shells.nrads,
                                  e->stream),
              "read out counts");
    else
      HIPCALL(e, readout_counts(e->counts_fresh ? nullptr : e->counts.as<uint32_t>(),
                                e->c64_live ? e->counts64.as<uint64_t>() : nullptr,
                                e->halos.inv.as<uint32_t>(), e->ucounts.as<uint64_t>(),
                                e->halos.n, e->shells.nrads, e->stream),
              "read out counts");
    HIPCALL(e, hipEventRecord(e->rev[1], e->stream), "hipEventRecord");
    e->readout_timed = true;
    e->readout_gen = e->counts_gen;
    e->readout_u32 = u32;
  }
  return download(e, counts, e->ucounts.as<char>() + first * K * w, n * w);
}

int sjcs_get_counts_range(sjcs_engine* e, uint64_t first, uint64_t nhalos, uint64_t* counts) {
  return get_counts_rows(e, first, nhalos, counts, false);
}

int sjcs_get_counts_range_u32(sjcs_engine* e, uint64_t first, uint64_t nhalos,
                              uint32_t* counts) {
  return get_counts_rows(e, first, nhalos, counts, true);
}

int sjcs_last_chunk_pair_tests(sjcs_engine* e, uint64_t* pair_tests) {
  int rc = check_ready(e);
  if (rc != SJCS_OK) return rc;
  if (!pair_tests) return fail(e, SJCS_ERR_INVALID, "null output");
  *pair_tests = 0;
  if (!e->timed) return SJCS_OK;
  HIPCALL(e, fill_u64(e->inbox.as<uint64_t>(), 0ull, 1, e->stream), "fill");
  bool use_tiles = false;
  rc = ensure_tiles(e, e->halos, e->shells.view, &use_tiles);
  if (rc != SJCS_OK) return rc;
  HIPCALL(e, count_chunk(e, e->ibs[e->last].view(), e->halos, e->shells.view, use_tiles,
                         nullptr, true,
                         e->inbox.as<unsigned long long>(), e->stream),
          "pair-test count");
  HIPCALL(e, hipMemcpyAsync(pair_tests, e->inbox.p, 8, hipMemcpyDeviceToHost, e->stream),
          "hipMemcpy");
  HIPCALL(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
  return SJCS_OK;
}

int sjcs_get_option(sjcs_engine* e, const char* key, int64_t* value) {
  if (!e || !key || !value) return SJCS_ERR_INVALID;
  const struct {
    const char* k;
    int64_t v;
  } opts[] = {{"cell_size_milli", e->cell_size_milli}, {"x_split", e->x_split},
              {"kernel", e->kernel},                   {"lanes_per_halo", e->lanes_per_halo},
              {"segment", e->segment},                 {"tiles", e->tiles},
              {"graph", e->graph},                     {"fold_limit", e->fold_limit},
              {"overlap", e->overlap},                 {"index_sort", e->index_sort},
              {"piece_max", e->piece_max},             {"tile_level", e->tile_level},
              {"tile_halos", e->tile_halos},           {"autotune_cache", e->autotune_cache},
              {"stage_particles", e->stage_particles}, {"stage_halos", e->stage_halos},
              {"autotune_cache_hit", e->autotune_hit ? 1 : 0}};
  for (const auto& o : opts)
    if (strcmp(key, o.k) == 0) {
      *value = o.v;
      return SJCS_OK;
    }
  return fail(e, SJCS_ERR_INVALID, std::string("unknown option ") + key);
}

}  // extern "C"

namespace {

struct TuneSetting {
  int64_t cell, xs, lanes, seg, tiles, level, thalos, stage, shalos;
};

// Autotuned settings of the workloads seen by this process (all engines).
std::mutex g_tune_mu;
std::map<std::string, std::pair<TuneSetting, float>> g_tune_cache;

// The workload a tuned setting is kept for: particle and halo counts, the
// thresholds, and how particles and halos fill the box (kOccCells^3 cell
// occupancy histograms, each bin's share of the cells quantized to 1/255).
// ws: occupancy_workspace_words() + 2 * kOccBins words of device memory.
int workload_key(sjcs_engine* e, const float* d_xyz, uint64_t n, uint32_t* ws, std::string* key) {
  uint32_t* hist = ws + occupancy_workspace_words();
  const HaloState& hs = e->halos;
  HIPCALL(e, occupancy_histogram(d_xyz, d_xyz + 1, d_xyz + 2, 3, n, ws, hist, e->stream),
          "occupancy");
  HIPCALL(e, occupancy_histogram(hs.bx.as<float>(), hs.by.as<float>(), hs.bz.as<float>(), 1, hs.n,
                                 ws, hist + kOccBins, e->stream),
          "occupancy");
  uint32_t h[2 * kOccBins];
  HIPCALL(e, hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, e->stream), "hipMemcpy");
  HIPCALL(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
  char buf[96];
  snprintf(buf, sizeof(buf), "np=%llu nh=%llu K=%d runs=%d r2=", (unsigned long long)n,
           (unsigned long long)hs.n, e->shells.nrads, hs.runs0 ? 1 : 0);
  std::string k(buf);
  for (float v : e->shells.r2) {
    snprintf(buf, sizeof(buf), "%08x", float_bits(v));
    k += buf;
  }
  k += " occ=";
  const double cells = (double)occupancy_workspace_words();
  for (int i = 0; i < 2 * kOccBins; ++i) {
    snprintf(buf, sizeof(buf), "%02x", (unsigned)lround(255.0 * (double)h[i] / cells));
    k += buf;
  }
  *key = k;
  return SJCS_OK;
}

void apply_setting(sjcs_engine* e, const TuneSetting& st) {
  e->cell_size_milli = st.cell;
  e->x_split = st.xs;
  e->lanes_per_halo = st.lanes;
  e->segment = st.seg;
  e->tiles = st.tiles;
  e->tile_level = st.level;
  e->tile_halos = st.thalos;
  e->stage_particles = st.stage;
  e->stage_halos = st.shalos;
  e->kernel = 0;
  e->halos.tcap = 0;  // tiles cut again for this setting
  ++e->epoch;
}

}  // namespace

extern "C" {

int sjcs_autotune(sjcs_engine* e, const float* d_xyz, uint64_t n, float* best_ms) {
  int rc = check_ready(e);
  if (rc != SJCS_OK) return rc;
  if (!d_xyz && n) return fail(e, SJCS_ERR_INVALID, "particles: null pointer");
  if (n > 0x3fffffffull) return fail(e, SJCS_ERR_INVALID, "autotune: at most 2^30-1 particles");
  rc = quiesce(e);
  if (rc != SJCS_OK) return rc;
  const size_t nc = (size_t)e->halos.n * (size_t)e->shells.nrads;
  const size_t trial_words = std::max<size_t>(nc ? nc : 1, occupancy_workspace_words() + 2 * kOccBins);
  DevBuf trial_counts;
  HIPCALL(e, ensure(trial_counts, 4 * trial_words), "hipMalloc(autotune)");
  const float rbox = e->shells.view.rbox;
  TuneSetting cur{e->cell_size_milli, e->x_split,     e->lanes_per_halo,
                  e->segment,         e->tiles,       e->tile_level,
                  e->tile_halos,      e->stage_particles, e->stage_halos};
  if (e->kernel == 1) cur.lanes = 64;
  std::string key;
  if (e->autotune_cache) {
    rc = workload_key(e, d_xyz, n, trial_counts.as<uint32_t>(), &key);
    if (rc != SJCS_OK) return rc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      apply_setting(e, it->second.first);
      if (e->halos.level != (int)e->tile_level) rc = apply_tiles(e, e->halos, (int)e->tile_level);
      if (rc == SJCS_OK && best_ms) *best_ms = it->second.second;
      e->autotune_hit = true;
      return rc;
    }
  }
  e->autotune_hit = false;
  // one timed sort + count (second of two runs) of this chunk with setting
  // st; the counts go to a scratch buffer (count rows do not depend on the
  // tiling, so the totals are left as they are)
  auto run = [&](const TuneSetting& st, float* ms) -> int {
    apply_setting(e, st);
    if (e->halos.level != (int)st.level) {
      int r = apply_tiles(e, e->halos, (int)st.level);
      if (r != SJCS_OK) return r;
    }
    const Grid g = choose_grid(n, rbox, e->cell_size_milli, e->x_split);
    bool use_tiles = false;
    int r = ensure_tiles(e, e->halos, e->shells.view, &use_tiles);
    if (r != SJCS_OK) return r;
    for (int rep = 0; rep < 2; ++rep) {
      HIPCALL(e, hipEventRecord(e->ev[2], e->stream), "hipEventRecord");
      r = build_index(e, e->ibs[0], packed(d_xyz), n, g, e->stream);
      if (r != SJCS_OK) return r;
      HIPCALL(e, count_chunk(e, e->ibs[0].view(), e->halos, e->shells.view, use_tiles,
                             trial_counts.as<uint32_t>(), false,
                             nullptr, e->stream),
              "autotune count");
      HIPCALL(e, hipEventRecord(e->ev[3], e->stream), "hipEventRecord");
      HIPCALL(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
    }
    HIPCALL(e, hipEventElapsedTime(ms, e->ev[2], e->ev[3]), "hipEventElapsedTime");
    return SJCS_OK;
  };
  float best = 0.0f;
  rc = run(cur, &best);
  auto try_set = [&](TuneSetting st) -> int {
    float ms = 0.0f;
    int r = run(st, &ms);
    if (r == SJCS_OK && ms < best) {
      best = ms;
      cur = st;
    }
    return r;
  };
  // coordinate descent: lanes, y/z cell edge (as a fraction of r_box), x
  // split, segment, then the halo tiles -- whether, how far apart their
  // halos may lie (tile level) and how many a tile holds
  for (int64_t l : {16, 64})
    if (rc == SJCS_OK && l != cur.lanes) {
      TuneSetting st = cur;
      st.lanes = l;
      rc = try_set(st);
    }
  if (rc == SJCS_OK && rbox > 0.0f && isfinite(rbox)) {
    for (double f : {0.3, 0.45, 0.6, 0.75, 0.9, 1.2}) {
      const int64_t milli = std::max<int64_t>(977, (int64_t)(f * rbox * 1000.0));
      if (rc == SJCS_OK && milli != cur.cell) {
        TuneSetting st = cur;
        st.cell = milli;
        rc = try_set(st);
      }
    }
  }
  for (int64_t x : {1, 2, 4, 8})
    if (rc == SJCS_OK && x != cur.xs) {
      TuneSetting st = cur;
      st.xs = x;
      rc = try_set(st);
    }
  for (int64_t sg : {4, 8})
    if (rc == SJCS_OK && sg != cur.seg) {
      TuneSetting st = cur;
      st.seg = sg;
      rc = try_set(st);
    }
  const int cap = tile_cap(e->shells.nrads);
  if (rc == SJCS_OK && e->shells.view.sorted && cap >= 2) {
    for (int64_t lv : {0, 1, 2, 3, 4})
      if (rc == SJCS_OK && (lv != cur.level || !cur.tiles)) {
        TuneSetting st = cur;
        st.level = lv;
        st.tiles = 1;
        rc = try_set(st);
      }
    if (rc == SJCS_OK && cur.tiles) {
      TuneSetting st = cur;
      st.tiles = 0;
      rc = try_set(st);
    }
    if (cur.tiles)
      for (int64_t th : {8, 16, 32, 0})
        if (rc == SJCS_OK && th < cap && th != cur.thalos) {
          TuneSetting st = cur;
          st.thalos = th;
          rc = try_set(st);
        }
  }
  // particle staging in LDS (depth = particles per workgroup), then the
  // halos per staging workgroup
  for (int64_t sp : {0, 1024, 2048, 4096})
    if (rc == SJCS_OK && sp != cur.stage) {
      TuneSetting st = cur;
      st.stage = sp;
      rc = try_set(st);
    }
  if (cur.stage)
    for (int64_t sh : {16, 32, 128, 0})
      if (rc == SJCS_OK && sh != cur.shalos) {
        TuneSetting st = cur;
        st.shalos = sh;
        rc = try_set(st);
      }
  // keep the winner; the index buffers of slot 0 no longer hold the last chunk
  apply_setting(e, cur);
  if (rc == SJCS_OK && e->halos.level != (int)cur.level) rc = apply_tiles(e, e->halos, (int)cur.level);
  e->last = -1;
  e->timed = false;
  trial_counts.release();
  if (rc == SJCS_OK && e->autotune_cache) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = std::make_pair(cur, best);
  }
  if (rc == SJCS_OK && best_ms) *best_ms = best;
  return rc;
}

int sjcs_set_option(sjcs_engine* e, const char* key, int64_t value) {
  if (!e || !key) return SJCS_ERR_INVALID;
  // options change which stream sorts run on and what graphs hold: let the
  // work in flight finish first
  int rq = use_device(e);
  if (rq == SJCS_OK) rq = quiesce(e);
  if (rq != SJCS_OK) return rq;
  ++e->epoch;
  if (strcmp(key, "tiles") == 0) {
    if (value < 0 || value > 1) return fail(e, SJCS_ERR_INVALID, "tiles: 0 or 1");
    e->tiles = value;
    return SJCS_OK;
  }
  if (strcmp(key, "tile_level") == 0) {
    if (value < 0 || value > kTileLevelMax)
      return fail(e, SJCS_ERR_INVALID, "tile_level: 0 to " + std::to_string(kTileLevelMax));
    e->tile_level = value;
    // count rows stay the Morton positions: counts in progress are kept
    if (e->have_halos && e->halos.level != (int)value) return apply_tiles(e, e->halos, (int)value);
    return SJCS_OK;
  }
  if (strcmp(key, "stage_particles") == 0) {
    if (value < 0 || (value > 0 && value < 256) || value > (int64_t)kStageCapMax)
      return fail(e, SJCS_ERR_INVALID,
                  "stage_particles: 0 or 256 to " + std::to_string(kStageCapMax));
    e->stage_particles = value;
    return SJCS_OK;
  }
  if (strcmp(key, "stage_halos") == 0) {
    if (value < 0 || (value > 0 && value < 4) || value > 1024)
      return fail(e, SJCS_ERR_INVALID, "stage_halos: 0 or 4 to 1024");
    e->stage_halos = value;
    return SJCS_OK;
  }
  if (strcmp(key, "autotune_cache") == 0) {
    if (value < 0 || value > 1) return fail(e, SJCS_ERR_INVALID, "autotune_cache: 0 or 1");
    e->autotune_cache = value;
    return SJCS_OK;
  }
  if (strcmp(key, "tile_halos") == 0) {
    if (value < 0 || value == 1 || value > kTileMaxHalos)
      return fail(e, SJCS_ERR_INVALID, "tile_halos: 0 or 2 to " + std::to_string(kTileMaxHalos));
    e->tile_halos = value;
    e->halos.tcap = 0;  // tiles are cut again at the next count
    return SJCS_OK;
  }
  if (strcmp(key, "graph") == 0) {
    if (value < 0 || value > 1) return fail(e, SJCS_ERR_INVALID, "graph: 0 or 1");
    e->graph = value;
    return SJCS_OK;
  }
  if (strcmp(key, "cell_size_milli") == 0) {
    if (value < 0 || (value > 0 && value < 977))  // at most 1024 cells per axis
      return fail(e, SJCS_ERR_INVALID, "cell_size_milli: 0 or >= 977");
    e->cell_size_milli = value;
    return SJCS_OK;
  }
  if (strcmp(key, "segment") == 0) {
    if (value != 0 && value != 4 && value != 8)
      return fail(e, SJCS_ERR_INVALID, "segment: 0, 4 or 8");
    e->segment = value;
    return SJCS_OK;
  }
  if (strcmp(key, "lanes_per_halo") == 0) {
    if (value != 0 && value != 8 && value != 16 && value != 64)
      return fail(e, SJCS_ERR_INVALID, "lanes_per_halo: 0, 8, 16 or 64");
    e->lanes_per_halo = value;
    return SJCS_OK;
  }
  if (strcmp(key, "fold_limit") == 0) {
    if (value < 0 || value > 0xffffffffll) return fail(e, SJCS_ERR_INVALID, "fold_limit: 0 to 2^32-1");
    e->fold_limit = value;
    return SJCS_OK;
  }
  if (strcmp(key, "piece_max") == 0) {
    if (value < 0) return fail(e, SJCS_ERR_INVALID, "piece_max: 0 or more");
    e->piece_max = value;
    return SJCS_OK;
  }
  if (strcmp(key, "overlap") == 0) {
    if (value < 0 || value > 1) return fail(e, SJCS_ERR_INVALID, "overlap: 0 or 1");
    e->overlap = value;
    return SJCS_OK;
  }
  if (strcmp(key, "x_split") == 0) {
    if (value < 0 || value > 16) return fail(e, SJCS_ERR_INVALID, "x_split: 0 to 16");
    e->x_split = value;
    return SJCS_OK;
  }
  if (strcmp(key, "index_sort") == 0) {
    if (value < 0 || value > 1) return fail(e, SJCS_ERR_INVALID, "index_sort: 0 or 1");
    e->index_sort = value;
    return SJCS_OK;
  }
  if (strcmp(key, "kernel") == 0) {
    if (value < 0 || value > 1) return fail(e, SJCS_ERR_INVALID, "kernel: 0 or 1");
    e->kernel = value;
    return SJCS_OK;
  }
  return fail(e, SJCS_ERR_INVALID, std::string("unknown option ") + key);
}

int sjcs_last_chunk_timings(sjcs_engine* e, float ms[3]) {
  if (!e || !ms) return SJCS_ERR_INVALID;
  ms[0] = ms[1] = ms[2] = 0.0f;
  if (!e->timed) return SJCS_OK;
  int rc = use_device(e);
  if (rc != SJCS_OK) return rc;
  rc = quiesce(e);
  if (rc != SJCS_OK) return rc;
  HIPCALL(e, hipEventElapsedTime(&ms[0], e->ev[0], e->ev[1]), "hipEventElapsedTime");
  HIPCALL(e, hipEventElapsedTime(&ms[1], e->ev[2], e->ev[3]), "hipEventElapsedTime");
  HIPCALL(e, hipEventElapsedTime(&ms[2], e->ev[4], e->ev[5]), "hipEventElapsedTime");
  return SJCS_OK;
}

int sjcs_phase_timings(sjcs_engine* e, float ms[5]) {
  if (!e || !ms) return SJCS_ERR_INVALID;
  int rc = sjcs_last_chunk_timings(e, ms);
  if (rc != SJCS_OK) return rc;
  ms[3] = ms[4] = 0.0f;
  rc = use_device(e);
  if (rc == SJCS_OK) rc = quiesce(e);
  if (rc != SJCS_OK) return rc;
  if (e->halo_timed)
    HIPCALL(e, hipEventElapsedTime(&ms[3], e->hev[0], e->hev[1]), "hipEventElapsedTime");
  if (e->readout_timed)
    HIPCALL(e, hipEventElapsedTime(&ms[4], e->rev[0], e->rev[1]), "hipEventElapsedTime");
  return SJCS_OK;
}

int sjcs_device_alloc(sjcs_engine* e, size_t bytes, void** d_ptr) {
  if (!e || !d_ptr) return SJCS_ERR_INVALID;
  *d_ptr = nullptr;
  int rc = use_device(e);
  if (rc != SJCS_OK) return rc;
  HIPCALL(e, hipMalloc(d_ptr, bytes ? bytes : 16), "hipMalloc");
  return SJCS_OK;
}

int sjcs_device_free(sjcs_engine* e, void* d_ptr) {
  if (!e) return SJCS_ERR_INVALID;
  int rc = use_device(e);
  if (rc != SJCS_OK) return rc;
  HIPCALL(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
  if (d_ptr) HIPCALL(e, hipFree(d_ptr), "hipFree");
  return SJCS_OK;
}

int sjcs_copy_to_device(sjcs_engine* e, void* d_dst, const void* src, size_t bytes) {
  if (!e || (!d_dst && bytes) || (!src && bytes)) return SJCS_ERR_INVALID;
  int rc = use_device(e);
  if (rc != SJCS_OK) return rc;
  if (bytes == 0) return SJCS_OK;
  HIPCALL(e, hipMemcpy(d_dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
  return SJCS_OK;
}

int sjcs_host_alloc(size_t bytes, void** out) {
  if (!out) return SJCS_ERR_INVALID;
  *out = nullptr;
  const hipError_t h = hipHostMalloc(out, bytes ? bytes : 16, hipHostMallocDefault);
  if (h != hipSuccess) {
    *out = nullptr;
    g_create_error = std::string("hipHostMalloc: ") + hipGetErrorString(h);
    return h == hipErrorOutOfMemory ? SJCS_ERR_NOMEM : SJCS_ERR_DEVICE;
  }
  return SJCS_OK;
}

void sjcs_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

int sjcs_copy_to_host(sjcs_engine* e, void* dst, const void* d_src, size_t bytes) {
  if (!e || (!dst && bytes) || (!d_src && bytes)) return SJCS_ERR_INVALID;
  int rc = use_device(e);
  if (rc != SJCS_OK) return rc;
  if (bytes == 0) return SJCS_OK;
  HIPCALL(e, hipMemcpy(dst, d_src, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
  return SJCS_OK;
}

// ---- index API ----

int sjcs_index_create(sjcs_engine* e, uint64_t ndata, sjcs_index** out) {
  if (!e || !out) return SJCS_ERR_INVALID;
  *out = nullptr;
  int rc = use_device(e);
  if (rc != SJCS_OK) return rc;
  // the count kernels address particles with 32-bit byte offsets
  if (ndata > 0x3fffffffull) return fail(e, SJCS_ERR_INVALID, "at most 2^30-1 particles");
  sjcs_index* ix = new sjcs_index();
  ix->e = e;
  ix->cap = ndata;
  hipError_t h = ensure(ix->aos, 12 * (size_t)(ndata ? ndata : 1));
  if (h != hipSuccess) {
    delete ix;
    return hip_fail(e, h, "hipMalloc(index)");
  }
  *out = ix;
  return SJCS_OK;
}

int sjcs_index_insert(sjcs_index* ix, const float* xyz, uint64_t n) {
  if (!ix) return SJCS_ERR_INVALID;
  sjcs_engine* e = ix->e;
  if (!xyz && n) return fail(e, SJCS_ERR_INVALID, "null points");
  if (ix->built) return fail(e, SJCS_ERR_STATE, "index already built");
  if (n > ix->cap - ix->n) return fail(e, SJCS_ERR_INVALID, "more points than the index holds");
  int rc = use_device(e);
  if (rc != SJCS_OK) return rc;
  if (n == 0) return SJCS_OK;
  HIPCALL(e, hipMemcpyAsync(ix->aos.as<char>() + 12 * ix->n, xyz, 12 * (size_t)n,
                            hipMemcpyHostToDevice, e->stream),
          "hipMemcpy(points)");
  HIPCALL(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
  ix->n += n;
  return SJCS_OK;
}

int sjcs_index_build(sjcs_index* ix) {
  if (!ix) return SJCS_ERR_INVALID;
  sjcs_engine* e = ix->e;
  int rc = use_device(e);
  if (rc != SJCS_OK) return rc;
  if (ix->built) return SJCS_OK;
  // no query radius is known yet: size cells for about two particles each
  const Grid g = choose_grid(ix->n, 0.0f, e->cell_size_milli, e->x_split);
  rc = quiesce(e);  // the sort workspace is shared with the chunk path
  if (rc != SJCS_OK) return rc;
  rc = build_index(e, ix->ib, packed(ix->aos.as<float>()), ix->n, g, e->stream);
  if (rc != SJCS_OK) return rc;
  HIPCALL(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
  ix->built = true;
  return SJCS_OK;
}

int sjcs_index_count(sjcs_index* ix, const float* centrals, size_t stride_bytes, uint64_t n,
                     const float* radiuses_sq, int nrads, uint64_t* counts) {
  if (!ix) return SJCS_ERR_INVALID;
  sjcs_engine* e = ix->e;
  if (!ix->built) return fail(e, SJCS_ERR_STATE, "index not built (sjcs_index_build)");
  if ((!centrals && n) || stride_bytes < 12 || (!counts && n))
    return fail(e, SJCS_ERR_INVALID, "null pointer or stride < 12");
  int rc = use_device(e);
  if (rc == SJCS_OK) rc = quiesce(e);  // the sort workspace is shared with the chunk path
  if (rc != SJCS_OK) return rc;
  if (n == 0) return SJCS_OK;
  ShellState sh;
  HaloState hs;
  DevBuf pts, dcounts, ucounts;
  rc = make_shells(e, sh, radiuses_sq, nrads);
  if (rc == SJCS_OK && !(sh.view.rbox < 499.0f) && ix->n > 0xffffffffull / 8)
    rc = fail(e, SJCS_ERR_INVALID, "index too large for this radius (counts could pass 2^32)");
  if (rc == SJCS_OK) rc = upload_points(e, pts, centrals, stride_bytes, n);
  // the query halos in Morton order, tiled as the engine's options say:
  // halos at one spot (or, at a tile level above 0, near one another) share
  // their candidates, as the chunk loop's halos do
  if (rc == SJCS_OK) rc = sort_halos(e, hs, pts.as<float>(), 12, n, true, false);
  bool use_tiles = false;
  if (rc == SJCS_OK) rc = ensure_tiles(e, hs, sh.view, &use_tiles);
  const size_t nc = (size_t)n * (size_t)nrads;
  std::vector<uint64_t> host;
  if (rc == SJCS_OK) {
    hipError_t h = ensure(dcounts, nc * sizeof(uint32_t));
    if (h == hipSuccess) h = ensure(ucounts, nc * sizeof(uint64_t));
    if (h == hipSuccess)
      h = count_chunk(e, ix->ib.view(), hs, sh.view, use_tiles, dcounts.as<uint32_t>(), false,
                      nullptr, e->stream);
    if (h == hipSuccess)
      h = readout_counts(dcounts.as<uint32_t>(), nullptr, hs.inv.as<uint32_t>(),
                         ucounts.as<uint64_t>(), n, nrads, e->stream);
    if (h == hipSuccess) {
      host.resize(nc);
      h = hipMemcpyAsync(host.data(), ucounts.p, nc * sizeof(uint64_t), hipMemcpyDeviceToHost,
                         e->stream);
    }
    if (h == hipSuccess) h = hipStreamSynchronize(e->stream);
    if (h != hipSuccess) rc = hip_fail(e, h, "index count");
  }
  if (rc == SJCS_OK)
    for (size_t i = 0; i < nc; ++i) counts[i] += host[i];
  sh.d_r2.release();
  sh.d_lut.release();
  hs.release();
  pts.release();
  dcounts.release();
  ucounts.release();
  return rc;
}

void sjcs_index_destroy(sjcs_index* ix) {
  if (!ix) return;
  (void)hipSetDevice(ix->e->device);
  (void)hipStreamSynchronize(ix->e->stream);
  ix->aos.release();
  ix->ib.release();
  delete ix;
}

int sjcs_index_count_noalloc(sjcs_index* ix, const float* d_centrals, uint64_t n,
                             const float* radiuses_sq, int nrads, uint64_t* d_counts) {
  if (!ix) return SJCS_ERR_INVALID;
  sjcs_engine* e = ix->e;
  if (!ix->built) return fail(e, SJCS_ERR_STATE, "index not built (sjcs_index_build)");
  if ((!d_centrals && n) || (!d_counts && n) || !radiuses_sq || nrads <= 0)
    return fail(e, SJCS_ERR_INVALID, "null pointer or nrads < 1");
  if (nrads > kMaxLdsShells)
    return fail(e, SJCS_ERR_INVALID, "noalloc count: at most 1024 shells");
  // scratch layout: shell table, thresholds, then per halo its Morton key,
  // x, y, z and row (u32 each) and its u32 counts, each part 256-aligned
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t off_r2 = up(sizeof(ShellEntry) * kLutSize);
  const size_t off_h = off_r2 + up(4 * (size_t)kMaxLdsShells);
  const size_t hb = n <= kNoAllocScratchBytes ? up(4 * (size_t)n) : kNoAllocScratchBytes;
  const size_t off_c = off_h + 5 * hb;
  const double need = (double)off_c + 4.0 * (double)n * (double)nrads;
  if (need > (double)kNoAllocScratchBytes)
    return fail(e, SJCS_ERR_INVALID,
                "noalloc count: " + std::to_string((uint64_t)need) +
                    " bytes of device scratch needed, at most " +
                    std::to_string(kNoAllocScratchBytes) +
                    " (20 + 4*nrads bytes per halo plus 12 KiB)");
  int rc = use_device(e);
  if (rc != SJCS_OK) return rc;
  if (!e->scratch) return fail(e, SJCS_ERR_STATE, "noalloc count: no device scratch");
  if (n == 0) return SJCS_OK;
  Shells v{};
  ShellEntry lut[kLutSize];
  int nlut = 1;
  shell_table_into(radiuses_sq, nrads, &v, lut, &nlut);
  if (!(v.rbox < 499.0f) && ix->n > 0xffffffffull / 8)
    return fail(e, SJCS_ERR_INVALID, "index too large for this radius (counts could pass 2^32)");
  char* base = static_cast<char*>(e->scratch);
  uint32_t* hw = reinterpret_cast<uint32_t*>(base + off_h);
  const size_t hwords = hb / 4;
  HIPCALL(e, hipMemcpyAsync(base, lut, sizeof(ShellEntry) * (size_t)nlut, hipMemcpyHostToDevice,
                            e->stream),
          "hipMemcpy(shell table)");
  HIPCALL(e, hipMemcpyAsync(base + off_r2, radiuses_sq, 4 * (size_t)nrads, hipMemcpyHostToDevice,
                            e->stream),
          "hipMemcpy(shells)");
  // centres -> SoA + rows, counted in input order (no halo sort: no sort
  // workspace), so the u32 counts add straight into the caller's u64 rows
  HIPCALL(e, halo_keys(d_centrals, 12, n, hw, reinterpret_cast<float*>(hw + hwords),
                       reinterpret_cast<float*>(hw + 2 * hwords),
                       reinterpret_cast<float*>(hw + 3 * hwords), hw + 4 * hwords, e->stream),
          "halo keys");
  v.r2 = reinterpret_cast<const float*>(base + off_r2);
  v.lut = reinterpret_cast<const ShellEntry*>(base);
  HaloView hv;
  hv.x = reinterpret_cast<const float*>(hw + hwords);
  hv.y = reinterpret_cast<const float*>(hw + 2 * hwords);
  hv.z = reinterpret_cast<const float*>(hw + 3 * hwords);
  hv.slot = nullptr;  // input order: halo i's counts are row i
  hv.n = n;
  uint32_t* c32 = reinterpret_cast<uint32_t*>(base + off_c);
  HIPCALL(e, count_halos(ix->ib.view(), hv, v, c32, false, nullptr, lanes(e, ix->n, v.rbox),
                         seg_len(e), e->stream),
          "noalloc count");
  HIPCALL(e, widen_counts(c32, d_counts, n * (uint64_t)nrads, false, e->stream), "add counts");
  HIPCALL(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
  return SJCS_OK;
}

}  // extern "C"

// ---- serving loop ----
//
// Requests (a halo batch with its own shells) against one built index.  A
// worker thread admits queued requests in arrival order while a slot (own
// stream and buffers) is free and the requests in flight fit the device
// memory budget; admitted requests run concurrently on their streams.  The
// worker retires finished requests (event query), hands their counts to a
// pool of copy-out threads (page-locked slot buffer -> caller's buffer, in
// pieces) and admits the next ones, so requests that arrive while others run
// start as soon as room frees up; a slot is free again once its counts are
// copied out.  Each request is computed on its own, so its counts do not
// depend on what else is in flight.

namespace {

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct ServerRequest {
  uint64_t ticket = 0;
  uint64_t n = 0;
  int nrads = 0;
  std::vector<float> xyz;  // packed halo centres
  std::vector<float> r2;
  Shells shells{};         // device pointers filled at launch
  std::vector<ShellEntry> lut;
  uint64_t* out = nullptr;
  size_t need = 0;         // device bytes while in flight
  int status = 1;          // 1 queued or running, 0 done, < 0 failed
  std::string err;
  double t_submit = 0, t_admit = 0, t_done = 0;
};

// Device bytes a request holds while in flight.
size_t request_bytes(uint64_t n, int nrads) {
  return 12 * (size_t)n + 2 * 20 * (size_t)n + 4 * radix_workspace_words(n) +
         12 * (size_t)n * (size_t)nrads + 8 * (size_t)kLutSize + 4 * (size_t)nrads + 4096;
}

// Below this many halos a request is counted in input order (a sort costs
// more launches than it saves).
constexpr uint64_t kServeSortMin = 4096;
// A request's halos are ordered by the top 15 bits of their Morton keys
// (cells ~31 units wide): two sort passes instead of four, and the count's
// neighbouring halos still share most cell rows.
constexpr int kServeKeyBits = 15;

struct ServerSlot {
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  DevBuf xyz, key[2], x[2], y[2], z[2], row[2], ws, c32, c64;  // xyz: the request's upload
  void* pin = nullptr;
  size_t pin_bytes = 0;
  // the finished counts are being copied out of `pin` (pieces left; under
  // the server mutex): the slot stays taken until the last piece is done
  bool copying = false;
  int pieces_left = 0;
  std::shared_ptr<ServerRequest> req;
  void release() {
    xyz.release();
    for (int i = 0; i < 2; ++i) {
      key[i].release();
      x[i].release();
      y[i].release();
      z[i].release();
      row[i].release();
    }
    ws.release();
    c32.release();
    c64.release();
    if (pin) (void)hipHostFree(pin);
    pin = nullptr;
    pin_bytes = 0;
    if (done) (void)hipEventDestroy(done);
    if (st) (void)hipStreamDestroy(st);
    done = nullptr;
    st = nullptr;
  }
};

}  // namespace

struct sjcs_server {
  sjcs_index* ix = nullptr;
  int device = 0;
  std::vector<ServerSlot> slots;
  uint64_t budget = 0, used = 0;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;  // a request finished
  std::deque<std::shared_ptr<ServerRequest>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<ServerRequest>> reqs;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread worker;
  std::string err;
  // copy-out pool: a request's counts go from its slot's page-locked buffer
  // to the caller's memory in pieces on these threads, so the worker keeps
  // launching while first-touch page faults of the callers' buffers are
  // taken on several cores at once
  struct CopyJob {
    ServerSlot* slot;
    size_t off, bytes;
  };
  std::mutex cmu;
  std::condition_variable ccv;
  std::deque<CopyJob> cjobs;
  bool cstop = false;
  std::vector<std::thread> copiers;
};

namespace {

#define SRVCALL(expr, what)                                          \
  do {                                                               \
    hipError_t h_ = (expr);                                          \
    if (h_ != hipSuccess) {                                          \
      r.err = std::string(what) + ": " + hipGetErrorString(h_);      \
      return h_ == hipErrorOutOfMemory ? SJCS_ERR_NOMEM : SJCS_ERR_DEVICE; \
    }                                                                \
  } while (0)

// Enqueues one request on its slot: upload, halo order, count, read-out in
// input order, download into the slot's pinned buffer, event.
int serve_launch(sjcs_server* sv, ServerSlot& sl, ServerRequest& r) {
  const uint64_t n = r.n;
  const int K = r.nrads;
  const size_t b = 4 * (size_t)(n ? n : 1);
  const size_t nc = (size_t)n * (size_t)K;
  // one upload from page-locked memory: [halos 12n | thresholds 4K | shell
  // table], each part 256-byte aligned
  const size_t off_r2 = (12 * (size_t)n + 255) & ~(size_t)255;
  const size_t off_lut = (off_r2 + 4 * (size_t)K + 255) & ~(size_t)255;
  const size_t in_bytes = off_lut + sizeof(ShellEntry) * r.lut.size();
  // slot buffers grow by half at least, so requests of other shapes seldom
  // reallocate (a device free synchronises the whole device)
  SRVCALL(ensure_grow(sl.xyz, in_bytes), "hipMalloc");
  for (int i = 0; i < 2; ++i) {
    SRVCALL(ensure_grow(sl.key[i], b), "hipMalloc");
    SRVCALL(ensure_grow(sl.x[i], b), "hipMalloc");
    SRVCALL(ensure_grow(sl.y[i], b), "hipMalloc");
    SRVCALL(ensure_grow(sl.z[i], b), "hipMalloc");
    SRVCALL(ensure_grow(sl.row[i], b), "hipMalloc");
  }
  SRVCALL(ensure_grow(sl.ws, 4 * radix_workspace_words(n)), "hipMalloc");
  SRVCALL(ensure_grow(sl.c32, 4 * (nc ? nc : 1)), "hipMalloc");
  SRVCALL(ensure_grow(sl.c64, 8 * (nc ? nc : 1)), "hipMalloc");
  const size_t pin = std::max(in_bytes, 8 * nc) + 64;
  if (sl.pin_bytes < pin) {
    const size_t grown = std::max(pin, sl.pin_bytes + sl.pin_bytes / 2);
    if (sl.pin) (void)hipHostFree(sl.pin);
    sl.pin = nullptr;
    sl.pin_bytes = 0;
    SRVCALL(hipHostMalloc(&sl.pin, grown, hipHostMallocDefault), "hipHostMalloc");
    sl.pin_bytes = grown;
  }
  char* pin_in = static_cast<char*>(sl.pin);
  if (n) memcpy(pin_in, r.xyz.data(), 12 * (size_t)n);
  memcpy(pin_in + off_r2, r.r2.data(), 4 * (size_t)K);
  memcpy(pin_in + off_lut, r.lut.data(), sizeof(ShellEntry) * r.lut.size());
  SRVCALL(hipMemcpyAsync(sl.xyz.p, sl.pin, in_bytes, hipMemcpyHostToDevice, sl.st), "hipMemcpy");
  SRVCALL(halo_keys(sl.xyz.as<float>(), 12, n, sl.key[0].as<uint32_t>(), sl.x[0].as<float>(),
                    sl.y[0].as<float>(), sl.z[0].as<float>(), sl.row[0].as<uint32_t>(), sl.st,
                    kServeKeyBits),
          "halo keys");
  int sel = 0;
  if (n >= kServeSortMin) {
    uint32_t* keys[2] = {sl.key[0].as<uint32_t>(), sl.key[1].as<uint32_t>()};
    uint32_t* pay[2][4] = {{sl.x[0].as<uint32_t>(), sl.y[0].as<uint32_t>(), sl.z[0].as<uint32_t>(),
                            sl.row[0].as<uint32_t>()},
                           {sl.x[1].as<uint32_t>(), sl.y[1].as<uint32_t>(), sl.z[1].as<uint32_t>(),
                            sl.row[1].as<uint32_t>()}};
    SRVCALL(radix_sort_pairs(keys, pay, 4, n, kServeKeyBits, sl.ws.as<uint32_t>(), sl.st, &sel),
            "halo sort");
  }
  HaloView hv;
  hv.x = sl.x[sel].as<float>();
  hv.y = sl.y[sel].as<float>();
  hv.z = sl.z[sel].as<float>();
  hv.slot = nullptr;  // count rows in sorted order
  hv.n = n;
  r.shells.r2 = reinterpret_cast<float*>(sl.xyz.as<char>() + off_r2);
  r.shells.lut = reinterpret_cast<ShellEntry*>(sl.xyz.as<char>() + off_lut);
  const IndexView ix = sv->ix->ib.view();
  const double side = 2.0 * (double)r.shells.rbox / 1000.0;
  const int lanes_auto = !((double)ix.n * side * side * side <= 4096.0) ? 64 : 16;
  SRVCALL(count_halos(ix, hv, r.shells, sl.c32.as<uint32_t>(), false, nullptr, lanes_auto, 4, sl.st),
          "count");
  if (n >= kServeSortMin) {
    // the other key buffer is free after the sort: it takes the inverse rows
    uint32_t* inv = sl.key[sel ^ 1].as<uint32_t>();
    SRVCALL(invert_rows(sl.row[sel].as<uint32_t>(), inv, n, sl.st), "read out");
    SRVCALL(readout_counts(sl.c32.as<uint32_t>(), nullptr, inv, sl.c64.as<uint64_t>(), n, K, sl.st),
            "read out");
  } else {
    // unsorted: the count rows are the input rows, widened in place order
    SRVCALL(widen_counts(sl.c32.as<uint32_t>(), sl.c64.as<uint64_t>(), nc, true, sl.st),
            "read out");
  }
  if (nc)
    SRVCALL(hipMemcpyAsync(sl.pin, sl.c64.p, 8 * nc, hipMemcpyDeviceToHost, sl.st), "hipMemcpy");
  SRVCALL(hipEventRecord(sl.done, sl.st), "hipEventRecord");
  return SJCS_OK;
}
#undef SRVCALL

// Bytes of counts one copy-out job moves, and the copy-out threads.
constexpr size_t kCopyPiece = (size_t)4 << 20;
constexpr int kCopyThreads = 8;

// The request on slot sl is over (status set): frees the slot and wakes the
// waiters.  Caller holds sv->mu.
void finish_request(sjcs_server* sv, ServerSlot& sl) {
  ServerRequest& r = *sl.req;
  r.t_done = now_s(sv->t0);
  r.xyz.clear();
  r.xyz.shrink_to_fit();
  sv->used -= r.need;
  sl.req.reset();
  sl.copying = false;
  sv->cv.notify_all();
}

// Copy-out thread: pieces of finished requests' counts, page-locked buffer
// -> caller's memory; the last piece of a request completes it.
void copy_loop(sjcs_server* sv) {
  for (;;) {
    sjcs_server::CopyJob job;
    {
      std::unique_lock<std::mutex> lk(sv->cmu);
      sv->ccv.wait(lk, [&] { return sv->cstop || !sv->cjobs.empty(); });
      if (sv->cjobs.empty()) return;  // stopping, nothing left
      job = sv->cjobs.front();
      sv->cjobs.pop_front();
    }
    ServerSlot& sl = *job.slot;
    char* dst = reinterpret_cast<char*>(sl.req->out);
    memcpy(dst + job.off, static_cast<const char*>(sl.pin) + job.off, job.bytes);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (--sl.pieces_left == 0) {
      sl.req->status = 0;
      finish_request(sv, sl);
    }
  }
}

void serve_loop(sjcs_server* sv) {
  (void)hipSetDevice(sv->device);
  for (;;) {
    bool busy = false;
    // retire finished requests: their counts go to the copy-out pool
    for (ServerSlot& sl : sv->slots) {
      std::shared_ptr<ServerRequest> rq;
      {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (!sl.req) continue;
        busy = true;
        if (sl.copying) continue;
        rq = sl.req;
      }
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      ServerRequest& r = *rq;
      const size_t bytes = 8 * (size_t)r.n * (size_t)r.nrads;
      if (q == hipSuccess && bytes > 0) {
        const int pieces = (int)((bytes + kCopyPiece - 1) / kCopyPiece);
        {
          std::lock_guard<std::mutex> lk(sv->mu);
          sl.copying = true;
          sl.pieces_left = pieces;
        }
        {
          std::lock_guard<std::mutex> lk(sv->cmu);
          for (int i = 0; i < pieces; ++i) {
            const size_t off = (size_t)i * kCopyPiece;
            sv->cjobs.push_back({&sl, off, std::min(kCopyPiece, bytes - off)});
          }
        }
        sv->ccv.notify_all();
        continue;
      }
      std::lock_guard<std::mutex> lk(sv->mu);
      if (q == hipSuccess) {
        r.status = 0;
      } else {
        r.err = std::string("request: ") + hipGetErrorString(q);
        r.status = SJCS_ERR_DEVICE;
      }
      finish_request(sv, sl);
    }
    // admit queued requests in order while a slot and the budget allow
    for (;;) {
      std::shared_ptr<ServerRequest> next;
      ServerSlot* free_slot = nullptr;
      {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->queue.empty()) break;
        for (ServerSlot& sl : sv->slots)
          if (!sl.req) {
            free_slot = &sl;
            break;
          }
        if (!free_slot) break;
        const auto& f = sv->queue.front();
        // a request larger than the whole budget still runs, alone
        if (sv->used && sv->used + f->need > sv->budget) break;
        next = f;
        sv->queue.pop_front();
        sv->used += next->need;
        next->t_admit = now_s(sv->t0);
      }
      busy = true;
      const int rc = serve_launch(sv, *free_slot, *next);
      if (rc != SJCS_OK) {
        (void)hipStreamSynchronize(free_slot->st);
        std::lock_guard<std::mutex> lk(sv->mu);
        next->status = rc;
        next->t_done = now_s(sv->t0);
        sv->used -= next->need;
        sv->cv.notify_all();
      } else {
        std::lock_guard<std::mutex> lk(sv->mu);
        free_slot->req = next;
      }
    }
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      bool active = false;
      for (const ServerSlot& sl : sv->slots) active = active || (bool)sl.req;
      if (sv->stop && sv->queue.empty() && !active) return;
      if (!busy && sv->queue.empty())
        sv->cv.wait_for(lk, std::chrono::milliseconds(1));  // idle: a submit wakes us
    }
    if (busy) std::this_thread::yield();
  }
}

}  // namespace

extern "C" {

int sjcs_server_create(sjcs_index* ix, int slots, uint64_t mem_budget, sjcs_server** out) {
  if (!ix || !out) return SJCS_ERR_INVALID;
  *out = nullptr;
  sjcs_engine* e = ix->e;
  if (!ix->built) return fail(e, SJCS_ERR_STATE, "index not built (sjcs_index_build)");
  if (slots < 0 || slots > 64) return fail(e, SJCS_ERR_INVALID, "slots: 0 (automatic) to 64");
  int rc = use_device(e);
  if (rc != SJCS_OK) return rc;
  size_t free_b = 0, total_b = 0;
  HIPCALL(e, hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
  sjcs_server* sv = new sjcs_server();
  sv->ix = ix;
  sv->device = e->device;
  sv->budget = mem_budget ? mem_budget : free_b / 2;
  sv->t0 = std::chrono::steady_clock::now();
  sv->slots.resize(slots ? slots : 8);
  hipError_t h = hipSuccess;
  for (ServerSlot& sl : sv->slots) {
    if (h == hipSuccess) h = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking);
    if (h == hipSuccess) h = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
  }
  if (h != hipSuccess) {
    for (ServerSlot& sl : sv->slots) sl.release();
    delete sv;
    return hip_fail(e, h, "server streams");
  }
  sv->worker = std::thread(serve_loop, sv);
  for (int i = 0; i < kCopyThreads; ++i) sv->copiers.emplace_back(copy_loop, sv);
  *out = sv;
  return SJCS_OK;
}

int sjcs_server_submit(sjcs_server* sv, const float* centrals, size_t stride_bytes, uint64_t n,
                       const float* radiuses_sq, int nrads, uint64_t* counts, uint64_t* ticket) {
  if (!sv || !ticket) return SJCS_ERR_INVALID;
  *ticket = 0;
  if ((!centrals && n) || stride_bytes < 12 || (!counts && n) || !radiuses_sq || nrads <= 0) {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->err = "submit: null pointer, stride < 12 or nrads < 1";
    return SJCS_ERR_INVALID;
  }
  if (n > 0x3fffffffull) {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->err = "submit: at most 2^30-1 halos per request";
    return SJCS_ERR_INVALID;
  }
  auto r = std::make_shared<ServerRequest>();
  shell_table(radiuses_sq, nrads, &r->shells, &r->lut);
  // a request's counts are u32 on the device: a particle counts once per
  // image it falls in, 8 when the box is about as wide as the periodic box
  if (!(r->shells.rbox < 499.0f) && sv->ix->n > 0xffffffffull / 8) {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->err = "submit: index too large for this radius (counts could pass 2^32)";
    return SJCS_ERR_INVALID;
  }
  r->n = n;
  r->nrads = nrads;
  r->xyz.resize(3 * (size_t)n);
  const char* src = reinterpret_cast<const char*>(centrals);
  for (uint64_t i = 0; i < n; ++i) memcpy(&r->xyz[3 * i], src + i * stride_bytes, 12);
  r->r2.assign(radiuses_sq, radiuses_sq + nrads);
  r->out = counts;
  r->need = request_bytes(n, nrads);
  std::lock_guard<std::mutex> lk(sv->mu);
  r->ticket = sv->next_ticket++;
  r->t_submit = now_s(sv->t0);
  sv->reqs[r->ticket] = r;
  sv->queue.push_back(r);
  *ticket = r->ticket;
  sv->cv.notify_all();
  return SJCS_OK;
}

int sjcs_server_test(sjcs_server* sv, uint64_t ticket) {
  if (!sv) return SJCS_ERR_INVALID;
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return SJCS_ERR_INVALID;
  const int st = it->second->status;
  return st == 0 ? 1 : st > 0 ? 0 : st;
}

int sjcs_server_wait(sjcs_server* sv, uint64_t ticket) {
  if (!sv) return SJCS_ERR_INVALID;
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return SJCS_ERR_INVALID;
  std::shared_ptr<ServerRequest> r = it->second;
  sv->cv.wait(lk, [&] { return r->status <= 0; });
  if (r->status < 0) sv->err = r->err;
  return r->status;
}

int sjcs_server_times(sjcs_server* sv, uint64_t ticket, double* submitted, double* admitted,
                      double* done) {
  if (!sv) return SJCS_ERR_INVALID;
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return SJCS_ERR_INVALID;
  if (submitted) *submitted = it->second->t_submit;
  if (admitted) *admitted = it->second->t_admit;
  if (done) *done = it->second->t_done;
  return SJCS_OK;
}

int sjcs_server_release(sjcs_server* sv, uint64_t ticket) {
  if (!sv) return SJCS_ERR_INVALID;
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return SJCS_ERR_INVALID;
  if (it->second->status > 0) return SJCS_ERR_STATE;  // still queued or running
  sv->reqs.erase(it);
  return SJCS_OK;
}

const char* sjcs_server_error(const sjcs_server* sv) { return sv ? sv->err.c_str() : ""; }

void sjcs_server_destroy(sjcs_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
    sv->cv.notify_all();
  }
  if (sv->worker.joinable()) sv->worker.join();
  {
    std::lock_guard<std::mutex> lk(sv->cmu);
    sv->cstop = true;
  }
  sv->ccv.notify_all();
  for (std::thread& t : sv->copiers)
    if (t.joinable()) t.join();
  (void)hipSetDevice(sv->device);
  for (ServerSlot& sl : sv->slots) {
    if (sl.st) (void)hipStreamSynchronize(sl.st);
    sl.release();
  }
  delete sv;
}


// ---- R-tree API (rt_rcount's index and count) ----

}  // extern "C"

struct sjcs_rtree {
  sjcs_engine* e = nullptr;  // null: a host-only tree (build and inspect)
  RtTree tree;
  std::string err;
  bool on_device = false;  // the device arrays hold the current tree
  DevBuf lo[3], hi[3], child;
  uint32_t root = 0, inner = 0;  // the root's reference, the inner nodes
  void release() {
    child.release();
    for (int d = 0; d < 3; ++d) {
      lo[d].release();
      hi[d].release();
    }
    on_device = false;
  }
};

namespace {

int rt_fail(sjcs_rtree* rt, int code, const std::string& msg) {
  rt->err = msg;
  if (rt->e) rt->e->err = msg;
  return code;
}

int rt_hip_fail(sjcs_rtree* rt, hipError_t h, const char* what) {
  return rt_fail(rt, h == hipErrorOutOfMemory ? SJCS_ERR_NOMEM : SJCS_ERR_DEVICE,
                 std::string(what) + ": " + hipGetErrorString(h));
}

template <class T>
hipError_t upload_vec(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  hipError_t h = ensure(b, sizeof(T) * (v.size() ? v.size() : 1));
  if (h == hipSuccess && !v.empty())
    h = hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s);
  return h;
}

// The tree's node arrays onto the device (once per state of the tree).
int rt_upload(sjcs_rtree* rt) {
  if (rt->on_device) return SJCS_OK;
  sjcs_engine* e = rt->e;
  RtFlat f;
  try {
    rt->tree.flatten(&f);
  } catch (const std::bad_alloc&) {
    return rt_fail(rt, SJCS_ERR_NOMEM, "host memory for the R-tree layout");
  }
  if (rt->tree.nodes() > kRtNodeMask)
    return rt_fail(rt, SJCS_ERR_INVALID, "R-tree with more than 2^28-1 nodes");
  rt->release();
  hipError_t h = hipSuccess;
  for (int d = 0; d < 3 && h == hipSuccess; ++d) {
    h = upload_vec(rt->lo[d], f.lo[d], e->stream);
    if (h == hipSuccess) h = upload_vec(rt->hi[d], f.hi[d], e->stream);
  }
  if (h == hipSuccess) h = upload_vec(rt->child, f.child, e->stream);
  // the host vectors go out of scope here: wait for the copies
  if (h == hipSuccess) h = hipStreamSynchronize(e->stream);
  if (h != hipSuccess) {
    rt->release();
    return rt_hip_fail(rt, h, "R-tree upload");
  }
  rt->root = f.root;
  rt->inner = f.inner;
  rt->on_device = true;
  return SJCS_OK;
}

}  // namespace

extern "C" {

int sjcs_rtree_create(sjcs_engine* e, sjcs_rtree** out) {
  if (!out) return SJCS_ERR_INVALID;
  *out = nullptr;
  sjcs_rtree* rt = nullptr;
  try {
    rt = new sjcs_rtree();
  } catch (const std::bad_alloc&) {
    return SJCS_ERR_NOMEM;
  }
  rt->e = e;
  *out = rt;
  return SJCS_OK;
}

int sjcs_rtree_insert(sjcs_rtree* rt, const float* xyz, uint64_t n) {
  if (!rt) return SJCS_ERR_INVALID;
  if (!xyz && n) return rt_fail(rt, SJCS_ERR_INVALID, "null points");
  std::string why;
  try {
    if (!rt->tree.insert(xyz, n, &why)) return rt_fail(rt, SJCS_ERR_INVALID, why);
  } catch (const std::bad_alloc&) {
    return rt_fail(rt, SJCS_ERR_NOMEM, "host memory for the R-tree");
  }
  if (n) rt->on_device = false;
  return SJCS_OK;
}

int sjcs_rtree_shape(const sjcs_rtree* rt, uint64_t* npoints, uint64_t* nleaves,
                     uint64_t* nnodes, int* height) {
  if (!rt) return SJCS_ERR_INVALID;
  if (npoints) *npoints = rt->tree.points();
  if (nleaves) *nleaves = rt->tree.leaves();
  if (nnodes) *nnodes = rt->tree.nodes();
  if (height) *height = rt->tree.height();
  return SJCS_OK;
}

int sjcs_rtree_leaves(const sjcs_rtree* rt, uint64_t* ids, uint32_t* leaf_len) {
  if (!rt || ((!ids || !leaf_len) && rt->tree.points())) return SJCS_ERR_INVALID;
  // a failed insert can leave entries of a split in the tree that no id
  // count accounts for
  if (!rt->tree.ok()) return SJCS_ERR_STATE;
  if (rt->tree.points() == 0) return SJCS_OK;
  rt->tree.leaf_ids(ids, leaf_len);
  return SJCS_OK;
}

const char* sjcs_rtree_error(const sjcs_rtree* rt) { return rt ? rt->err.c_str() : ""; }

}  // extern "C"

namespace {

int rt_check(sjcs_rtree* rt, int nrads) {
  if (!rt->e) return rt_fail(rt, SJCS_ERR_STATE, "a host-only R-tree cannot count (no engine)");
  if (!rt->tree.ok())
    return rt_fail(rt, SJCS_ERR_STATE, "the R-tree is unusable after a failed insert");
  if (nrads < 1 || nrads > kRtMaxShells)
    return rt_fail(rt, SJCS_ERR_INVALID, "R-tree count: 1 to 1024 shells");
  if (9 * rt->tree.height() + 1 > kRtStack)
    return rt_fail(rt, SJCS_ERR_INVALID, "R-tree too deep for the count's stack");
  return SJCS_OK;
}

// d_out[h * nrads + k] += the shell-k count of centre h, for n centres
// (packed xyz) at d_pts, both on the device; returns when done.
int rt_count_into(sjcs_rtree* rt, const float* d_pts, uint64_t n, const float* radiuses_sq,
                  int nrads, uint64_t* d_out) {
  sjcs_engine* e = rt->e;
  bool sorted = true;
  for (int k = 1; k < nrads; ++k) sorted &= radiuses_sq[k - 1] <= radiuses_sq[k];
  // MultiRangeCount.c:17: sqrt of the last threshold, in double, to FP32
  const float rbox = (float)sqrt((double)radiuses_sq[nrads - 1]);
  DevBuf r2, dcounts;
  HaloState hs;
  // the queries in Morton order: neighbours visit the same nodes
  int rc = sort_halos(e, hs, d_pts, 12, n, false, false);
  if (rc == SJCS_OK) {
    hipError_t h = ensure(r2, sizeof(float) * nrads);
    if (h == hipSuccess)
      h = hipMemcpyAsync(r2.p, radiuses_sq, sizeof(float) * nrads, hipMemcpyHostToDevice,
                         e->stream);
    if (h == hipSuccess) h = ensure(dcounts, (size_t)n * (size_t)nrads * sizeof(uint32_t));
    if (h == hipSuccess) {
      RtView v;
      v.lox = rt->lo[0].as<float>();
      v.loy = rt->lo[1].as<float>();
      v.loz = rt->lo[2].as<float>();
      v.hix = rt->hi[0].as<float>();
      v.hiy = rt->hi[1].as<float>();
      v.hiz = rt->hi[2].as<float>();
      v.child = rt->child.as<uint32_t>();
      h = rt_count(v, rt->root, rt->inner, hs.bx.as<float>(), hs.by.as<float>(),
                   hs.bz.as<float>(), n, r2.as<float>(), nrads, sorted, rbox,
                   dcounts.as<uint32_t>(), e->stream);
    }
    if (h == hipSuccess)
      h = rt_add_counts(dcounts.as<uint32_t>(), hs.inv.as<uint32_t>(), d_out, n, nrads,
                        e->stream);
    if (h == hipSuccess) h = hipStreamSynchronize(e->stream);
    if (h != hipSuccess) rc = rt_hip_fail(rt, h, "R-tree count");
  }
  r2.release();
  dcounts.release();
  hs.release();
  return rc;
}

}  // namespace

extern "C" {

int sjcs_rtree_count(sjcs_rtree* rt, const float* centrals, size_t stride_bytes, uint64_t n,
                     const float* radiuses_sq, int nrads, uint64_t* counts) {
  if (!rt) return SJCS_ERR_INVALID;
  if ((!centrals && n) || stride_bytes < 12 || (!counts && n) || !radiuses_sq)
    return rt_fail(rt, SJCS_ERR_INVALID, "null pointer or stride < 12");
  int rc = rt_check(rt, nrads);
  if (rc != SJCS_OK) return rc;
  sjcs_engine* e = rt->e;
  rc = use_device(e);
  if (rc == SJCS_OK) rc = quiesce(e);  // the pinned staging is shared with the engine
  if (rc == SJCS_OK) rc = rt_upload(rt);
  if (rc != SJCS_OK || n == 0) return rc;
  const size_t nc = (size_t)n * (size_t)nrads;
  DevBuf pts, ucounts;
  std::vector<uint64_t> host;
  rc = upload_points(e, pts, centrals, stride_bytes, n);
  if (rc == SJCS_OK) {
    hipError_t h = ensure(ucounts, nc * sizeof(uint64_t));
    if (h == hipSuccess) h = hipMemsetAsync(ucounts.p, 0, nc * sizeof(uint64_t), e->stream);
    if (h != hipSuccess) rc = rt_hip_fail(rt, h, "R-tree count");
  }
  if (rc == SJCS_OK)
    rc = rt_count_into(rt, pts.as<float>(), n, radiuses_sq, nrads, ucounts.as<uint64_t>());
  if (rc == SJCS_OK) {
    hipError_t h = hipSuccess;
    try {
      host.resize(nc);
    } catch (const std::bad_alloc&) {
      h = hipErrorOutOfMemory;
    }
    if (h == hipSuccess)
      h = hipMemcpy(host.data(), ucounts.p, nc * sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (h != hipSuccess) rc = rt_hip_fail(rt, h, "R-tree counts to the host");
  }
  if (rc == SJCS_OK)
    for (size_t i = 0; i < nc; ++i) counts[i] += host[i];
  pts.release();
  ucounts.release();
  if (rc != SJCS_OK) rt->err = e->err;  // the engine's helpers report there
  return rc;
}

int sjcs_rtree_count_device(sjcs_rtree* rt, const float* d_centrals, uint64_t n,
                            const float* radiuses_sq, int nrads, uint64_t* d_counts) {
  if (!rt) return SJCS_ERR_INVALID;
  if ((!d_centrals && n) || (!d_counts && n) || !radiuses_sq)
    return rt_fail(rt, SJCS_ERR_INVALID, "null pointer");
  int rc = rt_check(rt, nrads);
  if (rc != SJCS_OK) return rc;
  sjcs_engine* e = rt->e;
  rc = use_device(e);
  if (rc == SJCS_OK) rc = quiesce(e);  // the sort workspace is shared with the engine
  if (rc == SJCS_OK) rc = rt_upload(rt);
  if (rc == SJCS_OK && n) rc = rt_count_into(rt, d_centrals, n, radiuses_sq, nrads, d_counts);
  if (rc != SJCS_OK) rt->err = e->err;
  return rc;
}

void sjcs_rtree_destroy(sjcs_rtree* rt) {
  if (!rt) return;
  if (rt->e) {
    (void)hipSetDevice(rt->e->device);
    (void)hipStreamSynchronize(rt->e->stream);
  }
  rt->release();
  delete rt;
}

}  // extern "C"
