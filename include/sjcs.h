/*
 * sjcs.h -- C-ABI of the MI355X spatial-join-count-over-shells engine.
 *
 * The reference (astr_rcount of ryumt/SpatialJoinCountOverShells) has no FFI:
 * its hot path sits behind the ArraySTR index API and the per-halo
 * multi-range count.  This header replaces those, one entry point each,
 * with plain pointers and sizes:
 *
 *   reference                                         here
 *   constructArraySTR        ArraySTR.c:16-34          sjcs_index_create
 *   insertDataArraySTR  x n  Build.c:702-714,          sjcs_index_insert
 *                            astr_rcount/main.c:774-778
 *   makeTree*ArraySTR        Build.c:716-963           sjcs_index_build
 *   threadJoin + multiRangeCountWithPriodicBoundArraySTR
 *                            astr_rcount/main.c:795-826,
 *                            ArraySTR/MultiRangeCount.c:13-73
 *                                                      sjcs_index_count
 *   destructArraySTR         ArraySTR.c:36-41          sjcs_index_destroy
 *
 * and adds the chunk engine the drop-in CLI drives: the halo set and shell
 * thresholds stay on the device, each particle file is indexed and counted
 * on arrival, and the u64 counts accumulate on the device across files
 * (the reference's per-file loop, astr_rcount/main.c:235-264, with its
 * counts calloc'ed once at main.c:651 and never reset between files).
 *
 * Conventions.
 *  - Every call returns SJCS_OK (0) or a negative SJCS_ERR_*; nothing
 *    exits the process.  sjcs_last_error() gives the message of the last
 *    failure on that engine (or of the last failed sjcs_engine_create when
 *    passed NULL).
 *  - Host buffers belong to the caller and are copied before the call
 *    returns; device buffers passed to the *_device variants must stay
 *    valid until sjcs_sync() returns.
 *  - Shell thresholds are the reference's mrcc->radiuses after squaring
 *    (astr_rcount/main.c:646-649): FP32, one per shell, normally ascending.
 *    A pair lands in the first shell k with dist < radiuses_sq[k]
 *    (ArraySTR/MultiRangeCount.c:107-113).
 *  - Counts are uint64, row-major [nhalos][nrads] (main.c:651,821).
 *  - One engine per host thread at a time; engines are independent.
 */
#ifndef SJCS_H
#define SJCS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SJCS_OK 0
#define SJCS_ERR_INVALID (-1) /* bad argument */
#define SJCS_ERR_DEVICE (-2)  /* HIP runtime or kernel failure */
#define SJCS_ERR_NOMEM (-3)   /* device or pinned host allocation failed */
#define SJCS_ERR_STATE (-4)   /* call out of order (e.g. no shells set) */

typedef struct sjcs_engine sjcs_engine;
typedef struct sjcs_index sjcs_index;

/* ---- engine: device context, shells, halos, device-side accumulation ---- */

/* Opens an engine on HIP device `device`.  No reference counterpart: the
 * reference's state lives in main()'s JoinContext (astr_rcount/main.c:48-60). */
int sjcs_engine_create(int device, sjcs_engine **out);
void sjcs_engine_destroy(sjcs_engine *e);
const char *sjcs_last_error(const sjcs_engine *e);

/* Shell thresholds (squared radii).  Replaces the threshold template
 * mrcc->radiuses (astr_rcount/main.c:645-649).  Resets the counts. */
int sjcs_set_shells(sjcs_engine *e, const float *radiuses_sq, int nrads);

/* Halo centres: nhalos points of 3 FP32 each, `stride_bytes` apart (24 for
 * the reference's Halo{u64 id; f32 mass; f32 point[3]} with point at byte
 * 12, include/Gadget/BufHalo.h:26-30; 12 for packed xyz).  Replaces the
 * BufHalo handed to threadJoin (main.c:635-642).  Resets the counts. */
int sjcs_set_halos(sjcs_engine *e, const float *points, size_t stride_bytes,
                   uint64_t nhalos);
int sjcs_set_halos_device(sjcs_engine *e, const float *d_xyz, uint64_t nhalos);

/* Zeroes the device-side counts. */
int sjcs_reset_counts(sjcs_engine *e);

/* One particle chunk (packed FP32 xyz, the decoded particle file format,
 * src/Gadget/BufVector3D.c:285-308): builds the cell index and adds every
 * halo's shell counts for these particles to the device-side totals.
 * Replaces one iteration of noBacklogExecution (main.c:246-262):
 * constructIndexWithSameThreadPool + threadJoin fan-out + destructIndex.
 * Asynchronous: returns after the upload has been queued. */
int sjcs_add_particles(sjcs_engine *e, const float *xyz, uint64_t n);
int sjcs_add_particles_device(sjcs_engine *e, const float *d_xyz, uint64_t n);

/* ---- compressed (Gadget) particle files ----
 * A compressed particle file (`--compressed`, src/Gadget/BufVector3D.c:
 * 111-136) holds a 4-byte record marker, the 256-byte header below, 8 bytes
 * of record markers, then npart[1] triplets of uint16 displacements.  The
 * file's particle n has id = id_start + n; the id names a point of a regular
 * grid (src/Gadget/GadId.c:70-86, 112-137) and the triplet displaces it
 * (decodeCompressedParticle, BufVector3D.c:168-201):
 *   pos[d] = cell_interval * ip[d] + disp_min + unitsep * uv[d]. */

/* The file header, byte for byte (include/Gadget/GadgetHeader.h:4-27). */
typedef struct sjcs_gadget_header {
    int32_t npart[6];
    double mass[6];
    double time;
    double redshift;
    int32_t flag_sfr;
    int32_t flag_feedback;
    int32_t npartTotal[6];
    int32_t flag_cooling;
    int32_t num_files;
    double BoxSize;
    double Omega0;
    double OmegaLambda;
    double HubbleParam;
    int32_t flag_stellarage;
    int32_t flag_metals;
    int32_t hashsize;
    float disp_min;
    float disp_max;
    int32_t pos_bits;
    int64_t id_start;
    int64_t id_end;
    char fill[56];
} sjcs_gadget_header;

/* What the decode needs: GadgetDecodeParams + GadId
 * (include/Gadget/GadgetHeader.h:30-37, include/Gadget/GadId.h:5-14). */
typedef struct sjcs_gadget_decode {
    int64_t id_start;     /* id of the file's first particle */
    int64_t ng1d;         /* grid points per tile side (nparticles_partile) */
    int32_t nt1d;         /* tiles per side (astr_rcount uses 1) */
    int32_t fused;        /* 1: cell_interval*ip + disp_min is one FMA, as gcc
                             contracts it for an FMA-capable -march (the
                             reference's -march=native); 0: two roundings */
    float cell_interval;
    float disp_min;
    double unitsep;
} sjcs_gadget_decode;

/* initGadgetDecodeParams + initGadId (GadgetHeader.c:7-24, GadId.c:38-45,
 * astr_rcount/main.c:694-696) for `ntile` tiles per side, fused = 1. */
int sjcs_gadget_decode_params(const sjcs_gadget_header *hdr, int ntile,
                              sjcs_gadget_decode *out);

/* One compressed particle file's triplets (uv[3n], host) as one particle
 * chunk: uploaded as 6 bytes per particle, decoded on the device, then
 * indexed and counted as sjcs_add_particles does.  A file of one tile per
 * side is decoded by the index build's first pass as it reads the triplets
 * (no decoded copy is stored); other files, and the packed-key sort's input,
 * go through a decode into device staging first.  Replaces the compressed
 * branch of constructIndexWithSameThreadPool (astr_rcount/main.c:739-763).
 * SJCS_ERR_INVALID when an id of the file is out of the range the decode
 * handles exactly (grid points beyond +-2^31, |id| >= 2^50). */
int sjcs_add_compressed_particles(sjcs_engine *e, const uint16_t *uv, uint64_t n,
                                  const sjcs_gadget_decode *p);
/* The same with the triplets already on the device (d_uv, 3n uint16); the
 * upload phase of sjcs_last_chunk_timings is then empty, and a decode that
 * runs before the build (files of several tiles per side) is in no phase. */
int sjcs_add_compressed_particles_device(sjcs_engine *e, const uint16_t *d_uv, uint64_t n,
                                         const sjcs_gadget_decode *p);

/* Decodes n triplets at d_uv (device) into packed xyz at d_xyz (device);
 * returns when done.  The decode alone, for callers that keep the points. */
int sjcs_decode_compressed_device(sjcs_engine *e, const uint16_t *d_uv, uint64_t n,
                                  const sjcs_gadget_decode *p, float *d_xyz);

/* Waits for all queued work of this engine. */
int sjcs_sync(sjcs_engine *e);

/* Copies the accumulated counts [nhalos][nrads] to `counts` (host).
 * Synchronises first. */
int sjcs_get_counts(sjcs_engine *e, uint64_t *counts);
/* The rows of halos [first, first + nhalos) only, into counts[nhalos][nrads]:
 * a result writer can fetch the next rows while it formats these.  The rows
 * are put back in input order once per state of the totals.  Both calls copy
 * more than 64 MiB through the engine's page-locked staging in pieces,
 * several host threads moving each piece into `counts` (any host memory). */
int sjcs_get_counts_range(sjcs_engine *e, uint64_t first, uint64_t nhalos, uint64_t *counts);
/* The same rows as uint32, half the bytes over the link, while every total
 * is still one of the device's 32-bit running counts -- no fold into 64
 * bits has happened since the last reset (sjcs_reset_counts, sjcs_set_*):
 * the values are then exact.  Otherwise SJCS_ERR_STATE (take
 * sjcs_get_counts_range); nhalos = 0 asks which case holds.  No reference
 * counterpart: the read-back of a result writer that formats 32-bit rows. */
int sjcs_get_counts_range_u32(sjcs_engine *e, uint64_t first, uint64_t nhalos,
                              uint32_t *counts);

/* Exact number of in-box (halo image, particle) pair tests the last chunk
 * made -- the reference's CalcDistance calls (MultiRangeCount.c:107).  Runs
 * one extra untimed pass; a measurement aid, not part of the count.  When
 * "piece_max" (or the 2^30-1 particle limit of one piece) cuts a chunk into
 * pieces, this is the last piece's figure. */
int sjcs_last_chunk_pair_tests(sjcs_engine *e, uint64_t *pair_tests);

/* Tuning knobs (no reference counterpart).  Keys:
 *   "cell_size_milli"   cell edge along y and z in 1/1000 box units,
 *                       0 = automatic
 *   "x_split"           cells along x per cell along y/z (1 to 16),
 *                       0 = automatic
 *   "kernel"            0 = automatic, 1 = one halo per wavefront (plain)
 *   "lanes_per_halo"    8, 16 or 64 lanes of a wavefront per halo, 0 =
 *                       automatic (16 up to 256 shells)
 *   "segment"           particles of a cell row one lane walks: 4 or 8,
 *                       0 = automatic
 *   "tiles"             1 (default) = halos sharing a Morton cell (runs of
 *                       at least 4) are counted as tiles that share their
 *                       candidate particles; 0 = one halo at a time
 *   "graph"             1 = each chunk's sort + count is captured once as
 *                       a HIP graph and replayed (timings then give the
 *                       pair as the count phase, sort phase 0)
 *   "fold_limit"        bound on a running u32 count before it is folded
 *                       into the u64 totals (0 = 2^32-1; small values
 *                       exercise the folds)
 *   "overlap"           1 = sort chunk i+1 while chunk i is counted
 *                       (0, default: the upload still overlaps, the sort
 *                       waits)
 *   "index_sort"        0 = automatic (axis by axis up to 256 cells per
 *                       axis), 1 = always by the packed cell key
 *   "piece_max"         most particles of a chunk indexed and counted as
 *                       one piece (0 = the most that 32-bit offsets and
 *                       counts allow; small values exercise the split)
 *   "tile_level"        0 (default) = a tile is a run of at least 4 halos
 *                       in one cell of the halos' 1024^3 Morton grid (about
 *                       one unit); L = 1..6: a run of at least 2 halos in
 *                       one cell of 2^L grid cells per side, so neighbouring
 *                       halos share their candidates too.  Counts already
 *                       accumulated are kept.
 *   "tile_halos"        most halos per tile, 2 to 64 (0 = as many as the
 *                       tile kernel's LDS histograms hold: 2048 / nrads);
 *                       smaller tiles need less LDS per wavefront
 *   "stage_particles"   0 (default) = the count kernel reads candidates
 *                       from HBM/L2; 256 to 8192 = each workgroup of 8 or 16
 *                       lanes per halo first copies the particles of its
 *                       halos' union box (at most this many) into LDS and
 *                       reads the base images' candidates from there
 *   "stage_halos"       halos per such workgroup (0 = 64; 4 to 1024)
 *   "autotune_cache"    1 (default) = sjcs_autotune reuses the setting it
 *                       found for the same workload in this process
 * sjcs_get_option also reads "autotune_cache_hit": 1 when the last
 * sjcs_autotune took its setting from that cache.
 * Returns SJCS_ERR_INVALID for an unknown key or value. */
int sjcs_set_option(sjcs_engine *e, const char *key, int64_t value);

/* Current value of an option above (0 = automatic where that applies). */
int sjcs_get_option(sjcs_engine *e, const char *key, int64_t *value);

/* Autotuner (no reference counterpart): times the cell sort + count of the
 * chunk at d_xyz (device, n packed points) against the current halos and
 * shells under candidate settings -- lanes per halo, cell edge along y/z as
 * a fraction of r_box, x split, segment, then halo tiles on or off, tile
 * level and halos per tile, then the particles staged in LDS per count
 * workgroup and its halos, by coordinate descent -- and keeps the fastest
 * as the engine's options (sjcs_get_option reads them).  The setting is
 * remembered for the workload -- particle and halo counts, the thresholds,
 * and histograms of how particles and halos fill the box -- and reused
 * without timing when the same workload is tuned again in this process
 * ("autotune_cache").  The counts are not touched; the last chunk's index
 * is discarded.  *best_ms (may be NULL) receives the winner's sort + count
 * time. */
int sjcs_autotune(sjcs_engine *e, const float *d_xyz, uint64_t n, float *best_ms);

/* Milliseconds the device spent on the last chunk, by phase (HIP events):
 * [0] upload, [1] index build (key + sort + cell table), [2] count.  For a
 * chunk cut into pieces, [1] and [2] are the last piece's. */
int sjcs_last_chunk_timings(sjcs_engine *e, float ms[3]);

/* The three phases above, then [3] the last halo preparation on the device
 * (sjcs_set_halos*: Morton keys, their sort and the read-out's row map;
 * the tiling that follows on the host is not included) and [4] the last read-out of the counts into input row order
 * (the first sjcs_get_counts* after the totals changed).  No reference
 * counterpart: the reference's own timers are the CLI's >>>> lines. */
int sjcs_phase_timings(sjcs_engine *e, float ms[5]);

/* ---- Serving loop (no reference counterpart) ----
 * Requests -- a batch of halo centres with its own squared shell thresholds
 * -- counted against one built index.  A worker thread admits queued
 * requests in arrival order while one of `slots` streams is free and the
 * requests in flight fit `mem_budget` device bytes; admitted requests run
 * concurrently, and each result is independent of what else is in flight.
 * The index must outlive the server.  All calls are thread-safe. */
typedef struct sjcs_server sjcs_server;

/* slots: concurrent requests (0 = 8); mem_budget: bytes (0 = half the free
 * device memory at creation). */
int sjcs_server_create(sjcs_index *ix, int slots, uint64_t mem_budget, sjcs_server **out);

/* Queues a request (centres and thresholds are copied); its counts
 * [n][nrads] (u64, caller-owned, not read) are written when it completes.
 * *ticket names it for the calls below. */
int sjcs_server_submit(sjcs_server *sv, const float *centrals, size_t stride_bytes, uint64_t n,
                       const float *radiuses_sq, int nrads, uint64_t *counts, uint64_t *ticket);

/* 1 = done, 0 = queued or running, < 0 = the request failed. */
int sjcs_server_test(sjcs_server *sv, uint64_t ticket);

/* Waits until the request is done; SJCS_OK or its error. */
int sjcs_server_wait(sjcs_server *sv, uint64_t ticket);

/* Seconds since the server started at which the request was submitted,
 * admitted and done. */
int sjcs_server_times(sjcs_server *sv, uint64_t ticket, double *submitted, double *admitted,
                      double *done);

/* Forgets a finished request. */
int sjcs_server_release(sjcs_server *sv, uint64_t ticket);

/* Message of the server's last error. */
const char *sjcs_server_error(const sjcs_server *sv);

/* Finishes every queued and running request, then frees the server. */
void sjcs_server_destroy(sjcs_server *sv);

/* Device memory helpers for callers without a HIP runtime of their own
 * (benchmarks, bindings).  No reference counterpart. */
int sjcs_device_alloc(sjcs_engine *e, size_t bytes, void **d_ptr);
int sjcs_device_free(sjcs_engine *e, void *d_ptr);
int sjcs_copy_to_device(sjcs_engine *e, void *d_dst, const void *src, size_t bytes);
int sjcs_copy_to_host(sjcs_engine *e, void *dst, const void *d_src, size_t bytes);
/* Page-locked host memory (fast transfers in both directions); no engine
 * needed.  A failure's message is sjcs_last_error(NULL). */
int sjcs_host_alloc(size_t bytes, void **out);
void sjcs_host_free(void *p);

/* ---- index: the ArraySTR lifecycle, one device-resident particle index ---- */

/* ndata: at most 2^30-1 particles. */
int sjcs_index_create(sjcs_engine *e, uint64_t ndata, sjcs_index **out);
int sjcs_index_insert(sjcs_index *ix, const float *xyz, uint64_t n);
int sjcs_index_build(sjcs_index *ix);
/* counts[h*nrads + k] += shell-k count of centre h, for n centres given as
 * in sjcs_set_halos.  Uses its own shells; leaves the engine's untouched.
 * Centres are tiled as the engine's "tiles" / "tile_level" / "tile_halos"
 * options say (centres at one spot share their candidate particles). */
int sjcs_index_count(sjcs_index *ix, const float *centrals, size_t stride_bytes,
                     uint64_t n, const float *radiuses_sq, int nrads,
                     uint64_t *counts);
void sjcs_index_destroy(sjcs_index *ix);

/* sjcs_index_count without any device allocation.  No reference
 * counterpart: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  d_centrals: n packed FP32 xyz on the device; d_counts:
 * u64 [n][nrads] on the device, counts[h*nrads + k] += shell-k count;
 * radiuses_sq: host, at most 1024 shells.  Its working memory is a fixed
 * 32 MiB array in the library, one per device: n * (20 + 4*nrads) bytes
 * plus 12 KiB must fit (SJCS_ERR_INVALID otherwise).  Every call on a
 * device shares that array: calls that overlap on one device (two host
 * threads, two engines) race.  Returns when the counts are added. */
int sjcs_index_count_noalloc(sjcs_index *ix, const float *d_centrals, uint64_t n,
                             const float *radiuses_sq, int nrads, uint64_t *d_counts);

/* ---- R-tree: rt_rcount's index and count ----
 * rt_rcount counts with a Guttman R-tree whose leaf walk differs from
 * astr_rcount's: a leaf under an overlapping parent tests every entry's
 * distance with no box test, and an entry at distance exactly 0 from the
 * (moved) centre ends that leaf's scan -- neither it nor the entries after
 * it in the leaf are counted (src/objects/RTree/MultiRangeCount.c:86-114).
 * Which particles share a leaf depends on the insertion order, so the tree
 * is built as the reference builds it, one point at a time on the host, and
 * walked on the GPU:
 *
 *   reference                                         here
 *   constructRTree           RTree.c:13-25            sjcs_rtree_create
 *   insertDataRTree  x n     RTree.c:35-49,           sjcs_rtree_insert
 *                            rt_rcount/main.c:711-715
 *   threadJoin + multiRangeCountWithPriodicBoundRTree
 *                            rt_rcount/main.c:735-764,
 *                            RTree/MultiRangeCount.c:13-114
 *                                                      sjcs_rtree_count
 *   destructRTree            RTree.c:27-33            sjcs_rtree_destroy */
typedef struct sjcs_rtree sjcs_rtree;

/* e may be NULL: a host-only tree that can be built and inspected but not
 * counted with. */
int sjcs_rtree_create(sjcs_engine *e, sjcs_rtree **out);
/* Inserts n points (packed FP32 xyz, host) in order; their ids continue
 * from the points already inserted.  At most 2^29-1 points per tree.
 * SJCS_ERR_INVALID when the reference's split is undefined for the
 * coordinates (every enlargement difference NaN, where it reads outside its
 * entries); after a failed insert the tree refuses sjcs_rtree_leaves and
 * sjcs_rtree_count (SJCS_ERR_STATE), and sjcs_rtree_shape's npoints counts
 * the points inserted before the failure. */
int sjcs_rtree_insert(sjcs_rtree *rt, const float *xyz, uint64_t n);
/* Points, leaves, nodes and height (levels above the leaves). */
int sjcs_rtree_shape(const sjcs_rtree *rt, uint64_t *npoints, uint64_t *nleaves,
                     uint64_t *nnodes, int *height);
/* The leaf grouping: ids[npoints] leaf by leaf (leaves depth first, children
 * and entries in node order), leaf_len[nleaves].  No reference counterpart;
 * for inspection and tests. */
int sjcs_rtree_leaves(const sjcs_rtree *rt, uint64_t *ids, uint32_t *leaf_len);
/* counts[h*nrads + k] += rt_rcount's shell-k count of centre h, for n
 * centres given as in sjcs_set_halos, 1 to 1024 shells; on the engine's
 * GPU (the tree is copied there on the first count after an insert). */
int sjcs_rtree_count(sjcs_rtree *rt, const float *centrals, size_t stride_bytes, uint64_t n,
                     const float *radiuses_sq, int nrads, uint64_t *counts);
/* The same with the centres (n packed FP32 xyz) and the u64 counts
 * [n][nrads] on the device (d_counts += ...); radiuses_sq on the host.
 * Returns when the counts are added. */
int sjcs_rtree_count_device(sjcs_rtree *rt, const float *d_centrals, uint64_t n,
                            const float *radiuses_sq, int nrads, uint64_t *d_counts);
/* Message of the tree's last error (also sjcs_last_error of its engine). */
const char *sjcs_rtree_error(const sjcs_rtree *rt);
void sjcs_rtree_destroy(sjcs_rtree *rt);

#ifdef __cplusplus
}
#endif

#endif /* SJCS_H */
