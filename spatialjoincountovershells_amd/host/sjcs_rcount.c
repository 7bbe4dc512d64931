/*
 * sjcs_rcount.c -- drop-in replacement for the reference's astr_rcount
 * (src/astr_rcount/main.c:445-672) that counts on an MI355X.
 *
 * Same argv, same stdout lines, same stderr messages and exit status for
 * bad options, same text result.  What runs differs: the halos, shell
 * thresholds and u64 counts live on the GPU (include/sjcs.h); each particle
 * file is indexed (cell sort) and counted there.  --backlog N keeps up to N
 * particle files on the device ahead of the one being counted (the
 * reference's background index builds, main.c:266-350; "index n" and
 * "index getting time" lines as it prints them, in file order); with
 * --backlog 0 a file is read after the previous one is counted.
 *
 * Host side, in order of the wall time it saves:
 *  - the device is set up (runtime start, engine, pinned staging) on a
 *    thread of its own while the halo file is parsed;
 *  - a particle file goes from disk to the device in 64 MiB pieces, each
 *    read by parallel preads into page-locked memory and copied from there
 *    while the next piece is read;
 *  - the counts come back in batches of rows that the result writer asks
 *    for while its workers format the previous batch.
 * The per-file "particles reading time" line reports disk to device memory,
 * "index construction" the cell sort, "searching" the count kernel, the
 * last two from HIP events; the index buffers are reused from file to file,
 * so "index destruction" is ~0.  With --compressed, each file's uint16
 * triplets go to the device as they are and are decoded there ("particles
 * decoding + index construction" = decode + cell sort).
 *
 * Run under a name that starts with "rt_rcount" (a link named rt_rcount),
 * it is the drop-in for the reference's rt_rcount instead (src/rt_rcount/
 * main.c, the same argv, stdout lines and result format): each particle
 * file is inserted point by point into rt_rcount's R-tree on the host
 * (include/sjcs.h, sjcs_rtree_*), and the halos are counted against it on
 * the GPU with rt_rcount's leaf walk, whose counts differ from astr_rcount's
 * for halos that sit exactly on a particle.  Files are handled one after the
 * other with either --backlog setting.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <unistd.h>

#include "sjcs.h"
#include "sjcs_host.h"

#define TIMER_START(tv) gettimeofday(&(tv), NULL)
#define PIECE_BYTES ((uint64_t)64 << 20)

/* SJCS_TRACE=1: milliseconds since start at each step, on stderr (the
 * reference's timer lines leave process start-up, device set-up, count
 * read-back and clean-up out). */
static struct timeval g_t0;
static int g_trace;

static void trace(const char *what)
{
    struct timeval b;
    if (!g_trace)
        return;
    gettimeofday(&b, NULL);
    fprintf(stderr, "[sjcs trace] %8.1f ms  %s\n",
            (b.tv_sec - g_t0.tv_sec) * 1e3 + (b.tv_usec - g_t0.tv_usec) * 1e-3, what);
}

static void print_ms(const char *what, struct timeval *a)
{
    struct timeval b;
    gettimeofday(&b, NULL);
    printf("%s: %ld milli seconds\n", what, sjcs_elapsed_ms(a, &b));
}

static void die_engine(sjcs_engine *e, const char *what)
{
    fprintf(stderr, "ERROR: %s: %s\n", what, sjcs_last_error(e));
    exit(EXIT_FAILURE);
}

/* ---- device set-up, overlapped with the halo parse ---- */

typedef struct chunk_buf chunk_buf;
static int read_chunk(sjcs_engine *e, void *pin[2], const char *fname, chunk_buf *c,
                      int compressed, int nthreads, char *err, size_t errlen);

typedef struct {
    int device;
    sjcs_engine *e;
    void *pin[2];  /* page-locked pieces of a particle file */
    int rc;
    /* the first particle file, read while the halos are parsed */
    const char *first;
    chunk_buf *buf;
    int compressed, nthreads;
    int read_rc;
    char err[4096 + 256];
} setup_job;

static void *setup_run(void *arg)
{
    setup_job *s = arg;
    s->rc = sjcs_engine_create(s->device, &s->e);
    trace("  (set-up thread) engine created");
    if (s->rc == SJCS_OK)
        s->rc = sjcs_host_alloc(PIECE_BYTES, &s->pin[0]);
    if (s->rc == SJCS_OK)
        s->rc = sjcs_host_alloc(PIECE_BYTES, &s->pin[1]);
    trace("  (set-up thread) page-locked pieces");
    if (s->rc == SJCS_OK && s->first)
        s->read_rc = read_chunk(s->e, s->pin, s->first, s->buf, s->compressed, s->nthreads,
                                s->err, sizeof(s->err));
    trace("  (set-up thread) first particle file on the device");
    return NULL;
}

/* ---- particle files: disk -> device ---- */

struct chunk_buf {
    void *d;               /* device buffer */
    uint64_t cap;          /* its bytes */
    uint64_t n;            /* particles of the file in it */
    sjcs_gadget_decode dec;
    long read_ms;
};

typedef struct {
    sjcs_engine *e;
    char *dst;
    const void *src;
    size_t len;
    int rc;
} piece_copy;

static void *piece_copy_run(void *arg)
{
    piece_copy *c = arg;
    c->rc = sjcs_copy_to_device(c->e, c->dst, c->src, c->len);
    return NULL;
}

/* Bytes [off, off + bytes) of fname into device memory at d, through the
 * two pinned pieces: piece i is copied to the device on a helper thread
 * while piece i + 1 is read into the other one.  0, the sjcs_pread_par
 * status, or -3 (copy failed). */
static int file_to_device(sjcs_engine *e, void *pin[2], const char *fname, uint64_t off,
                          uint64_t bytes, char *d, int nthreads)
{
    uint64_t done = 0;
    int k = 0, copying = 0, rc = 0;
    pthread_t th;
    piece_copy job;
    while (done < bytes) {
        uint64_t len = bytes - done < PIECE_BYTES ? bytes - done : PIECE_BYTES;
        rc = sjcs_pread_par(fname, off + done, len, pin[k], nthreads);
        if (copying) {  /* the other piece's copy; pin[k ^ 1] is free after it */
            pthread_join(th, NULL);
            copying = 0;
            if (job.rc != SJCS_OK)
                return -3;
        }
        if (rc != 0)
            return rc;
        job.e = e;
        job.dst = d + done;
        job.src = pin[k];
        job.len = (size_t)len;
        job.rc = SJCS_OK;
        if (pthread_create(&th, NULL, piece_copy_run, &job) == 0)
            copying = 1;
        else if (sjcs_copy_to_device(e, d + done, pin[k], (size_t)len) != SJCS_OK)
            return -3;
        done += len;
        k ^= 1;
    }
    if (copying) {
        pthread_join(th, NULL);
        if (job.rc != SJCS_OK)
            return -3;
    }
    return 0;
}

/* Halo centres [beg, end) packed as FP32 xyz at dst, by several threads. */
typedef struct {
    const sjcs_halo *h;
    uint64_t beg, end;
    float *dst;
} centre_pack;

static void *centre_pack_run(void *arg)
{
    const centre_pack *c = arg;
    uint64_t i;
    for (i = c->beg; i < c->end; i++)
        memcpy(c->dst + 3 * (i - c->beg), c->h[i].point, 12);
    return NULL;
}

static void pack_centres_par(const sjcs_halo *h, uint64_t n, float *dst, int nthreads)
{
    pthread_t th[64];
    centre_pack job[64];
    int t, nt = nthreads < 1 ? 1 : (nthreads > 64 ? 64 : nthreads), started = 0;
    for (t = 0; t < nt; t++) {
        job[t].h = h;
        job[t].beg = n * (uint64_t)t / (uint64_t)nt;
        job[t].end = n * (uint64_t)(t + 1) / (uint64_t)nt;
        job[t].dst = dst + 3 * job[t].beg;
    }
    for (t = 1; t < nt; t++) {
        if (pthread_create(&th[t], NULL, centre_pack_run, &job[t]) != 0)
            break;
        started = t;
    }
    centre_pack_run(&job[0]);
    for (t = started + 1; t < nt; t++)  /* threads that could not be started */
        centre_pack_run(&job[t]);
    for (t = 1; t <= started; t++)
        pthread_join(th[t], NULL);
}

/* The halo centres into device memory at d (12 bytes each) through the two
 * pinned pieces: piece i is copied to the device on a helper thread while
 * piece i + 1 is packed into the other one.  0, or -3 (copy failed). */
static int centres_to_device(sjcs_engine *e, void *pin[2], const sjcs_halos *halos, char *d,
                             int nthreads)
{
    const uint64_t per = PIECE_BYTES / 12;
    uint64_t first = 0;
    int k = 0, copying = 0;
    pthread_t th;
    piece_copy job;
    while (first < halos->n) {
        const uint64_t cnt = halos->n - first < per ? halos->n - first : per;
        pack_centres_par(halos->halo + first, cnt, pin[k], nthreads);
        if (copying) {  /* the other piece's copy; pin[k ^ 1] is free after it */
            pthread_join(th, NULL);
            copying = 0;
            if (job.rc != SJCS_OK)
                return -3;
        }
        job.e = e;
        job.dst = d + 12 * first;
        job.src = pin[k];
        job.len = (size_t)(12 * cnt);
        job.rc = SJCS_OK;
        if (pthread_create(&th, NULL, piece_copy_run, &job) == 0)
            copying = 1;
        else if (sjcs_copy_to_device(e, job.dst, pin[k], job.len) != SJCS_OK)
            return -3;
        first += cnt;
        k ^= 1;
    }
    if (copying) {
        pthread_join(th, NULL);
        if (job.rc != SJCS_OK)
            return -3;
    }
    return 0;
}

static int ensure_device(sjcs_engine *e, chunk_buf *c, uint64_t bytes)
{
    if (bytes == 0)
        bytes = 16;
    if (c->cap >= bytes && c->d != NULL)
        return 0;
    if (c->d)
        sjcs_device_free(e, c->d);
    c->d = NULL;
    c->cap = 0;
    if (sjcs_device_alloc(e, (size_t)bytes, &c->d) != SJCS_OK)
        return -1;
    c->cap = bytes;
    return 0;
}

/* Reads particle file fname into c's device buffer.  0, or -1 with the
 * message the reference's failure would print in err (src/Gadget/
 * BufVector3D.c:111-136, 285-308). */
static int read_chunk(sjcs_engine *e, void *pin[2], const char *fname, chunk_buf *c,
                      int compressed, int nthreads, char *err, size_t errlen)
{
    struct timeval a, b;
    uint64_t n = 0, off, bytes;
    const char *fn = compressed ? "readCompressedParticles" : "readDecodedParticles";
    int rc;
    TIMER_START(a);
    if (compressed) {
        sjcs_gadget_header hdr;
        struct stat st;
        rc = sjcs_read_compressed_header(fname, &hdr);
        if (rc == -1) {
            snprintf(err, errlen, "ERROR: file %s not found in %s()\n", fname, fn);
            return -1;
        }
        n = rc == 0 ? (uint64_t)hdr.npart[1] : 0;
        if (rc != 0 || stat(fname, &st) != 0 ||
            (uint64_t)st.st_size < SJCS_GADGET_DATA_OFFSET + 6 * n) {
            snprintf(err, errlen, "ERROR: file shorter than its header says in %s()\n", fn);
            return -1;
        }
        /* one tile per side, as constructIndex* fixes it (main.c:744) */
        if (sjcs_gadget_decode_params(&hdr, 1, &c->dec) != SJCS_OK) {
            snprintf(err, errlen, "ERROR: bad particle totals in the header of %s\n", fname);
            return -1;
        }
        off = SJCS_GADGET_DATA_OFFSET;
        bytes = 6 * n;
    } else {
        if (sjcs_particle_count(fname, &n) != 0) {
            snprintf(err, errlen, "ERROR: file %s not found in %s()\n", fname, fn);
            return -1;
        }
        off = 0;
        bytes = 12 * n;
    }
    if (ensure_device(e, c, bytes) != 0) {
        snprintf(err, errlen, "ERROR: device memory for %s: %s\n", fname, sjcs_last_error(e));
        return -1;
    }
    rc = file_to_device(e, pin, fname, off, bytes, c->d, nthreads);
    if (rc == -1) {
        snprintf(err, errlen, "ERROR: file %s not found in %s()\n", fname, fn);
        return -1;
    }
    if (rc == -3) {
        snprintf(err, errlen, "ERROR: particle upload: %s\n", sjcs_last_error(e));
        return -1;
    }
    if (rc != 0) {
        snprintf(err, errlen, "ERROR: while reading %s in %s()\n", fname, fn);
        return -1;
    }
    c->n = n;
    gettimeofday(&b, NULL);
    c->read_ms = sjcs_elapsed_ms(&a, &b);
    return 0;
}

static void read_chunk_or_die(sjcs_engine *e, void *pin[2], const char *fname, chunk_buf *c,
                              int compressed, int nthreads)
{
    char err[4096 + 256];
    if (read_chunk(e, pin, fname, c, compressed, nthreads, err, sizeof(err)) != 0) {
        fputs(err, stderr);
        exit(EXIT_FAILURE);
    }
}

/* Waits for the chunk in flight and prints its lines. */
static void finish_chunk(sjcs_engine *e, uint64_t n, int compressed)
{
    float ms[3];
    struct timeval a;
    if (sjcs_last_chunk_timings(e, ms) != SJCS_OK)
        die_engine(e, "search");
    printf(">>>> %s time: %ld milli seconds\n",
           compressed ? "particles decoding + index construction" : "index construction",
           (long)(ms[0] + ms[1]));
    printf("  # of particles = %llu\n", (unsigned long long)n);
    printf(">>>> searching time: %ld milli seconds\n", (long)ms[2]);
    TIMER_START(a);
    print_ms(">>>> index destruction time", &a);
}

/* ---- rt_rcount mode ---- */

static int rt_named(const char *argv0)
{
    const char *b = strrchr(argv0, '/');
    b = b ? b + 1 : argv0;
    return strncmp(b, "rt_rcount", 9) == 0;
}

static void *grow_host(void *p, uint64_t *cap, uint64_t bytes)
{
    void *q;
    if (bytes <= *cap)
        return p;
    q = realloc(p, (size_t)(bytes ? bytes : 16));
    if (!q) {
        fprintf(stderr, "ERROR: memory shortage\n");
        exit(EXIT_FAILURE);
    }
    *cap = bytes;
    return q;
}

/* The chunk loop of rt_rcount (src/rt_rcount/main.c:200-226, 618-673):
 * per file, its points on the host (a compressed file is decoded on the
 * device and copied back), an R-tree built by inserting them in order, the
 * halos counted against it on the GPU (counts += ...), the tree dropped. */
static void rt_chunks(sjcs_engine *e, void *pin[2], const sjcs_cli *cli, char **argv,
                      const sjcs_halos *halos, const float *r2, uint64_t *counts, int nthreads)
{
    const int nfiles = cli->particle_end - cli->particle_beg;
    const char *fn = cli->compressed ? "readCompressedParticles" : "readDecodedParticles";
    chunk_buf cb;
    float *xyz = NULL;
    uint64_t cap = 0, dcap = 0;
    void *d_xyz = NULL;
    int i;
    memset(&cb, 0, sizeof(cb));
    for (i = 0; i < nfiles; i++) {
        const char *fname = argv[cli->particle_beg + i];
        struct timeval a, b;
        sjcs_rtree *rt = NULL;
        uint64_t n = 0;
        long read_ms;
        int rc;
        TIMER_START(a);
        if (cli->compressed) {
            read_chunk_or_die(e, pin, fname, &cb, 1, nthreads);
            n = cb.n;
            read_ms = cb.read_ms;
        } else {
            if (sjcs_particle_count(fname, &n) != 0) {
                fprintf(stderr, "ERROR: file %s not found in %s()\n", fname, fn);
                exit(EXIT_FAILURE);
            }
            xyz = grow_host(xyz, &cap, 12 * n);
            rc = n ? sjcs_pread_par(fname, 0, 12 * n, xyz, nthreads) : 0;
            if (rc == -1) {
                fprintf(stderr, "ERROR: file %s not found in %s()\n", fname, fn);
                exit(EXIT_FAILURE);
            }
            if (rc != 0) {
                fprintf(stderr, "ERROR: while reading %s in %s()\n", fname, fn);
                exit(EXIT_FAILURE);
            }
            gettimeofday(&b, NULL);
            read_ms = sjcs_elapsed_ms(&a, &b);
        }
        printf(">>>> particles reading time: %ld milli seconds\n", read_ms);
        TIMER_START(a);
        if (cli->compressed && n) {
            if (dcap < 12 * n) {
                if (d_xyz)
                    sjcs_device_free(e, d_xyz);
                d_xyz = NULL;
                dcap = 0;
                if (sjcs_device_alloc(e, (size_t)(12 * n), &d_xyz) != SJCS_OK)
                    die_engine(e, "device memory for the decode");
                dcap = 12 * n;
            }
            xyz = grow_host(xyz, &cap, 12 * n);
            if (sjcs_decode_compressed_device(e, cb.d, n, &cb.dec, d_xyz) != SJCS_OK ||
                sjcs_copy_to_host(e, xyz, d_xyz, (size_t)(12 * n)) != SJCS_OK)
                die_engine(e, "particle decode");
        }
        if (sjcs_rtree_create(e, &rt) != SJCS_OK)
            die_engine(e, "R-tree");
        if (sjcs_rtree_insert(rt, xyz, n) != SJCS_OK) {
            fprintf(stderr, "ERROR: R-tree: %s\n", sjcs_rtree_error(rt));
            exit(EXIT_FAILURE);
        }
        print_ms(cli->compressed ? ">>>> particles decoding + index construction time"
                                 : ">>>> index construction time", &a);
        printf("  # of particles = %llu\n", (unsigned long long)n);
        if (cli->maxbacklogs > 0)
            printf(">>>> index getting time: %ld milli seconds\n", 0L);
        TIMER_START(a);
        if (sjcs_rtree_count(rt, halos->n ? halos->halo[0].point : NULL, sizeof(sjcs_halo),
                             halos->n, r2, cli->nrads, counts) != SJCS_OK) {
            fprintf(stderr, "ERROR: search: %s\n", sjcs_rtree_error(rt));
            exit(EXIT_FAILURE);
        }
        print_ms(">>>> searching time", &a);
        TIMER_START(a);
        sjcs_rtree_destroy(rt);
        print_ms(">>>> index destruction time", &a);
        trace("file counted (R-tree)");
        fflush(stdout);
    }
    if (cb.d)
        sjcs_device_free(e, cb.d);
    if (d_xyz)
        sjcs_device_free(e, d_xyz);
    free(xyz);
}

/* ---- counts -> result writer ---- */

typedef struct {
    sjcs_engine *e;
    void *rows[2];  /* page-locked, one batch of rows each */
    int slot;
} row_source;

static const uint64_t *fetch_rows(void *ctx, uint64_t beg, uint64_t n)
{
    row_source *r = ctx;
    r->slot ^= 1;
    if (sjcs_get_counts_range(r->e, beg, n, r->rows[r->slot]) != SJCS_OK) {
        fprintf(stderr, "ERROR: counts: %s\n", sjcs_last_error(r->e));
        return NULL;
    }
    return r->rows[r->slot];
}

static const uint32_t *fetch_rows32(void *ctx, uint64_t beg, uint64_t n)
{
    row_source *r = ctx;
    r->slot ^= 1;
    if (sjcs_get_counts_range_u32(r->e, beg, n, r->rows[r->slot]) != SJCS_OK) {
        fprintf(stderr, "ERROR: counts: %s\n", sjcs_last_error(r->e));
        return NULL;
    }
    return r->rows[r->slot];
}

int main(int argc, char *argv[])
{
    sjcs_cli cli;
    sjcs_halos halos = {NULL, 0};
    sjcs_engine *e = NULL;
    struct timeval tv, tvtotal;
    char err[4096];
    float *rad, *r2;
    chunk_buf *bufs;
    int nbufs, depth, ahead, next;
    setup_job setup;
    pthread_t setup_th;
    row_source src;
    int nfiles, i, hthreads, rc, setup_threaded, rt_mode;
    const char *dev_env = getenv("SJCS_DEVICE");
    const char *trace_env = getenv("SJCS_TRACE");

    gettimeofday(&g_t0, NULL);
    g_trace = trace_env != NULL && trace_env[0] == '1';
    rt_mode = rt_named(argv[0]);
    sjcs_cli_parse(argc, argv, &cli);
    sjcs_cli_print_info(stdout, &cli, argv);
    if (cli.errflag) {
        fprintf(stderr, "\n => %s\n", cli.errmsg);
        sjcs_cli_print_usage(stderr, argv[0]);
        exit(EXIT_FAILURE);
    }
    /* --backlog (main.c:266-350): the reference builds up to maxbacklogs
     * indexes in the background while it searches.  Here that many particle
     * files are held on the device ahead of the one being counted (and the
     * next file's cell sort may then run while this one counts); 0 reads
     * each file only after the previous one is counted, as
     * noBacklogExecution builds each index after the previous search.  The
     * first file is read while the halos are parsed either way. */
    depth = cli.maxbacklogs > 0 ? cli.maxbacklogs : 0;
    nbufs = depth + 1;
    bufs = calloc((size_t)nbufs, sizeof(chunk_buf));
    if (!bufs) {
        fprintf(stderr, "ERROR: memory shortage\n");
        exit(EXIT_FAILURE);
    }
    memset(&setup, 0, sizeof(setup));
    hthreads = cli.nthreads < 256 ? cli.nthreads : 256;
    setup.device = dev_env ? atoi(dev_env) : 0;
    /* rt mode reads its files into host memory, one at a time */
    setup.first = rt_mode ? NULL : argv[cli.particle_beg];
    setup.buf = &bufs[0];
    setup.compressed = cli.compressed;
    /* half the threads: the halo parse runs meanwhile */
    setup.nthreads = hthreads / 2 > 0 ? hthreads / 2 : 1;
    setup_threaded = pthread_create(&setup_th, NULL, setup_run, &setup) == 0;
    if (!setup_threaded)
        setup_run(&setup);

    /* halos (main.c:634-642); one thread fewer than the host threads, so
     * the device set-up (runtime start, code load) is not starved meanwhile */
    TIMER_START(tv);
    if (cli.filter_file == NULL)
        rc = sjcs_read_halos(cli.halo_file, &halos, hthreads > 1 ? hthreads - 1 : 1, err,
                             sizeof(err));
    else
        rc = sjcs_read_and_filter_halos(cli.halo_file, cli.filter_file, &halos, err, sizeof(err));
    if (rc != 0) {
        /* the device set-up finishes first: exiting under its runtime calls
         * would race the runtime's teardown */
        if (setup_threaded)
            pthread_join(setup_th, NULL);
        fputs(err, stderr);
        exit(EXIT_FAILURE);
    }
    print_ms(">>>> halo data reading time", &tv);
    printf("  # of halos = %lu\n", (unsigned long)halos.n);
    trace("halos read");

    /* shells (main.c:645-649) */
    rad = malloc(sizeof(float) * (size_t)cli.nrads);
    r2 = malloc(sizeof(float) * (size_t)cli.nrads);
    if (setup_threaded)
        pthread_join(setup_th, NULL);
    if (!rad || !r2) {
        fprintf(stderr, "ERROR: memory shortage\n");
        exit(EXIT_FAILURE);
    }
    sjcs_host_shell_thresholds(cli.minrad, cli.maxrad, cli.nrads, rad, r2);

    e = setup.e;
    if (setup.rc != SJCS_OK)
        die_engine(e, "GPU engine");
    trace("device set up");
    if (rt_mode) {
        uint64_t *counts = calloc(halos.n ? (size_t)halos.n * (size_t)cli.nrads : 1,
                                  sizeof(uint64_t));
        if (!counts) {
            fprintf(stderr, "ERROR: memory shortage\n");
            exit(EXIT_FAILURE);
        }
        TIMER_START(tvtotal);
        puts(cli.maxbacklogs == 0 ? "  -- with no backlog execution"
                                  : "  -- with backlog execution");
        fflush(stdout);
        rt_chunks(e, setup.pin, &cli, argv, &halos, r2, counts, hthreads);
        print_ms(">>>> total searching time", &tvtotal);
        TIMER_START(tv);
        if (sjcs_write_result(cli.outfile, &halos, counts, r2, cli.nrads, hthreads) < 0)
            exit(EXIT_FAILURE);
        print_ms(">>>> result writing time", &tv);
        free(counts);
        free(bufs);
        sjcs_host_free(setup.pin[0]);
        sjcs_host_free(setup.pin[1]);
        sjcs_engine_destroy(e);
        sjcs_free_halos(&halos);
        free(rad);
        free(r2);
        return 0;
    }
    if (sjcs_set_shells(e, r2, cli.nrads) != SJCS_OK)
        die_engine(e, "shells");
    trace("shells on the device");
    /* the centres go up packed (12 of the records' 24 bytes) through the
     * page-locked pieces: a copy of the records from pageable memory runs
     * at a fraction of the link's rate */
    {
        void *d_centres = NULL;
        if (halos.n && (sjcs_device_alloc(e, 12 * (size_t)halos.n, &d_centres) != SJCS_OK ||
                        centres_to_device(e, setup.pin, &halos, d_centres, hthreads) != 0))
            die_engine(e, "halos");
        if (sjcs_set_halos_device(e, d_centres, halos.n) != SJCS_OK)
            die_engine(e, "halos");
        if (d_centres)
            sjcs_device_free(e, d_centres);
    }
    trace("shells and halos on the device");

    /* chunk loop (main.c:235-368) */
    nfiles = cli.particle_end - cli.particle_beg;
    TIMER_START(tvtotal);
    puts(cli.maxbacklogs == 0 ? "  -- with no backlog execution" : "  -- with backlog execution");
    fflush(stdout);
    if (setup.read_rc != 0) {
        fputs(setup.err, stderr);
        exit(EXIT_FAILURE);
    }
    if (depth > 0 && sjcs_set_option(e, "overlap", 1) != SJCS_OK)
        die_engine(e, "options");
    next = 1;  /* the next file to read */
    if (depth > 0) {
        /* --first-setup a:b: a files read before the first count */
        ahead = cli.first_nbacklogs < depth ? cli.first_nbacklogs : depth;
        printf("index 0\n>>>> particles reading time: %ld milli seconds\n", bufs[0].read_ms);
        for (; next < nfiles && next <= ahead; next++) {
            printf("index %d\n", next);
            read_chunk_or_die(e, setup.pin, argv[cli.particle_beg + next], &bufs[next % nbufs],
                              cli.compressed, hthreads);
            printf(">>>> particles reading time: %ld milli seconds\n", bufs[next % nbufs].read_ms);
        }
    }
    for (i = 0; i < nfiles; i++) {
        chunk_buf *cur = &bufs[i % nbufs];
        if (depth > 0) {
            printf(">>>> index getting time: %ld milli seconds\n", 0L);
        } else {
            if (i > 0)
                read_chunk_or_die(e, setup.pin, argv[cli.particle_beg + i], cur, cli.compressed,
                                  hthreads);
            printf(">>>> particles reading time: %ld milli seconds\n", cur->read_ms);
        }
        if (cur->n == 0)
            fprintf(stderr, "%s(): No data loaded. Cannot make tree.\n",
                    "MakeTreeWithSameThreadPoolArraySTR");
        rc = cli.compressed
                 ? sjcs_add_compressed_particles_device(e, cur->d, cur->n, &cur->dec)
                 : sjcs_add_particles_device(e, cur->d, cur->n);
        if (rc != SJCS_OK)
            die_engine(e, "particle chunk");
        trace("chunk queued");
        /* up to `depth` files ahead while this one is sorted and counted;
         * a file's buffer was last read by the chunk `depth` + 1 earlier,
         * which is done */
        for (; depth > 0 && next < nfiles && next <= i + depth; next++) {
            printf("index %d\n", next);
            read_chunk_or_die(e, setup.pin, argv[cli.particle_beg + next], &bufs[next % nbufs],
                              cli.compressed, hthreads);
            printf(">>>> particles reading time: %ld milli seconds\n", bufs[next % nbufs].read_ms);
            trace("file ahead on the device");
        }
        finish_chunk(e, cur->n, cli.compressed);
        trace("chunk counted");
        fflush(stdout);
    }
    print_ms(">>>> total searching time", &tvtotal);

    /* result (main.c:660-664): rows come back while earlier rows are
     * formatted; as u32 (half the bytes over the link) while the totals are
     * still the device's 32-bit running counts */
    TIMER_START(tv);
    src.e = e;
    src.slot = 1;
    src.rows[0] = src.rows[1] = NULL;
    {
        const int narrow = sjcs_get_counts_range_u32(e, 0, 0, NULL) == SJCS_OK;
        const size_t rb = (narrow ? sizeof(uint32_t) : sizeof(uint64_t)) * SJCS_RESULT_BATCH *
                          (size_t)cli.nrads;
        if (sjcs_host_alloc(rb, &src.rows[0]) != SJCS_OK ||
            sjcs_host_alloc(rb, &src.rows[1]) != SJCS_OK)
            die_engine(NULL, "counts staging");
        if ((narrow ? sjcs_write_result_rows32(cli.outfile, &halos, fetch_rows32, &src, r2,
                                               cli.nrads, hthreads)
                    : sjcs_write_result_rows(cli.outfile, &halos, fetch_rows, &src, r2,
                                             cli.nrads, hthreads)) < 0)
            exit(EXIT_FAILURE);
    }
    print_ms(">>>> result writing time", &tv);
    trace("result written");

    sjcs_host_free(src.rows[0]);
    sjcs_host_free(src.rows[1]);
    for (i = 0; i < nbufs; i++)
        if (bufs[i].d)
            sjcs_device_free(e, bufs[i].d);
    free(bufs);
    sjcs_host_free(setup.pin[0]);
    sjcs_host_free(setup.pin[1]);
    sjcs_engine_destroy(e);
    sjcs_free_halos(&halos);
    free(rad);
    free(r2);
    trace("exit");
    return 0;
}
