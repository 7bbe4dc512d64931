/*
 * sjcs_rcount.c -- drop-in replacement for the reference's astr_rcount
 * (src/astr_rcount/main.c:445-672) that counts on an MI355X.
 *
 * Same argv, same stdout lines, same stderr messages and exit status for
 * bad options, same text result.  What runs differs: the halos, shell
 * thresholds and u64 counts live on the GPU (include/sjcs.h); each particle
 * file is indexed (cell sort) and counted there, and the next file is read
 * from disk while the GPU works on the current one.  The per-file
 * "index construction" line reports upload + cell sort, "searching" the
 * count kernel, both from HIP events; the index buffers are reused from
 * file to file, so "index destruction" is ~0.  With --compressed, each
 * file's uint16 triplets are uploaded as they are and decoded on the
 * device ("particles decoding + index construction" = upload + decode +
 * cell sort).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <unistd.h>

#include "sjcs.h"
#include "sjcs_host.h"

#define TIMER_START(tv) gettimeofday(&(tv), NULL)

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

typedef struct {
    float *xyz;
    uint16_t *uv;             /* --compressed: the file's triplets */
    sjcs_gadget_decode dec;   /* --compressed: their decode parameters */
    uint64_t n, cap;
    long read_ms;
} chunk_buf;

/* Reads compressed particle file fname into c (main.c:739-756). */
static void read_compressed_chunk(const char *fname, chunk_buf *c)
{
    struct timeval a, b;
    sjcs_gadget_header hdr;
    uint64_t n = 0;
    int rc;
    TIMER_START(a);
    rc = sjcs_read_compressed(fname, &hdr, &c->uv, &c->cap, &n);
    if (rc == -1) {
        fprintf(stderr, "ERROR: file %s not found in %s()\n", fname, "readCompressedParticles");
        exit(EXIT_FAILURE);
    }
    if (rc != 0) {
        fprintf(stderr, "ERROR: %s in %s()\n",
                rc == -3 ? "memory shortage" : "file shorter than its header says",
                "readCompressedParticles");
        exit(EXIT_FAILURE);
    }
    /* one tile per side, as constructIndex* fixes it (main.c:744) */
    if (sjcs_gadget_decode_params(&hdr, 1, &c->dec) != SJCS_OK) {
        fprintf(stderr, "ERROR: bad particle totals in the header of %s\n", fname);
        exit(EXIT_FAILURE);
    }
    c->n = n;
    gettimeofday(&b, NULL);
    c->read_ms = sjcs_elapsed_ms(&a, &b);
}

/* Reads particle file fname into c; exits like the reference would fail. */
static void read_chunk(const char *fname, chunk_buf *c)
{
    struct timeval a, b;
    uint64_t n = 0;
    TIMER_START(a);
    if (sjcs_particle_count(fname, &n) != 0) {
        fprintf(stderr, "ERROR: file %s not found in %s()\n", fname, "readDecodedParticles");
        exit(EXIT_FAILURE);
    }
    if (n > c->cap) {
        free(c->xyz);
        c->xyz = malloc(12 * (size_t)n);
        if (c->xyz == NULL) {
            fprintf(stderr, "ERROR: memory shortage in %s()\n", "readDecodedParticles");
            exit(EXIT_FAILURE);
        }
        c->cap = n;
    }
    if (n && sjcs_read_particles(fname, c->xyz, n) != 0) {
        fprintf(stderr, "ERROR: while reading %s in %s()\n", fname, "readDecodedParticles");
        exit(EXIT_FAILURE);
    }
    c->n = n;
    gettimeofday(&b, NULL);
    c->read_ms = sjcs_elapsed_ms(&a, &b);
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

static void read_any(const char *fname, chunk_buf *c, int compressed)
{
    if (compressed)
        read_compressed_chunk(fname, c);
    else
        read_chunk(fname, c);
}

int main(int argc, char *argv[])
{
    sjcs_cli cli;
    sjcs_halos halos = {NULL, 0};
    sjcs_engine *e = NULL;
    struct timeval tv, tvtotal;
    char err[4096];
    float *rad, *r2;
    uint64_t *counts;
    chunk_buf bufs[2];
    int nfiles, i, hthreads, device = 0, rc;
    const char *dev_env = getenv("SJCS_DEVICE");

    sjcs_cli_parse(argc, argv, &cli);
    sjcs_cli_print_info(stdout, &cli, argv);
    if (cli.errflag) {
        fprintf(stderr, "\n => %s\n", cli.errmsg);
        sjcs_cli_print_usage(stderr, argv[0]);
        exit(EXIT_FAILURE);
    }
    memset(bufs, 0, sizeof(bufs));
    hthreads = cli.nthreads < 256 ? cli.nthreads : 256;
    if (dev_env)
        device = atoi(dev_env);

    /* halos (main.c:634-642) */
    TIMER_START(tv);
    if (cli.filter_file == NULL)
        rc = sjcs_read_halos(cli.halo_file, &halos, hthreads, err, sizeof(err));
    else
        rc = sjcs_read_and_filter_halos(cli.halo_file, cli.filter_file, &halos, err, sizeof(err));
    if (rc != 0) {
        fputs(err, stderr);
        exit(EXIT_FAILURE);
    }
    print_ms(">>>> halo data reading time", &tv);
    printf("  # of halos = %lu\n", (unsigned long)halos.n);

    /* shells (main.c:645-651) */
    rad = malloc(sizeof(float) * (size_t)cli.nrads);
    r2 = malloc(sizeof(float) * (size_t)cli.nrads);
    counts = calloc((size_t)halos.n * (size_t)cli.nrads + 1, sizeof(uint64_t));
    if (!rad || !r2 || !counts) {
        fprintf(stderr, "ERROR: memory shortage\n");
        exit(EXIT_FAILURE);
    }
    sjcs_host_shell_thresholds(cli.minrad, cli.maxrad, cli.nrads, rad, r2);

    if (sjcs_engine_create(device, &e) != SJCS_OK)
        die_engine(NULL, "GPU engine");
    if (sjcs_set_shells(e, r2, cli.nrads) != SJCS_OK)
        die_engine(e, "shells");
    if (sjcs_set_halos(e, halos.n ? halos.halo[0].point : NULL, sizeof(sjcs_halo), halos.n) !=
        SJCS_OK)
        die_engine(e, "halos");

    /* chunk loop (main.c:235-368): file i+1 is read while file i counts */
    nfiles = cli.particle_end - cli.particle_beg;
    TIMER_START(tvtotal);
    puts(cli.maxbacklogs == 0 ? "  -- with no backlog execution" : "  -- with backlog execution");
    fflush(stdout);
    read_any(argv[cli.particle_beg], &bufs[0], cli.compressed);
    for (i = 0; i < nfiles; i++) {
        chunk_buf *cur = &bufs[i & 1];
        printf(">>>> particles reading time: %ld milli seconds\n", cur->read_ms);
        if (cur->n == 0)
            fprintf(stderr, "%s(): No data loaded. Cannot make tree.\n",
                    "MakeTreeWithSameThreadPoolArraySTR");
        rc = cli.compressed ? sjcs_add_compressed_particles(e, cur->uv, cur->n, &cur->dec)
                            : sjcs_add_particles(e, cur->xyz, cur->n);
        if (rc != SJCS_OK)
            die_engine(e, "particle chunk");
        if (i + 1 < nfiles)
            read_any(argv[cli.particle_beg + i + 1], &bufs[(i + 1) & 1], cli.compressed);
        finish_chunk(e, cur->n, cli.compressed);
        fflush(stdout);
    }
    print_ms(">>>> total searching time", &tvtotal);

    if (sjcs_get_counts(e, counts) != SJCS_OK)
        die_engine(e, "counts");
    sjcs_engine_destroy(e);

    /* result (main.c:660-664) */
    TIMER_START(tv);
    if (sjcs_write_result(cli.outfile, &halos, counts, r2, cli.nrads, hthreads) < 0)
        exit(EXIT_FAILURE);
    print_ms(">>>> result writing time", &tv);

    for (i = 0; i < 2; i++) {
        free(bufs[i].xyz);
        free(bufs[i].uv);
    }
    sjcs_free_halos(&halos);
    free(counts);
    free(rad);
    free(r2);
    return 0;
}
