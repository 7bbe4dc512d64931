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
 * file to file, so "index destruction" is ~0.
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
    uint64_t n, cap;
    long read_ms;
} chunk_buf;

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
static void finish_chunk(sjcs_engine *e, uint64_t n)
{
    float ms[3];
    struct timeval a;
    if (sjcs_last_chunk_timings(e, ms) != SJCS_OK)
        die_engine(e, "search");
    printf(">>>> index construction time: %ld milli seconds\n", (long)(ms[0] + ms[1]));
    printf("  # of particles = %llu\n", (unsigned long long)n);
    printf(">>>> searching time: %ld milli seconds\n", (long)ms[2]);
    TIMER_START(a);
    print_ms(">>>> index destruction time", &a);
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
    chunk_buf bufs[2] = {{NULL, 0, 0, 0}, {NULL, 0, 0, 0}};
    int nfiles, i, hthreads, device = 0, rc;
    const char *dev_env = getenv("SJCS_DEVICE");

    sjcs_cli_parse(argc, argv, &cli);
    sjcs_cli_print_info(stdout, &cli, argv);
    if (cli.errflag) {
        fprintf(stderr, "\n => %s\n", cli.errmsg);
        sjcs_cli_print_usage(stderr, argv[0]);
        exit(EXIT_FAILURE);
    }
    if (cli.compressed) {
        fprintf(stderr, "\n => --compressed (Gadget) particle input is not supported by this "
                        "engine; decode the files first.\n");
        exit(EXIT_FAILURE);
    }
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
    read_chunk(argv[cli.particle_beg], &bufs[0]);
    for (i = 0; i < nfiles; i++) {
        chunk_buf *cur = &bufs[i & 1];
        printf(">>>> particles reading time: %ld milli seconds\n", cur->read_ms);
        if (cur->n == 0)
            fprintf(stderr, "%s(): No data loaded. Cannot make tree.\n",
                    "MakeTreeWithSameThreadPoolArraySTR");
        if (sjcs_add_particles(e, cur->xyz, cur->n) != SJCS_OK)
            die_engine(e, "particle chunk");
        if (i + 1 < nfiles)
            read_chunk(argv[cli.particle_beg + i + 1], &bufs[(i + 1) & 1]);
        finish_chunk(e, cur->n);
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

    free(bufs[0].xyz);
    free(bufs[1].xyz);
    sjcs_free_halos(&halos);
    free(counts);
    free(rad);
    free(r2);
    return 0;
}
