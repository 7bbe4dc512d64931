/*
 * result_io.c -- the text result file, byte for byte as astr_rcount's
 * outputResult writes it (src/astr_rcount/main.c:371-442):
 *
 *   id <id>\n
 *     [0.0 ~ <r0>]: <count>\n
 *     [<r(k-1)> ~ <rk>]: <count>\n      (k = 1 .. K-1)
 *   ...
 *   \nTotal\n  and the same K lines with the totals.
 *
 * Radii are (float)sqrt((double)r2[k]) printed with %f; counts with %lu.
 * The reference's total array aliases halo 0's count row and is updated
 * while the rows are printed, so its Total is the sum over all halos plus
 * halo 0's counts once more; that is reproduced here.
 *
 * The K label strings are formatted once; the counts are formatted by
 * worker threads into per-thread buffers, batch by batch, and written in
 * order.  The rows come from a callback, asked for the next batch while the
 * workers format the current one (the device read-back overlaps the
 * formatting); the totals are summed by the workers as they go.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "sjcs_host.h"

#define BATCH_HALOS SJCS_RESULT_BATCH

typedef struct {
    char *s;     /* at least LABEL_SLOT bytes allocated */
    size_t len;
} label_t;

static const char digits2[] =
    "00010203040506070809101112131415161718192021222324252627282930313233343536373839"
    "40414243444546474849505152535455565758596061626364656667686970717273747576777879"
    "8081828384858687888990919293949596979899";

/* Decimal of v at p, as %lu prints it; returns the end. */
static inline char *put_u64(char *p, uint64_t v)
{
    int nd;
    char *q;
    if (v < 10) {
        *p = (char)('0' + v);
        return p + 1;
    }
    if (v < 100) {
        p[0] = digits2[2 * v];
        p[1] = digits2[2 * v + 1];
        return p + 2;
    }
    nd = 3;
    {
        uint64_t t = v / 1000;
        while (t) {
            nd++;
            t /= 10;
        }
    }
    q = p + nd;
    while (v >= 100) {
        unsigned r = (unsigned)(v % 100);
        v /= 100;
        q -= 2;
        q[0] = digits2[2 * r];
        q[1] = digits2[2 * r + 1];
    }
    if (v >= 10) {
        q -= 2;
        q[0] = digits2[2 * v];
        q[1] = digits2[2 * v + 1];
    } else {
        *--q = (char)('0' + v);
    }
    return p + nd;
}

/* A label of at most LABEL_SLOT bytes is copied as one fixed-size block
 * (the bytes past its end are overwritten by what follows). */
#define LABEL_SLOT 32

typedef struct {
    const sjcs_halos *h;
    const void *rows;       /* rows of the batch, from halo `first` */
    int wide;               /* u64 rows (else u32) */
    uint64_t first;
    const label_t *labels;
    const char *slots;      /* every label in a LABEL_SLOT-byte slot, or NULL */
    const unsigned char *slot_len;
    int nrads;
    uint64_t beg, end;
    uint64_t *tot;          /* this worker's share of the totals */
    char *buf;
    size_t cap, len;
} fmt_job;

static void fmt_batch(fmt_job *j);

/* Workers that stay up for the whole result: the main thread bumps `gen`
 * for each batch, every worker formats its job and the last one to finish
 * wakes the main thread. */
typedef struct {
    pthread_mutex_t mu;
    pthread_cond_t go, finished;
    uint64_t gen;
    int remaining, stop;
} pool_t;

typedef struct {
    pool_t *pool;
    fmt_job *job;
} worker_arg;

static void *pool_run(void *arg)
{
    worker_arg *w = arg;
    pool_t *pl = w->pool;
    uint64_t seen = 0;
    for (;;) {
        pthread_mutex_lock(&pl->mu);
        while (pl->gen == seen && !pl->stop)
            pthread_cond_wait(&pl->go, &pl->mu);
        if (pl->stop) {
            pthread_mutex_unlock(&pl->mu);
            break;
        }
        seen = pl->gen;
        pthread_mutex_unlock(&pl->mu);
        fmt_batch(w->job);
        pthread_mutex_lock(&pl->mu);
        if (--pl->remaining == 0)
            pthread_cond_signal(&pl->finished);
        pthread_mutex_unlock(&pl->mu);
    }
    return NULL;
}

/* Count i of a block of u64 (wide) or u32 rows. */
#define ROW_AT(rows, wide, i) \
    ((wide) ? ((const uint64_t *)(rows))[i] : (uint64_t)((const uint32_t *)(rows))[i])

/* The rows of job j when every label fits a slot: the labels are read from
 * one table (no pointer per line, nothing reloaded after each store through
 * the output pointer), counts below 100 take one path whichever their
 * digit count, and the totals are summed in a second pass over the rows.
 * Inlined with `wide` a constant, one copy per row width. */
static inline __attribute__((always_inline)) void fmt_rows_slots_w(fmt_job *j, const int wide)
{
    char *restrict p = j->buf;
    const int K = j->nrads;
    const char *restrict slots = j->slots;
    const unsigned char *restrict slen = j->slot_len;
    const sjcs_halo *restrict halo = j->h->halo;
    const void *rows = j->rows;
    uint64_t *restrict tot = j->tot;
    uint64_t n, i = (j->beg - j->first) * (uint64_t)K;
    int k;
    for (n = j->beg; n < j->end; n++) {
        memcpy(p, "id ", 3);
        p = put_u64(p + 3, halo[n].id);
        *p++ = '\n';
        for (k = 0; k < K; k++, i++) {
            const uint64_t v = ROW_AT(rows, wide, i);
            __builtin_memcpy(p, slots + (size_t)k * LABEL_SLOT, LABEL_SLOT);
            p += slen[k];
            if (v < 100) {
                const unsigned two = v >= 10;
                p[0] = digits2[2 * v + 1 - two];
                p[1] = digits2[2 * v + 1];
                p[1 + two] = '\n';
                p += 2 + two;
            } else {
                p = put_u64(p, v);
                *p++ = '\n';
            }
        }
    }
    j->len = (size_t)(p - j->buf);
    i = (j->beg - j->first) * (uint64_t)K;
    for (n = j->beg; n < j->end; n++)
        for (k = 0; k < K; k++, i++)
            tot[k] += ROW_AT(rows, wide, i);
}

static void fmt_rows_slots(fmt_job *j)
{
    if (j->wide)
        fmt_rows_slots_w(j, 1);
    else
        fmt_rows_slots_w(j, 0);
}

static void fmt_batch(fmt_job *j)
{
    size_t maxline = 0, need;
    uint64_t n;
    char *p;
    int k;
    for (k = 0; k < j->nrads; k++)
        if (j->labels[k].len > maxline)
            maxline = j->labels[k].len;
    need = (j->end - j->beg) * ((size_t)j->nrads * (maxline + 22) + 24) + LABEL_SLOT + 1;
    if (need > j->cap) {
        free(j->buf);
        j->buf = malloc(need);
        j->cap = j->buf ? need : 0;
    }
    if (j->buf == NULL) {
        j->len = (size_t)-1;
        return;
    }
    if (j->slots) {
        fmt_rows_slots(j);
        return;
    }
    p = j->buf;
    for (n = j->beg; n < j->end; n++) {
        const uint64_t r0 = (n - j->first) * (uint64_t)j->nrads;
        memcpy(p, "id ", 3);
        p = put_u64(p + 3, j->h->halo[n].id);
        *p++ = '\n';
        for (k = 0; k < j->nrads; k++) {
            const label_t *l = &j->labels[k];
            if (l->len <= LABEL_SLOT)
                __builtin_memcpy(p, l->s, LABEL_SLOT);
            else
                memcpy(p, l->s, l->len);
            const uint64_t v = ROW_AT(j->rows, j->wide, r0 + (uint64_t)k);
            p = put_u64(p + l->len, v);
            *p++ = '\n';
            j->tot[k] += v;
        }
    }
    j->len = (size_t)(p - j->buf);
}

typedef struct {
    const uint64_t *counts;
    int nrads;
} array_rows;

static const uint64_t *rows_of_array(void *ctx, uint64_t beg, uint64_t n)
{
    const array_rows *a = ctx;
    (void)n;
    return a->counts + beg * (uint64_t)a->nrads;
}

static int write_rows(const char *fname, const sjcs_halos *h, sjcs_rows_fn rows64,
                      sjcs_rows32_fn rows32, void *ctx, const float *radiuses_sq, int nrads,
                      int nthreads);

int sjcs_write_result(const char *fname, const sjcs_halos *h, const uint64_t *counts,
                      const float *radiuses_sq, int nrads, int nthreads)
{
    array_rows a = {counts, nrads};
    return write_rows(fname, h, rows_of_array, NULL, &a, radiuses_sq, nrads, nthreads);
}

int sjcs_write_result_rows(const char *fname, const sjcs_halos *h, sjcs_rows_fn rows, void *ctx,
                           const float *radiuses_sq, int nrads, int nthreads)
{
    return write_rows(fname, h, rows, NULL, ctx, radiuses_sq, nrads, nthreads);
}

int sjcs_write_result_rows32(const char *fname, const sjcs_halos *h, sjcs_rows32_fn rows,
                             void *ctx, const float *radiuses_sq, int nrads, int nthreads)
{
    return write_rows(fname, h, NULL, rows, ctx, radiuses_sq, nrads, nthreads);
}

/* The rows of halos [beg, beg + n) from whichever callback is set. */
static const void *get_rows(sjcs_rows_fn rows64, sjcs_rows32_fn rows32, void *ctx, uint64_t beg,
                            uint64_t n)
{
    return rows64 ? (const void *)rows64(ctx, beg, n) : (const void *)rows32(ctx, beg, n);
}

static int write_rows(const char *fname, const sjcs_halos *h, sjcs_rows_fn rows64,
                      sjcs_rows32_fn rows32, void *ctx, const float *radiuses_sq, int nrads,
                      int nthreads)
{
    label_t *labels;
    char *slots;
    unsigned char *slot_len;
    uint64_t *total, *row0;
    const void *cur, *next;
    const int wide = rows64 != NULL;
    fmt_job *jobs;
    pthread_t *th;
    worker_arg *wargs;
    pool_t pool;
    FILE *fp;
    char msg[1024];
    uint64_t b;
    int k, t, nworkers, rc = 0;

    snprintf(msg, sizeof(msg), "ERROR: while writing %s in %s()\n", fname ? fname : "(null)",
             "outputResult");
    labels = calloc((size_t)nrads, sizeof(*labels));
    for (k = 0; k < nrads; k++) {
        float r = (float)sqrt((double)radiuses_sq[k]);
        char tmp[256];
        int len;
        if (k == 0)
            len = snprintf(tmp, sizeof(tmp), "  [0.0 ~ %f]: ", r);
        else
            len = snprintf(tmp, sizeof(tmp), "  [%f ~ %f]: ",
                           (float)sqrt((double)radiuses_sq[k - 1]), r);
        if (len < 0 || (size_t)len >= sizeof(tmp)) {
            /* a radius too large for the scratch buffer: format exactly */
            size_t need;
            if (k == 0)
                need = (size_t)snprintf(NULL, 0, "  [0.0 ~ %f]: ", r) + 1;
            else
                need = (size_t)snprintf(NULL, 0, "  [%f ~ %f]: ",
                                        (float)sqrt((double)radiuses_sq[k - 1]), r) + 1;
            labels[k].s = calloc(need > LABEL_SLOT ? need : LABEL_SLOT, 1);
            if (k == 0)
                snprintf(labels[k].s, need, "  [0.0 ~ %f]: ", r);
            else
                snprintf(labels[k].s, need, "  [%f ~ %f]: ",
                         (float)sqrt((double)radiuses_sq[k - 1]), r);
            labels[k].len = need - 1;
        } else {
            labels[k].s = calloc((size_t)len + 1 > LABEL_SLOT ? (size_t)len + 1 : LABEL_SLOT, 1);
            memcpy(labels[k].s, tmp, (size_t)len + 1);
            labels[k].len = (size_t)len;
        }
    }

    slots = malloc((size_t)nrads * LABEL_SLOT + 1);
    slot_len = malloc((size_t)nrads + 1);
    for (k = 0; k < nrads && slots && slot_len; k++) {
        if (labels[k].len > LABEL_SLOT)
            break;
        memcpy(slots + (size_t)k * LABEL_SLOT, labels[k].s, LABEL_SLOT);
        slot_len[k] = (unsigned char)labels[k].len;
    }
    if (k < nrads) {
        free(slots);
        free(slot_len);
        slots = NULL;
        slot_len = NULL;
    }

    if (fname == NULL) {
        fp = stdout;
    } else {
        fp = fopen(fname, "w");
        if (fp == NULL) {
            fprintf(stderr, "Cannot create %s in %s().\n", fname, "outputResult");
            for (k = 0; k < nrads; k++)
                free(labels[k].s);
            free(labels);
            free(slots);
            free(slot_len);
            return -1;
        }
    }

    /* totals: every row once, halo 0's row twice (main.c:400,413); the
     * workers sum the rows they format */
    total = calloc((size_t)nrads, sizeof(uint64_t));
    row0 = calloc((size_t)nrads, sizeof(uint64_t));

    if (nthreads < 1)
        nthreads = 1;
    jobs = calloc((size_t)nthreads, sizeof(*jobs));
    th = calloc((size_t)nthreads, sizeof(*th));
    wargs = calloc((size_t)nthreads, sizeof(*wargs));
    for (t = 0; t < nthreads; t++)
        jobs[t].tot = calloc((size_t)nrads, sizeof(uint64_t));
    pthread_mutex_init(&pool.mu, NULL);
    pthread_cond_init(&pool.go, NULL);
    pthread_cond_init(&pool.finished, NULL);
    pool.gen = 0;
    pool.remaining = 0;
    pool.stop = 0;
    /* jobs [0, nworkers) run on workers, the rest (if a thread could not be
     * started) on this thread */
    nworkers = 0;
    for (t = 0; t < nthreads; t++) {
        wargs[t].pool = &pool;
        wargs[t].job = &jobs[t];
        if (pthread_create(&th[t], NULL, pool_run, &wargs[t]) != 0)
            break;
        nworkers++;
    }
    cur = h->n ? get_rows(rows64, rows32, ctx, 0, h->n < BATCH_HALOS ? h->n : BATCH_HALOS) : NULL;
    if (h->n && cur == NULL)
        rc = -1;
    else if (h->n)
        for (k = 0; k < nrads; k++)
            row0[k] = ROW_AT(cur, wide, k);
    for (b = 0; b < h->n && rc == 0; b += BATCH_HALOS) {
        uint64_t e = b + BATCH_HALOS < h->n ? b + BATCH_HALOS : h->n;
        for (t = 0; t < nthreads; t++) {
            jobs[t].h = h;
            jobs[t].rows = cur;
            jobs[t].wide = wide;
            jobs[t].first = b;
            jobs[t].labels = labels;
            jobs[t].slots = slots;
            jobs[t].slot_len = slot_len;
            jobs[t].nrads = nrads;
            jobs[t].beg = b + (e - b) * (uint64_t)t / (uint64_t)nthreads;
            jobs[t].end = b + (e - b) * (uint64_t)(t + 1) / (uint64_t)nthreads;
        }
        pthread_mutex_lock(&pool.mu);
        pool.remaining = nworkers;
        pool.gen++;
        pthread_cond_broadcast(&pool.go);
        pthread_mutex_unlock(&pool.mu);
        /* the next batch's rows arrive while this one is formatted */
        next = NULL;
        if (e < h->n) {
            uint64_t e2 = e + BATCH_HALOS < h->n ? e + BATCH_HALOS : h->n;
            next = get_rows(rows64, rows32, ctx, e, e2 - e);
            if (next == NULL)
                rc = -1;
        }
        for (t = nworkers; t < nthreads; t++)
            fmt_batch(&jobs[t]);
        pthread_mutex_lock(&pool.mu);
        while (pool.remaining > 0)
            pthread_cond_wait(&pool.finished, &pool.mu);
        pthread_mutex_unlock(&pool.mu);
        for (t = 0; t < nthreads && rc == 0; t++) {
            if (jobs[t].len == (size_t)-1 ||
                fwrite(jobs[t].buf, 1, jobs[t].len, fp) != jobs[t].len || ferror(fp)) {
                perror(msg);
                rc = -1;
            }
        }
        cur = next;
    }
    pthread_mutex_lock(&pool.mu);
    pool.stop = 1;
    pthread_cond_broadcast(&pool.go);
    pthread_mutex_unlock(&pool.mu);
    for (t = 0; t < nworkers; t++)
        pthread_join(th[t], NULL);
    pthread_mutex_destroy(&pool.mu);
    pthread_cond_destroy(&pool.go);
    pthread_cond_destroy(&pool.finished);
    for (t = 0; t < nthreads; t++)
        for (k = 0; k < nrads; k++)
            total[k] += jobs[t].tot[k];
    if (h->n > 0)
        for (k = 0; k < nrads; k++)
            total[k] += row0[k];
    if (rc == 0) {
        fputs("\nTotal\n", fp);
        for (k = 0; k < nrads; k++) {
            char line[64];
            char *p = put_u64(line, total[k]);
            *p++ = '\n';
            fwrite(labels[k].s, 1, labels[k].len, fp);
            fwrite(line, 1, (size_t)(p - line), fp);
            if (ferror(fp)) {
                perror(msg);
                rc = -1;
                break;
            }
        }
    }
    if (fp != stdout) {
        if (fclose(fp) != 0 && rc == 0) {
            perror(msg);
            rc = -1;
        }
    } else {
        fflush(fp);
    }
    for (t = 0; t < nthreads; t++) {
        free(jobs[t].buf);
        free(jobs[t].tot);
    }
    free(jobs);
    free(th);
    free(wargs);
    free(total);
    free(row0);
    for (k = 0; k < nrads; k++)
        free(labels[k].s);
    free(labels);
    free(slots);
    free(slot_len);
    return rc;
}

long sjcs_elapsed_ms(const void *tv_begin, const void *tv_end)
{
    const struct timeval *a = tv_begin, *b = tv_end;
    long diff = (b->tv_sec - a->tv_sec) * 1000 * 1000 + (b->tv_usec - a->tv_usec);
    return diff / 1000;
}
