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
 * order.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "sjcs_host.h"

#define BATCH_HALOS 32768

typedef struct {
    char *s;
    size_t len;
} label_t;

static const char digits2[] =
    "00010203040506070809101112131415161718192021222324252627282930313233343536373839"
    "40414243444546474849505152535455565758596061626364656667686970717273747576777879"
    "8081828384858687888990919293949596979899";

/* Decimal of v at p, as %lu prints it; returns the end. */
static char *put_u64(char *p, uint64_t v)
{
    char tmp[24];
    char *e = tmp + sizeof(tmp);
    char *q = e;
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
    memcpy(p, q, (size_t)(e - q));
    return p + (e - q);
}

typedef struct {
    const sjcs_halos *h;
    const uint64_t *counts;
    const label_t *labels;
    int nrads;
    uint64_t beg, end;
    char *buf;
    size_t cap, len;
} fmt_job;

static void *fmt_run(void *arg)
{
    fmt_job *j = arg;
    size_t maxline = 0, need;
    uint64_t n;
    char *p;
    int k;
    for (k = 0; k < j->nrads; k++)
        if (j->labels[k].len > maxline)
            maxline = j->labels[k].len;
    need = (j->end - j->beg) * ((size_t)j->nrads * (maxline + 22) + 24) + 1;
    if (need > j->cap) {
        free(j->buf);
        j->buf = malloc(need);
        j->cap = j->buf ? need : 0;
    }
    if (j->buf == NULL) {
        j->len = (size_t)-1;
        return NULL;
    }
    p = j->buf;
    for (n = j->beg; n < j->end; n++) {
        const uint64_t *row = j->counts + n * (uint64_t)j->nrads;
        memcpy(p, "id ", 3);
        p = put_u64(p + 3, j->h->halo[n].id);
        *p++ = '\n';
        for (k = 0; k < j->nrads; k++) {
            memcpy(p, j->labels[k].s, j->labels[k].len);
            p = put_u64(p + j->labels[k].len, row[k]);
            *p++ = '\n';
        }
    }
    j->len = (size_t)(p - j->buf);
    return NULL;
}

int sjcs_write_result(const char *fname, const sjcs_halos *h, const uint64_t *counts,
                      const float *radiuses_sq, int nrads, int nthreads)
{
    label_t *labels;
    uint64_t *total;
    fmt_job *jobs;
    pthread_t *th;
    FILE *fp;
    char msg[1024];
    uint64_t b, n;
    int k, t, rc = 0;

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
            labels[k].s = malloc(need);
            if (k == 0)
                snprintf(labels[k].s, need, "  [0.0 ~ %f]: ", r);
            else
                snprintf(labels[k].s, need, "  [%f ~ %f]: ",
                         (float)sqrt((double)radiuses_sq[k - 1]), r);
            labels[k].len = need - 1;
        } else {
            labels[k].s = malloc((size_t)len + 1);
            memcpy(labels[k].s, tmp, (size_t)len + 1);
            labels[k].len = (size_t)len;
        }
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
            return -1;
        }
    }

    /* totals: every row once, halo 0's row twice (main.c:400,413) */
    total = calloc((size_t)nrads, sizeof(uint64_t));
    for (n = 0; n < h->n; n++)
        for (k = 0; k < nrads; k++)
            total[k] += counts[n * (uint64_t)nrads + k];
    if (h->n > 0)
        for (k = 0; k < nrads; k++)
            total[k] += counts[k];

    if (nthreads < 1)
        nthreads = 1;
    jobs = calloc((size_t)nthreads, sizeof(*jobs));
    th = calloc((size_t)nthreads, sizeof(*th));
    for (b = 0; b < h->n && rc == 0; b += BATCH_HALOS) {
        uint64_t e = b + BATCH_HALOS < h->n ? b + BATCH_HALOS : h->n;
        for (t = 0; t < nthreads; t++) {
            jobs[t].h = h;
            jobs[t].counts = counts;
            jobs[t].labels = labels;
            jobs[t].nrads = nrads;
            jobs[t].beg = b + (e - b) * (uint64_t)t / (uint64_t)nthreads;
            jobs[t].end = b + (e - b) * (uint64_t)(t + 1) / (uint64_t)nthreads;
            pthread_create(&th[t], NULL, fmt_run, &jobs[t]);
        }
        for (t = 0; t < nthreads; t++)
            pthread_join(th[t], NULL);
        for (t = 0; t < nthreads && rc == 0; t++) {
            if (jobs[t].len == (size_t)-1 ||
                fwrite(jobs[t].buf, 1, jobs[t].len, fp) != jobs[t].len || ferror(fp)) {
                perror(msg);
                rc = -1;
            }
        }
    }
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
    for (t = 0; t < nthreads; t++)
        free(jobs[t].buf);
    free(jobs);
    free(th);
    free(total);
    for (k = 0; k < nrads; k++)
        free(labels[k].s);
    free(labels);
    return rc;
}

long sjcs_elapsed_ms(const void *tv_begin, const void *tv_end)
{
    const struct timeval *a = tv_begin, *b = tv_end;
    long diff = (b->tv_sec - a->tv_sec) * 1000 * 1000 + (b->tv_usec - a->tv_usec);
    return diff / 1000;
}
