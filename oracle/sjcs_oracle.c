/*
 * sjcs_oracle.c -- CPU restatement of astr_rcount's per-halo periodic
 * multi-shell range count.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links it and has
 * no CPU fallback.
 *
 * It is a brute force over every (halo, periodic image, particle) triple, so
 * its result is the set definition of the count, independent of any index
 * structure (the reference's STR tree only prunes; internal MBRs are exact
 * min/max of their children, src/objects/ArraySTR/Node.c:21-31, and the box
 * test is monotone, so pruning never changes which leaves pass).
 *
 * Reference behaviour restated, with the file:line each step follows:
 *   thresholds   include/support_functions.h:18-34 (Logspace), with the
 *                exponent a + i*c contracted to fma() as gcc -O2 on an FMA
 *                host compiles it; squared in FP32 at
 *                src/astr_rcount/main.c:648-649.
 *   box radius   src/objects/ArraySTR/MultiRangeCount.c:19
 *                (float)sqrt((double)R2[K-1]).
 *   query box    src/objects/RTree/Mbr.c:207-239 (c - r, c + r in FP32).
 *   images       src/objects/ArraySTR/MultiRangeCount.c:31-71 and
 *                src/objects/RTree/Mbr.c:284-313: axis d is out of bounds
 *                low if lo[d] < 0, high if hi[d] > 1000; low wins when both
 *                hold; every non-empty subset of out-of-bounds axes is one
 *                more query, centre shifted by +-1000 on those axes.
 *   box test     src/objects/RTree/Mbr.c:167-205 (cmpnlt both ways: a NaN
 *                coordinate passes).
 *   distance     include/objects/RTree/Distance.h:5-25, DPPS imm 0x77:
 *                (dx*dx + dy*dy) + dz*dz, every op rounded to FP32, no FMA.
 *   shell        src/objects/ArraySTR/MultiRangeCount.c:107-113: first k
 *                with dist < R2[k]; nothing counted when none matches.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).  The
 * -ffp-contract=off is what keeps the distance free of FMAs.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "sjcs_oracle.h"

#define ORACLE_BOX_LO 0.0f
#define ORACLE_BOX_HI 1000.0f

/* Logspace + squaring (include/support_functions.h:18-34,
 * src/astr_rcount/main.c:646-649). */
void oracle_shell_thresholds(double minrad, double maxrad, int nrads,
                             float *radiuses, float *r2)
{
    double a = log2(minrad);
    double b = log2(maxrad);
    double c = (b - a) / (nrads - 1);
    int i;

    for (i = 0; i < nrads - 1; i++)
        radiuses[i] = (float)pow(2.0, fma((double)i, c, a));
    radiuses[nrads - 1] = (float)pow(2.0, b);
    for (i = 0; i < nrads; i++)
        r2[i] = radiuses[i] * radiuses[i];
}

/* The same without the FMA, for the test that shows the difference matters. */
void oracle_shell_thresholds_nofma(double minrad, double maxrad, int nrads,
                                   float *radiuses, float *r2)
{
    double a = log2(minrad);
    double b = log2(maxrad);
    double c = (b - a) / (nrads - 1);
    volatile double prod;
    int i;

    for (i = 0; i < nrads - 1; i++) {
        prod = (double)i * c;
        radiuses[i] = (float)pow(2.0, a + prod);
    }
    radiuses[nrads - 1] = (float)pow(2.0, b);
    for (i = 0; i < nrads; i++)
        r2[i] = radiuses[i] * radiuses[i];
}

float oracle_box_radius(const float *r2, int nrads)
{
    return (float)sqrt((double)r2[nrads - 1]);
}

float oracle_distance(const float p[3], const float c[3])
{
    float dx = p[0] - c[0];
    float dy = p[1] - c[1];
    float dz = p[2] - c[2];
    float sxy = dx * dx + dy * dy;
    return sxy + dz * dz;
}

int oracle_shell_of(float dist, const float *r2, int nrads)
{
    int k;
    for (k = 0; k < nrads; k++)
        if (dist < r2[k])
            return k;
    return -1;
}

/* Fills shifts[0..n) with the centre shifts of every query the reference
 * issues for centre c (the original query first), returns n (1..8). */
int oracle_images(const float c[3], float rbox, float centres[8][3])
{
    float lo[3], hi[3], shift[3];
    int oob[3], axes[3], naxes = 0, d, s, n = 0;

    for (d = 0; d < 3; d++) {
        lo[d] = c[d] - rbox;
        hi[d] = c[d] + rbox;
        oob[d] = 0;
        shift[d] = 0.0f;
        if (lo[d] < ORACLE_BOX_LO) {
            oob[d] = 1;
            shift[d] = ORACLE_BOX_HI;
        } else if (hi[d] > ORACLE_BOX_HI) {
            oob[d] = 1;
            shift[d] = -ORACLE_BOX_HI;
        }
        if (oob[d])
            axes[naxes++] = d;
    }
    for (s = 0; s < (1 << naxes); s++) {
        int j;
        for (d = 0; d < 3; d++)
            centres[n][d] = c[d];
        for (j = 0; j < naxes; j++)
            if (s & (1 << j))
                centres[n][axes[j]] = c[axes[j]] + shift[axes[j]];
        n++;
    }
    return n;
}

static int in_box(const float p[3], const float lo[3], const float hi[3])
{
    int d;
    for (d = 0; d < 3; d++)
        if (p[d] < lo[d] || hi[d] < p[d])
            return 0;
    return 1;
}

static void count_one_halo(const float *particles, uint64_t np,
                           const float c[3], const float *r2, int nrads,
                           float rbox, uint64_t *row, uint64_t *inbox)
{
    float centres[8][3];
    int nimg = oracle_images(c, rbox, centres);
    int img, d;
    uint64_t i;

    for (img = 0; img < nimg; img++) {
        float lo[3], hi[3];
        for (d = 0; d < 3; d++) {
            lo[d] = centres[img][d] - rbox;
            hi[d] = centres[img][d] + rbox;
        }
        for (i = 0; i < np; i++) {
            const float *p = &particles[3 * i];
            int k;
            if (!in_box(p, lo, hi))
                continue;
            if (inbox)
                (*inbox)++;
            k = oracle_shell_of(oracle_distance(p, centres[img]), r2, nrads);
            if (k >= 0 && row)
                row[k]++;
        }
    }
}

typedef struct {
    const float *particles;
    uint64_t np;
    const char *halos;
    size_t stride;
    uint64_t nh;
    const float *r2;
    int nrads;
    float rbox;
    uint64_t *counts;
    uint64_t inbox;
    int tid, nthreads;
} oracle_job;

static void *oracle_worker(void *arg)
{
    oracle_job *j = arg;
    uint64_t h;
    for (h = (uint64_t)j->tid; h < j->nh; h += (uint64_t)j->nthreads) {
        float c[3];
        memcpy(c, j->halos + h * j->stride, sizeof(c));
        count_one_halo(j->particles, j->np, c, j->r2, j->nrads, j->rbox,
                       j->counts ? &j->counts[h * (uint64_t)j->nrads] : NULL,
                       &j->inbox);
    }
    return NULL;
}

/* counts[h*nrads + k] += pairs of halo h in shell k; returns the number of
 * in-box pair tests (the work unit W).  counts may be NULL. */
uint64_t oracle_count(const float *particles, uint64_t np,
                      const void *halo_points, size_t stride_bytes, uint64_t nh,
                      const float *r2, int nrads, uint64_t *counts,
                      int nthreads)
{
    oracle_job *jobs;
    pthread_t *th;
    uint64_t total = 0;
    float rbox;
    int t;

    if (nrads <= 0)
        return 0;
    if (nthreads < 1)
        nthreads = 1;
    rbox = oracle_box_radius(r2, nrads);
    jobs = calloc((size_t)nthreads, sizeof(*jobs));
    th = calloc((size_t)nthreads, sizeof(*th));
    for (t = 0; t < nthreads; t++) {
        jobs[t].particles = particles;
        jobs[t].np = np;
        jobs[t].halos = halo_points;
        jobs[t].stride = stride_bytes;
        jobs[t].nh = nh;
        jobs[t].r2 = r2;
        jobs[t].nrads = nrads;
        jobs[t].rbox = rbox;
        jobs[t].counts = counts;
        jobs[t].tid = t;
        jobs[t].nthreads = nthreads;
        pthread_create(&th[t], NULL, oracle_worker, &jobs[t]);
    }
    for (t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        total += jobs[t].inbox;
    }
    free(jobs);
    free(th);
    return total;
}
