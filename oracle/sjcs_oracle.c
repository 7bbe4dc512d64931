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
 * test is monotone, so pruning never changes which leaves pass).  The
 * triples are visited in cache-sized blocks of particles per group of halos,
 * each one once.
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
 *   compressed   src/Gadget/GadgetHeader.c:7-24 (decode parameters),
 *   particles    src/Gadget/GadId.c:38-45, 70-86, 112-137 (id -> grid point,
 *                C's truncating division), src/Gadget/BufVector3D.c:168-201
 *                (the non-AVX branch, the one the reference's build takes:
 *                nothing defines USE_AVX), with cell_interval*ip + disp_min
 *                fused into fmaf() as gcc -std=gnu99 contracts it for an FMA
 *                host, or left as two roundings (fused = 0).
 *
 * Build: gcc -O3 -ffp-contract=off -fPIC -shared (oracle/Makefile).  The
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

/* One query (a halo's periodic image): its centre, closed box and row. */
typedef struct {
    float c[3], lo[3], hi[3];
    uint64_t *row;
} oracle_query;

/* Every (query, particle) pair of a group of queries: the particles are
 * walked in blocks that stay in cache while every query of the group tests
 * them, so the particle array is read once per group rather than once per
 * query.  Each pair is still tested exactly once, so the counts are the
 * same sums as one query at a time. */
#define ORACLE_BLOCK 4096
static void count_queries(const float *particles, uint64_t np,
                          const oracle_query *q, int nq, const float *r2,
                          int nrads, uint64_t *inbox)
{
    float bx[ORACLE_BLOCK], by[ORACLE_BLOCK], bz[ORACLE_BLOCK];
    unsigned char pass[ORACLE_BLOCK];
    uint64_t b0;
    int j, i, m;

    for (b0 = 0; b0 < np; b0 += ORACLE_BLOCK) {
        m = (int)(b0 + ORACLE_BLOCK < np ? ORACLE_BLOCK : np - b0);
        for (i = 0; i < m; i++) {
            bx[i] = particles[3 * (b0 + i)];
            by[i] = particles[3 * (b0 + i) + 1];
            bz[i] = particles[3 * (b0 + i) + 2];
        }
        for (j = 0; j < nq; j++) {
            const float lo = q[j].lo[0], hi = q[j].hi[0];
            /* the x half of in_box for the whole block (a pre-filter: the
             * full test below decides) */
            for (i = 0; i < m; i++)
                pass[i] = (unsigned char)(!(bx[i] < lo) & !(hi < bx[i]));
            for (i = 0; i < m; i++) {
                float p[3];
                int k;
                if (!pass[i])
                    continue;
                p[0] = bx[i];
                p[1] = by[i];
                p[2] = bz[i];
                if (!in_box(p, q[j].lo, q[j].hi))
                    continue;
                if (inbox)
                    (*inbox)++;
                k = oracle_shell_of(oracle_distance(p, q[j].c), r2, nrads);
                if (k >= 0 && q[j].row)
                    q[j].row[k]++;
            }
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

/* Halos per group of queries (each halo makes 1 to 8 queries). */
#define ORACLE_GROUP 64

static void *oracle_worker(void *arg)
{
    oracle_job *j = arg;
    oracle_query q[8 * ORACLE_GROUP];
    uint64_t h = (uint64_t)j->tid;

    while (h < j->nh) {
        int nq = 0, g, img, d;
        for (g = 0; g < ORACLE_GROUP && h < j->nh; g++, h += (uint64_t)j->nthreads) {
            float c[3], centres[8][3];
            int nimg;
            memcpy(c, j->halos + h * j->stride, sizeof(c));
            nimg = oracle_images(c, j->rbox, centres);
            for (img = 0; img < nimg; img++, nq++) {
                for (d = 0; d < 3; d++) {
                    q[nq].c[d] = centres[img][d];
                    q[nq].lo[d] = centres[img][d] - j->rbox;
                    q[nq].hi[d] = centres[img][d] + j->rbox;
                }
                q[nq].row = j->counts ? &j->counts[h * (uint64_t)j->nrads] : NULL;
            }
        }
        count_queries(j->particles, j->np, q, nq, j->r2, j->nrads, &j->inbox);
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

/* ---- compressed (Gadget) particles ---- */

/* Header fields by byte offset (include/Gadget/GadgetHeader.h:4-27: six
 * int npart, six double mass, time, redshift, two int flags, six int
 * npartTotal, two int, four double from BoxSize, three int, float disp_min,
 * float disp_max, int pos_bits, long long id_start, ...). */
#define GH_NPART_TOTAL 96
#define GH_BOXSIZE 128
#define GH_DISP_MIN 172
#define GH_DISP_MAX 176
#define GH_POS_BITS 180
#define GH_ID_START 184

void oracle_gadget_params(const unsigned char *hdr, int ntile, long long *id_start,
                          long long *ng1d, int *nt1d, float *cell_interval,
                          float *disp_min, double *unitsep)
{
    int tot[6], bits;
    double box;
    float dmin, dmax;
    long long total;
    int at1d;

    memcpy(tot, hdr + GH_NPART_TOTAL, sizeof(tot));
    memcpy(&box, hdr + GH_BOXSIZE, sizeof(box));
    memcpy(&dmin, hdr + GH_DISP_MIN, sizeof(dmin));
    memcpy(&dmax, hdr + GH_DISP_MAX, sizeof(dmax));
    memcpy(&bits, hdr + GH_POS_BITS, sizeof(bits));
    memcpy(id_start, hdr + GH_ID_START, sizeof(*id_start));
    total = (long long)tot[1] + (long long)((unsigned long long)(long long)tot[2] << 32);
    at1d = (int)pow((double)total + 0.5, 1 / 3.);
    *cell_interval = (float)(box / (double)at1d);
    *ng1d = at1d / ntile;
    *nt1d = ntile;
    *disp_min = dmin;
    *unitsep = (double)(float)(dmax - dmin) / pow(2, bits);
}

/* id2posGadId: id - 1 splits into a tile number and a number within the
 * tile; each splits into (slowest, middle, fastest) digits; axis d's grid
 * coordinate is tile digit d * points per side + point digit d. */
void oracle_gadget_gridpos(long long id, long long ng1d, int nt1d, int ip[3])
{
    const long long ng2d = (long long)(int)(ng1d * ng1d);   /* an int there */
    const long long ng3d = ng1d * ng1d * ng1d;
    const int nt2d = nt1d * nt1d;
    const long long tile = (id - 1) / ng3d;
    const long long inner = id - 1 - tile * ng3d;
    int g[3], t[3], d;

    g[0] = (int)(inner / ng2d);
    g[1] = (int)((inner - g[0] * ng2d) / ng1d);
    g[2] = (int)(inner - g[0] * ng2d - g[1] * ng1d);
    t[0] = (int)(tile / nt2d);
    t[1] = (int)((tile - t[0] * nt2d) / nt1d);
    t[2] = (int)(tile - t[0] * nt2d - t[1] * nt1d);
    for (d = 0; d < 3; d++)
        ip[d] = (int)(t[d] * ng1d + g[d]);
}

void oracle_gadget_decode(const uint16_t *uv, uint64_t n, long long id_start, long long ng1d,
                          int nt1d, float cell_interval, float disp_min, double unitsep,
                          int fused, float *xyz)
{
    uint64_t i;
    int d, ip[3];

    for (i = 0; i < n; i++) {
        oracle_gadget_gridpos((long long)((unsigned long long)id_start + i), ng1d, nt1d, ip);
        for (d = 0; d < 3; d++) {
            float s = fused ? fmaf(cell_interval, (float)ip[d], disp_min)
                            : cell_interval * (float)ip[d] + disp_min;
            xyz[3 * i + d] = (float)((double)s + unitsep * (double)uv[3 * i + d]);
        }
    }
}
