/*
 * halo_io.c -- the halo text file.
 *
 * Format and reading rules of the reference (src/Gadget/BufHalo.c:76-249):
 *   - 16 header lines are skipped; fewer is an error;
 *   - every further line is one halo: id = strtoul of the line, mass =
 *     strtof after the 2nd space, x, y, z = strtof after the 8th, 9th and
 *     10th space (single ' ' characters, counted as strchr finds them);
 *   - lines are read as fgets(buf, 16384) returns them, and the loop stops
 *     as soon as fgets reaches end-of-file, so a last line without '\n' is
 *     not a halo.
 * The whole file is read at once (parallel preads) and its records are cut
 * exactly where those fgets calls would cut them, then parsed in parallel.
 * Every value is what the C library's strtoul / strtof give: plain decimal
 * numbers take a fast path that is exact (below), anything else -- signs of
 * strtoul, spaces, hex floats, inf, nan, 20 or more significant digits,
 * large exponents, float subnormals and overflow, decimal float midpoints --
 * goes to the library functions themselves.
 */
#include <errno.h>
#include <float.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "sjcs_host.h"

#define REC_MAX (SJCS_HALO_LINE_BUF - 1) /* fgets reads at most size-1 chars */

static const double k_pow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                   1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                   1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

#define IS_DIGIT(c) ((unsigned)((c) - '0') < 10u)

/* strtof(s, NULL).  For [+-]digits[.digits][(e|E)[+-]digits] with a
 * significand m < 2^53 (at most 19 digits read) and a decimal exponent
 * |e| <= 22, m and 10^|e| are exact doubles, so m * 10^e (or m / 10^-e) is the decimal's value
 * rounded once to double.  Float midpoints are doubles, and rounding is
 * monotone, so that double and the decimal lie on the same side of every
 * float midpoint unless the double is one: then its rounding to float is
 * the decimal's correctly rounded float, as strtof's is.  A double on a
 * midpoint, outside the normal float range, or any other form goes to
 * strtof. */
static float parse_float(const char *s)
{
    const char *p = s, *d0;
    uint64_t m = 0, bits;
    int neg = 0, nd = 0, frac = 0, any = 0, e10;
    double q;
    float f;
    if (*p == '+' || *p == '-')
        neg = *p++ == '-';
    d0 = p;
    for (; IS_DIGIT(*p); p++) {
        any = 1;
        if (m || *p != '0') {
            if (nd == 19)
                return strtof(s, NULL);
            m = m * 10 + (uint64_t)(*p - '0');
            nd++;
        }
    }
    if (*p == '.') {
        for (p++; IS_DIGIT(*p); p++) {
            any = 1;
            frac++;
            if (m || *p != '0') {
                if (nd == 19)
                    return strtof(s, NULL);
                m = m * 10 + (uint64_t)(*p - '0');
                nd++;
            }
        }
    }
    /* "0x...", "inf", "nan", leading spaces, a lone "." */
    if (!any || ((*p == 'x' || *p == 'X') && p == d0 + 1 && *d0 == '0'))
        return strtof(s, NULL);
    e10 = -frac;
    if (*p == 'e' || *p == 'E') {
        const char *q2 = p + 1;
        int eneg = 0, ev = 0;
        if (*q2 == '+' || *q2 == '-')
            eneg = *q2++ == '-';
        if (IS_DIGIT(*q2)) {
            for (; IS_DIGIT(*q2); q2++)
                if (ev < 100000)
                    ev = ev * 10 + (*q2 - '0');
            e10 += eneg ? -ev : ev;
        }
    }
    if (m == 0)
        return neg ? -0.0f : 0.0f;
    if ((m >> 53) != 0 || e10 < -22 || e10 > 22)
        return strtof(s, NULL);
    q = e10 >= 0 ? (double)m * k_pow10[e10] : (double)m / k_pow10[-e10];
    if (!(q >= (double)FLT_MIN && q <= (double)FLT_MAX))
        return strtof(s, NULL);
    memcpy(&bits, &q, sizeof(bits));
    if ((bits & ((1ull << 29) - 1)) == (1ull << 28))
        return strtof(s, NULL);
    f = (float)q;
    return neg ? -f : f;
}

/* strtoul(s, NULL, 10); digits-only values below 10^19 directly. */
static uint64_t parse_u64(const char *s)
{
    const char *p = s;
    uint64_t v = 0;
    int nd = 0;
    if (!IS_DIGIT(*p))
        return strtoul(s, NULL, 10);
    for (; IS_DIGIT(*p); p++) {
        if (++nd > 19)
            return strtoul(s, NULL, 10);
        v = v * 10 + (uint64_t)(*p - '0');
    }
    return v;
}

int sjcs_parse_halo_line(const char *str, uint64_t *id, float *mass, float point[3])
{
    int i;
    *id = parse_u64(str);
    for (i = 0; i < 2; i++) {
        if ((str = strchr(str, ' ')) == NULL)
            return -1;
        str++;
    }
    *mass = parse_float(str);
    for (i = 0; i < 6; i++) {
        if ((str = strchr(str, ' ')) == NULL)
            return -1;
        str++;
    }
    point[0] = parse_float(str);
    if ((str = strchr(str, ' ')) == NULL)
        return -1;
    str++;
    point[1] = parse_float(str);
    if ((str = strchr(str, ' ')) == NULL)
        return -1;
    str++;
    point[2] = parse_float(str);
    return 0;
}

/* Length of the record fgets would return at buf[pos..size); *at_eof is
 * what feof() would say right after that call. */
static size_t next_record(const char *buf, size_t size, size_t pos, int *at_eof)
{
    size_t left = size - pos;
    size_t lim = left < REC_MAX ? left : REC_MAX;
    const char *nl = memchr(buf + pos, '\n', lim);
    if (nl) {
        *at_eof = 0;
        return (size_t)(nl - (buf + pos)) + 1;
    }
    *at_eof = left < REC_MAX; /* stopped by end of file, not by the size cap */
    return lim;
}

static int slurp(const char *fname, char **out, size_t *size, int nthreads)
{
    struct stat st;
    char *buf;
    if (stat(fname, &st) != 0)
        return -1;
    buf = malloc((size_t)st.st_size + 1);
    if (buf == NULL)
        return -2;
    if (sjcs_pread_par(fname, 0, (uint64_t)st.st_size, buf, nthreads) != 0) {
        free(buf);
        return -1;
    }
    buf[st.st_size] = '\0';
    *out = buf;
    *size = (size_t)st.st_size;
    return 0;
}

typedef struct {
    const char *buf;
    size_t size, beg, end; /* records starting in [beg, end) */
    uint64_t nrec;         /* complete records (not ended by EOF) */
    uint64_t first;        /* index of this range's first record */
    uint64_t first_bad;    /* global index of the first unparseable record */
    sjcs_halo *dst;
} parse_job;

static void *count_job(void *arg)
{
    parse_job *j = arg;
    size_t pos = j->beg;
    uint64_t n = 0;
    while (pos < j->end) {
        int eof;
        size_t len = next_record(j->buf, j->size, pos, &eof);
        if (eof)
            break;
        n++;
        pos += len;
    }
    j->nrec = n;
    return NULL;
}

static void *parse_job_run(void *arg)
{
    parse_job *j = arg;
    char line[SJCS_HALO_LINE_BUF];
    size_t pos = j->beg;
    uint64_t r;
    j->first_bad = UINT64_MAX;
    for (r = 0; r < j->nrec; r++) {
        int eof;
        size_t len = next_record(j->buf, j->size, pos, &eof);
        sjcs_halo *h = &j->dst[j->first + r];
        memcpy(line, j->buf + pos, len);
        line[len] = '\0';
        if (sjcs_parse_halo_line(line, &h->id, &h->mass, h->point) < 0 &&
            j->first_bad == UINT64_MAX)
            j->first_bad = j->first + r;
        pos += len;
    }
    return NULL;
}

/* Parses every record after the headers.  Returns 0 and the records, the
 * index of the first bad one (== *nrec when none), or -1 (no file), -2
 * (fewer than 16 header lines), -3 (out of memory). */
static int parse_file(const char *fname, int nthreads, sjcs_halo **out, uint64_t *nrec,
                      uint64_t *first_bad)
{
    char *buf = NULL;
    size_t size = 0, pos = 0, start;
    parse_job *jobs;
    pthread_t *th;
    uint64_t total = 0;
    int t, row, rc;

    rc = slurp(fname, &buf, &size, nthreads);
    if (rc == -1)
        return -1;
    if (rc != 0)
        return -3;
    for (row = 0; row < SJCS_HALO_HEADER_LINES; row++) {
        int eof;
        size_t len = next_record(buf, size, pos, &eof);
        if (eof) {
            free(buf);
            return -2;
        }
        pos += len;
    }
    start = pos;
    if (nthreads < 1)
        nthreads = 1;
    if ((size_t)nthreads > (size - start) / 65536 + 1)
        nthreads = (int)((size - start) / 65536 + 1);
    jobs = calloc((size_t)nthreads, sizeof(*jobs));
    th = calloc((size_t)nthreads, sizeof(*th));
    /* range starts: the first line start at or after an even split point */
    for (t = 0; t < nthreads; t++) {
        size_t b = start + (size - start) * (size_t)t / (size_t)nthreads;
        if (t > 0) {
            const char *nl = b < size ? memchr(buf + b - 1, '\n', size - (b - 1)) : NULL;
            b = nl ? (size_t)(nl - buf) + 1 : size;
            if (b < jobs[t - 1].beg)
                b = jobs[t - 1].beg;
        }
        jobs[t].buf = buf;
        jobs[t].size = size;
        jobs[t].beg = b;
    }
    for (t = 0; t < nthreads; t++)
        jobs[t].end = t + 1 < nthreads ? jobs[t + 1].beg : size;
    for (t = 0; t < nthreads; t++)
        pthread_create(&th[t], NULL, count_job, &jobs[t]);
    for (t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        jobs[t].first = total;
        total += jobs[t].nrec;
    }
    *out = malloc(sizeof(sjcs_halo) * (total ? total : 1));
    if (*out == NULL) {
        free(buf);
        free(jobs);
        free(th);
        return -3;
    }
    for (t = 0; t < nthreads; t++) {
        jobs[t].dst = *out;
        pthread_create(&th[t], NULL, parse_job_run, &jobs[t]);
    }
    *first_bad = total;
    for (t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        if (jobs[t].first_bad < *first_bad)
            *first_bad = jobs[t].first_bad;
    }
    *nrec = total;
    free(buf);
    free(jobs);
    free(th);
    return 0;
}

static void set_err(char *err, size_t errlen, const char *fmt, const char *fname,
                    const char *func)
{
    if (err && errlen)
        snprintf(err, errlen, fmt, fname, func);
}

int sjcs_read_halos(const char *fname, sjcs_halos *out, int nthreads, char *err, size_t errlen)
{
    sjcs_halo *h = NULL;
    uint64_t n = 0, bad = 0;
    int rc = parse_file(fname, nthreads, &h, &n, &bad);
    out->halo = NULL;
    out->n = 0;
    if (rc == -1) {
        set_err(err, errlen, "ERROR: file %s not found in %s().\n", fname, "readHalos");
        return -1;
    }
    if (rc == -2) {
        set_err(err, errlen, "ERROR: file %s does not contain headers in %s().\n", fname,
                "readHalos");
        return -1;
    }
    if (rc == -3) {
        set_err(err, errlen, "ERROR: memory shortage reading %s in %s().\n", fname, "readHalos");
        return -1;
    }
    if (bad < n) {
        free(h);
        set_err(err, errlen, "ERROR: contents in file %s may be incorrect in %s().\n", fname,
                "readHalos");
        return -1;
    }
    out->halo = h;
    out->n = n;
    return 0;
}

/* Filter ids: one per fgets record, strtol, counted as the reference
 * counts them (a last line without '\n' counts), then sorted as u64
 * (BufHalo.c:251-300, 348-380). */
static int read_filter(const char *fname, uint64_t **ids, uint64_t *n)
{
    char *buf = NULL;
    size_t size = 0, pos = 0;
    uint64_t cnt = 0, i;
    char line[SJCS_HALO_LINE_BUF];
    int rc = slurp(fname, &buf, &size, 4);
    if (rc != 0)
        return rc;
    while (pos < size) {
        int eof;
        pos += next_record(buf, size, pos, &eof);
        cnt++;
    }
    *ids = malloc(sizeof(uint64_t) * (cnt ? cnt : 1));
    pos = 0;
    for (i = 0; i < cnt; i++) {
        int eof;
        size_t len = next_record(buf, size, pos, &eof);
        memcpy(line, buf + pos, len);
        line[len] = '\0';
        (*ids)[i] = (uint64_t)strtol(line, NULL, 10);
        pos += len;
    }
    free(buf);
    *n = cnt;
    return 0;
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

int sjcs_read_and_filter_halos(const char *fname, const char *filter_fname, sjcs_halos *out,
                               char *err, size_t errlen)
{
    uint64_t *ids = NULL, nids = 0, n = 0, bad = 0, r, fr = 0, kept = 0;
    sjcs_halo *h = NULL;
    int rc;

    out->halo = NULL;
    out->n = 0;
    if (read_filter(filter_fname, &ids, &nids) != 0) {
        set_err(err, errlen, "ERROR: file %s not found in %s().\n", filter_fname,
                "ConstructFilterIdList");
        return -1;
    }
    qsort(ids, nids, sizeof(uint64_t), cmp_u64);
    rc = parse_file(fname, 16, &h, &n, &bad);
    if (rc == -1 || rc == -2 || rc == -3) {
        free(ids);
        set_err(err, errlen,
                rc == -1 ? "ERROR: file %s not found in %s().\n"
                         : rc == -2 ? "ERROR: file %s does not contain headers in %s().\n"
                                    : "ERROR: memory shortage reading %s in %s().\n",
                fname, "readAndFilterHalos");
        return -1;
    }
    /* merge join (BufHalo.c:165-210): the filter cursor only moves forward
     * past ids smaller than the halo's; a match keeps the halo and leaves
     * the cursor, so repeated halo ids all match */
    for (r = 0; r < n && nids > 0; r++) {
        if (r == bad) {
            free(ids);
            free(h);
            set_err(err, errlen, "ERROR: contents in file %s may be incorrect in %s().\n",
                    fname, "readAndFilterHalos");
            return -1;
        }
        if (h[r].id < ids[fr])
            continue;
        while (h[r].id > ids[fr]) {
            if (++fr >= nids)
                break;
        }
        if (fr >= nids)
            break;
        if (h[r].id == ids[fr])
            h[kept++] = h[r];
    }
    free(ids);
    out->halo = h;
    out->n = kept;
    return 0;
}

void sjcs_free_halos(sjcs_halos *h)
{
    free(h->halo);
    h->halo = NULL;
    h->n = 0;
}
