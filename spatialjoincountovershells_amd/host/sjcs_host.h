/*
 * sjcs_host.h -- host side of the drop-in astr_rcount replacement: argv
 * handling, the two input formats, shell thresholds and the text result.
 * Plain C (gnu99), no HIP: it builds into lib/libsjcs_host.so (so the CPU
 * tests can drive it) and into bin/sjcs_rcount.
 */
#ifndef SJCS_HOST_H
#define SJCS_HOST_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#include "sjcs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- argv (src/GetOption/GetOption.c) ---- */

/* Last occurrence of `opt` (exact length match) wins; the value is the
 * argument `offset + 1` places after it.  Returns -1, keeping what earlier
 * occurrences set, when an occurrence has no value (GetOption.c:7-69). */
int sjcs_opt_string(int argc, char **argv, const char *opt, int offset, char **dst);
int sjcs_opt_int(int argc, char **argv, const char *opt, int offset, int *dst);
/* [*beg, *end) = the arguments after the last `opt` up to the next one that
 * starts with '-', if there is at least one (GetOption.c:72-93). */
void sjcs_opt_range(int argc, char **argv, const char *opt, int *beg, int *end);
int sjcs_opt_flag(int argc, char **argv, const char *opt);

#define SJCS_ERRMSG_LEN 1024

typedef struct sjcs_cli {
    char *outfile;          /* --output-file, NULL = stdout */
    int particle_beg, particle_end;
    int compressed;         /* --compressed */
    char *halo_file;        /* --halo-file */
    char *filter_file;      /* --filter-file */
    int nrads;
    double minrad, maxrad;
    char *range_str;
    int maxbacklogs;        /* --backlog */
    int nthreads;           /* --thread */
    int nsetups;            /* --next-setup */
    int first_nbacklogs;    /* --first-setup a:b */
    int first_nsetups;
    char *setup_str;
    int errflag;
    char errmsg[SJCS_ERRMSG_LEN];
} sjcs_cli;

/* Option processing and validation of src/astr_rcount/main.c:492-602,
 * including its first-error-wins message. */
void sjcs_cli_parse(int argc, char **argv, sjcs_cli *cli);
/* The "--execution info--" block (main.c:604-621). */
void sjcs_cli_print_info(FILE *fp, const sjcs_cli *cli, char **argv);
/* The usage text (main.c:118-133). */
void sjcs_cli_print_usage(FILE *fp, const char *argv0);

/* ---- shells (include/support_functions.h:18-34, main.c:645-649) ---- */

void sjcs_host_shell_thresholds(double minrad, double maxrad, int nrads,
                                float *radiuses, float *radiuses_sq);

/* ---- halos (src/Gadget/BufHalo.c) ---- */

/* In-memory halo record, the reference's Halo (include/Gadget/BufHalo.h:
 * 26-30): 24 bytes, point at byte 12. */
typedef struct sjcs_halo {
    uint64_t id;
    float mass;
    float point[3];
} sjcs_halo;

typedef struct sjcs_halos {
    sjcs_halo *halo;
    uint64_t n;
} sjcs_halos;

#define SJCS_HALO_HEADER_LINES 16
#define SJCS_HALO_LINE_BUF (1024 * 16)

/* readHalos (BufHalo.c:76-119): 16 header lines, then one halo per line;
 * a last line without '\n' is dropped.  Parsed in parallel with the same
 * record boundaries fgets(buf, 16384) would give.  Returns 0, or -1 with
 * the reference's message in err. */
int sjcs_read_halos(const char *fname, sjcs_halos *out, int nthreads,
                    char *err, size_t errlen);
/* readAndFilterHalos (BufHalo.c:121-214): keeps the halos whose id is in
 * the filter list (merge join against the sorted list). */
int sjcs_read_and_filter_halos(const char *fname, const char *filter_fname,
                               sjcs_halos *out, char *err, size_t errlen);
/* GetDataFromLineFast (BufHalo.c:218-249) on one NUL-terminated line. */
int sjcs_parse_halo_line(const char *str, uint64_t *id, float *mass, float point[3]);
void sjcs_free_halos(sjcs_halos *h);

/* ---- particles (src/Gadget/BufVector3D.c:285-308) ---- */

/* Number of particles in a decoded particle file: size / 12 (trailing
 * bytes ignored).  -1 when the file cannot be opened. */
int sjcs_particle_count(const char *fname, uint64_t *n);
/* Reads the first n particles (3 x FP32 each) into dst. */
int sjcs_read_particles(const char *fname, float *dst, uint64_t n);
/* bytes [offset, offset + bytes) of a file into dst with nthreads parallel
 * preads.  0, -1 (cannot open), -2 (shorter). */
int sjcs_pread_par(const char *fname, uint64_t offset, uint64_t bytes, void *dst, int nthreads);

/* ---- compressed particles (src/Gadget/BufVector3D.c:111-136) ---- */

/* Reads a compressed particle file: its header into hdr and its npart[1]
 * uint16 triplets into *uv (grown with realloc when *cap, in particles, is
 * smaller; the caller frees it), *n = npart[1].  Returns 0; -1 when the file
 * cannot be opened; -2 when it is shorter than its header says (the
 * reference then decodes whatever the buffer held) or npart[1] < 0; -3 when
 * memory runs out. */
int sjcs_read_compressed(const char *fname, sjcs_gadget_header *hdr, uint16_t **uv,
                         uint64_t *cap, uint64_t *n);
/* The header alone; the triplets start at byte SJCS_GADGET_DATA_OFFSET.
 * 0, -1 (cannot open), -2 (shorter than a header, or npart[1] < 0). */
int sjcs_read_compressed_header(const char *fname, sjcs_gadget_header *hdr);
#define SJCS_GADGET_DATA_OFFSET 268

/* ---- result (src/astr_rcount/main.c:371-442) ---- */

/* Writes the text result byte for byte as outputResult does, including its
 * Total block, which adds halo 0's counts twice (main.c:400,413).
 * fname NULL = stdout.  Returns 0 or -1. */
int sjcs_write_result(const char *fname, const sjcs_halos *h, const uint64_t *counts,
                      const float *radiuses_sq, int nrads, int nthreads);
/* The same with the count rows from a callback: rows(ctx, beg, n) returns
 * the rows of halos [beg, beg + n) (n * nrads u64), or NULL on failure; the
 * writer asks for the next batch while it formats the current one, so a
 * returned block must stay valid until the call after next. */
typedef const uint64_t *(*sjcs_rows_fn)(void *ctx, uint64_t beg, uint64_t n);
int sjcs_write_result_rows(const char *fname, const sjcs_halos *h, sjcs_rows_fn rows, void *ctx,
                           const float *radiuses_sq, int nrads, int nthreads);
/* The same with u32 count rows (half the bytes to fetch and read). */
typedef const uint32_t *(*sjcs_rows32_fn)(void *ctx, uint64_t beg, uint64_t n);
int sjcs_write_result_rows32(const char *fname, const sjcs_halos *h, sjcs_rows32_fn rows,
                             void *ctx, const float *radiuses_sq, int nrads, int nthreads);
/* Halos per batch of the writer (the largest n it asks rows for). */
#define SJCS_RESULT_BATCH 32768

/* milliseconds between two gettimeofday samples, truncated as the
 * reference's printTimeMillis does (include/time.h:30-39) */
long sjcs_elapsed_ms(const void *tv_begin, const void *tv_end);

#ifdef __cplusplus
}
#endif

#endif /* SJCS_HOST_H */
