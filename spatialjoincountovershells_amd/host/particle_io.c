/*
 * particle_io.c -- the decoded particle file: packed little-endian FP32
 * x, y, z records, count = file size / 12, trailing bytes ignored
 * (src/Gadget/BufVector3D.c:285-308, 391-400).
 */
#include <stdio.h>
#include <sys/stat.h>

#include "sjcs_host.h"

int sjcs_particle_count(const char *fname, uint64_t *n)
{
    struct stat st;
    FILE *fp = fopen(fname, "r");
    if (fp == NULL)
        return -1;
    if (fstat(fileno(fp), &st) != 0) {
        fclose(fp);
        return -1;
    }
    fclose(fp);
    *n = (uint64_t)st.st_size / 12u;
    return 0;
}

int sjcs_read_particles(const char *fname, float *dst, uint64_t n)
{
    FILE *fp = fopen(fname, "r");
    size_t got;
    if (fp == NULL)
        return -1;
    got = fread(dst, 12, (size_t)n, fp);
    fclose(fp);
    return got == (size_t)n ? 0 : -1;
}
