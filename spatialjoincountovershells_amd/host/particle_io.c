/*
 * particle_io.c -- the decoded particle file: packed little-endian FP32
 * x, y, z records, count = file size / 12, trailing bytes ignored
 * (src/Gadget/BufVector3D.c:285-308, 391-400).
 */
#include <fcntl.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/stat.h>
#include <unistd.h>

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

typedef struct {
    int fd;
    uint64_t off, len;
    char *dst;
    int rc;
} pread_job;

static void *pread_run(void *arg)
{
    pread_job *j = arg;
    uint64_t done = 0;
    while (done < j->len) {
        ssize_t got = pread(j->fd, j->dst + done, (size_t)(j->len - done), (off_t)(j->off + done));
        if (got <= 0) {
            j->rc = -2;
            return NULL;
        }
        done += (uint64_t)got;
    }
    return NULL;
}

int sjcs_pread_par(const char *fname, uint64_t offset, uint64_t bytes, void *dst, int nthreads)
{
    pread_job jobs[64];
    pthread_t th[64];
    int created[64];
    int fd, t, rc = 0;
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > 64)
        nthreads = 64;
    if (bytes < ((uint64_t)nthreads << 20))  /* at least 1 MiB a thread */
        nthreads = (int)(bytes >> 20) + 1;
    fd = open(fname, O_RDONLY);
    if (fd < 0)
        return -1;
    for (t = 0; t < nthreads; t++) {
        /* slices on 4 KiB boundaries */
        uint64_t a = (bytes * (uint64_t)t / (uint64_t)nthreads) & ~(uint64_t)4095;
        uint64_t b = t + 1 == nthreads
                         ? bytes
                         : (bytes * (uint64_t)(t + 1) / (uint64_t)nthreads) & ~(uint64_t)4095;
        jobs[t].fd = fd;
        jobs[t].off = offset + a;
        jobs[t].len = b - a;
        jobs[t].dst = (char *)dst + a;
        jobs[t].rc = 0;
        /* the last slice runs on this thread, as does any whose thread
         * could not be started */
        created[t] = t + 1 < nthreads && pthread_create(&th[t], NULL, pread_run, &jobs[t]) == 0;
        if (!created[t])
            pread_run(&jobs[t]);
    }
    for (t = 0; t < nthreads; t++)
        if (created[t])
            pthread_join(th[t], NULL);
    for (t = 0; t < nthreads; t++)
        if (jobs[t].rc)
            rc = jobs[t].rc;
    close(fd);
    return rc;
}
