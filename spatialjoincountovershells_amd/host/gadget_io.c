/*
 * gadget_io.c -- the compressed (Gadget) particle file: a 4-byte record
 * marker, the 256-byte header, 8 bytes of record markers, then npart[1]
 * triplets of uint16 (readCompressedParticles, src/Gadget/BufVector3D.c:
 * 111-136).  Record markers are skipped, not checked, as there.
 */
#include <stdio.h>
#include <stdlib.h>

#include "sjcs_host.h"

int sjcs_read_compressed_header(const char *fname, sjcs_gadget_header *hdr)
{
    FILE *fp = fopen(fname, "r");
    int rc = 0;
    if (fp == NULL)
        return -1;
    if (fseek(fp, 4, SEEK_CUR) != 0 || fread(hdr, sizeof(*hdr), 1, fp) != 1 || hdr->npart[1] < 0)
        rc = -2;
    fclose(fp);
    return rc;
}

int sjcs_read_compressed(const char *fname, sjcs_gadget_header *hdr, uint16_t **uv,
                         uint64_t *cap, uint64_t *n)
{
    FILE *fp = fopen(fname, "r");
    uint64_t np;
    if (fp == NULL)
        return -1;
    if (fseek(fp, 4, SEEK_CUR) != 0 || fread(hdr, sizeof(*hdr), 1, fp) != 1 ||
        fseek(fp, 8, SEEK_CUR) != 0) {
        fclose(fp);
        return -2;
    }
    if (hdr->npart[1] < 0) {
        fclose(fp);
        return -2;
    }
    np = (uint64_t)hdr->npart[1];
    if (np > *cap) {
        uint16_t *p = realloc(*uv, 6 * (size_t)np);
        if (p == NULL) {
            fclose(fp);
            return -3;
        }
        *uv = p;
        *cap = np;
    }
    if (np && fread(*uv, 6, (size_t)np, fp) != (size_t)np) {
        fclose(fp);
        return -2;
    }
    fclose(fp);
    *n = np;
    return 0;
}
