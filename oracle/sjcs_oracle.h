/*
 * sjcs_oracle.h -- CPU restatement of astr_rcount's shell count and of
 * rt_rcount's R-tree index and count.
 * TEST INFRASTRUCTURE ONLY: the checker, never linked by the product.
 */
#ifndef SJCS_ORACLE_H
#define SJCS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

void oracle_shell_thresholds(double minrad, double maxrad, int nrads,
                             float *radiuses, float *r2);
void oracle_shell_thresholds_nofma(double minrad, double maxrad, int nrads,
                                   float *radiuses, float *r2);
float oracle_box_radius(const float *r2, int nrads);
float oracle_distance(const float p[3], const float c[3]);
int oracle_shell_of(float dist, const float *r2, int nrads);
int oracle_images(const float c[3], float rbox, float centres[8][3]);
uint64_t oracle_count(const float *particles, uint64_t np,
                      const void *halo_points, size_t stride_bytes, uint64_t nh,
                      const float *r2, int nrads, uint64_t *counts,
                      int nthreads);

void oracle_gadget_params(const unsigned char *hdr, int ntile, long long *id_start,
                          long long *ng1d, int *nt1d, float *cell_interval,
                          float *disp_min, double *unitsep);
void oracle_gadget_gridpos(long long id, long long ng1d, int nt1d, int ip[3]);
void oracle_gadget_decode(const uint16_t *uv, uint64_t n, long long id_start, long long ng1d,
                          int nt1d, float cell_interval, float disp_min, double unitsep,
                          int fused, float *xyz);

/* rt_rcount (sjcs_oracle_rt.c): the Guttman R-tree built by inserting
 * xyz[3n] in order, its leaf grouping, and its periodic shell count
 * (counts[h*nrads + k] += ...).  oracle_rt_build returns NULL when the
 * reference's split would read outside its entries (NaN areas) or memory
 * runs out. */
typedef struct oracle_rtree oracle_rtree;
oracle_rtree *oracle_rt_build(const float *xyz, uint64_t n);
void oracle_rt_free(oracle_rtree *t);
int oracle_rt_height(const oracle_rtree *t);
uint64_t oracle_rt_leaves(const oracle_rtree *t, uint64_t *ids, uint32_t *leaf_len);
void oracle_rt_count(const oracle_rtree *t, const void *halo_points, size_t stride_bytes,
                     uint64_t nh, const float *r2, int nrads, uint64_t *counts);

#endif
