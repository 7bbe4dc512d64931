/*
 * dpps_probe.c -- checks oracle_distance() against SSE4.1 DPPS (imm 0x77),
 * the instruction the reference computes its distance with
 * (include/objects/RTree/Distance.h:11-13), on the CPU this runs on.
 *
 * TEST INFRASTRUCTURE ONLY.  Usage: dpps_probe [n_random] [seed]
 * Prints "dpps ok <n>" and exits 0 when every result matches bit for bit
 * (all NaNs counted equal), otherwise prints the first mismatches, exits 1.
 */
#include <math.h>
#include <smmintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sjcs_oracle.h"

static uint64_t state;

static uint32_t next_u32(void)
{
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return (uint32_t)(state >> 11);
}

static float from_bits(uint32_t b)
{
    float f;
    memcpy(&f, &b, 4);
    return f;
}

static uint32_t to_bits(float f)
{
    uint32_t b;
    memcpy(&b, &f, 4);
    return b;
}

/* A value drawn to hit the interesting parts of FP32: plain box
 * coordinates, values a few ulps apart, subnormals, huge and special. */
static float draw(int kind)
{
    switch (kind % 6) {
    case 0: return (float)(next_u32() % 1000000) / 1000.0f;
    case 1: return from_bits(next_u32());
    case 2: return from_bits(next_u32() & 0x807fffffu); /* subnormal / zero */
    case 3: return 500.0f + (float)((int)(next_u32() % 64) - 32) * 6.103515625e-05f;
    case 4: return from_bits(0x7f000000u | (next_u32() & 0x80ffffffu)); /* huge */
    default: {
        static const float special[] = {0.0f, -0.0f, 1000.0f, INFINITY,
                                        -INFINITY, NAN, 1e-30f, 3e38f};
        return special[next_u32() % 8];
    }
    }
}

static float dpps_distance(const float p[3], const float c[3])
{
    float a[4] __attribute__((aligned(16))) = {p[0], p[1], p[2], 0.0f};
    float b[4] __attribute__((aligned(16))) = {c[0], c[1], c[2], 0.0f};
    __m128 d = _mm_sub_ps(_mm_load_ps(a), _mm_load_ps(b));
    return _mm_cvtss_f32(_mm_dp_ps(d, d, 0x77));
}

int main(int argc, char **argv)
{
    long n = argc > 1 ? atol(argv[1]) : 2000000;
    long i, bad = 0;
    state = argc > 2 ? strtoull(argv[2], NULL, 10) : 88172645463325252ull;
    if (state == 0)
        state = 1;
    for (i = 0; i < n; i++) {
        float p[3], c[3], g, o;
        int d, kind = (int)(next_u32() % 6);
        for (d = 0; d < 3; d++) {
            c[d] = draw(kind);
            /* most draws are nearby pairs, like the count's candidates */
            p[d] = (next_u32() & 3) ? c[d] + draw(0) * 0.01f - 5.0f : draw(kind);
        }
        g = dpps_distance(p, c);
        o = oracle_distance(p, c);
        if (!(isnan(g) && isnan(o)) && to_bits(g) != to_bits(o)) {
            if (bad < 10)
                fprintf(stderr, "mismatch p=(%a %a %a) c=(%a %a %a) dpps=%a oracle=%a\n",
                        p[0], p[1], p[2], c[0], c[1], c[2], g, o);
            bad++;
        }
    }
    if (bad) {
        printf("dpps mismatches %ld of %ld\n", bad, n);
        return 1;
    }
    printf("dpps ok %ld\n", n);
    return 0;
}
