/*
 * thresholds.c -- the K shell radii of `--radius min:max:K`.
 *
 * Logspace(2, log2(min), log2(max), K) of include/support_functions.h:18-34,
 * then squared in FP32 (src/astr_rcount/main.c:646-649).  The exponent
 * a + i*c is evaluated as fma(i, c, a): that is what gcc -O2 emits for the
 * reference on an FMA-capable host (its default -ffp-contract=fast), and
 * fma() keeps the same rounding however this file is compiled.
 */
#include <math.h>

#include "sjcs_host.h"

void sjcs_host_shell_thresholds(double minrad, double maxrad, int nrads,
                                float *radiuses, float *radiuses_sq)
{
    const double a = log2(minrad);
    const double b = log2(maxrad);
    const double c = (b - a) / (nrads - 1);
    int i;

    for (i = 0; i < nrads - 1; i++)
        radiuses[i] = (float)pow(2.0, fma((double)i, c, a));
    radiuses[nrads - 1] = (float)pow(2.0, b);
    for (i = 0; i < nrads; i++)
        radiuses_sq[i] = radiuses[i] * radiuses[i];
}
