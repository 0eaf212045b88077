/* Exhaustive check of the division the uint16 amplitude path uses
 * (div_u16 in fused-fourier-ptycho_amd/csrc/fpm_kernels.h): for every
 * integer numerator 0 <= a < 2^16 and denominator 1 <= b < 2^16,
 *     q = RN(a * RN(1/b));  r = fma(-q, b, a);  fma(r, RN(1/b), q)
 * must equal the IEEE single-precision quotient a / b bit for bit, the
 * quotient MATLAB forms in single for (I - min)/(max - min).
 * Built with -ffp-contract=off so that only the explicit fmaf calls fuse.
 * Prints the number of mismatches; exit status 0 when there are none. */
#include <math.h>
#include <stdio.h>

int main(void) {
    long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(dynamic, 64)
    for (int b = 1; b < 65536; ++b) {
        volatile float bv = (float)b;
        const float bf = bv;
        const float rb = 1.0f / bf;
        for (int a = 0; a < 65536; ++a) {
            const float af = (float)a;
            const float q = af * rb;
            const float r = fmaf(-q, bf, af);
            if (fmaf(r, rb, q) != af / bf) bad++;
        }
    }
    printf("%lld\n", bad);
    return bad != 0;
}
