/*
 * Real-input transforms through the C ABI (topeRealPlanInit/Exec/Destroy).
 *   usage: harness_real NX NY NZ
 * Input: cos(2 pi 3 x / NX) (so X[3] = NX/2 * NY * NZ at ky = kz = 0 and
 * the rest of the half spectrum is zero).  Prints "NX NY NZ kernel_s
 * peak_err maxerr": peak_err checks that closed form over the whole half
 * spectrum, maxerr is the largest deviation after R2C then C2R.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "topefft.h"

int main(int argc, char *argv[])
{
    if (argc != 4) {
        printf("<exec> NX NY NZ\n");
        return 1;
    }
    int NX = atoi(argv[1]);
    int NY = atoi(argv[2]);
    int NZ = atoi(argv[3]);
    long long i, n = (long long)NX * NY * NZ, h = (long long)(NX / 2 + 1) * NY * NZ;
    double PI = acos(-1);
    double *x = calloc((size_t)n, sizeof(double));
    double *spec = calloc((size_t)h * 2, sizeof(double));
    double *back = calloc((size_t)n, sizeof(double));
    for (i = 0; i < n; i++) x[i] = cos(2 * PI * 3.0 * (double)(i % NX) / NX);

    struct topeFFT framework;
    topeFFTInit(&framework);
    struct topeFFT *f = &framework;
    $CHECKERROR
    struct topePlanReal plan;
    topeRealPlanInit(&framework, &plan, NX, NY, NZ, x);
    $CHECKERROR
    topeRealExec(&framework, &plan, spec, FORWARD);
    $CHECKERROR
    const double peak = 0.5 * (double)NX * NY * NZ;
    double peak_err = 0;
    for (i = 0; i < h; i++) {
        const double want = (NX >= 8 && i == 3) ? peak : 0.0;
        const double e = (fabs(spec[2 * i] - want) + fabs(spec[2 * i + 1])) / peak;
        if (e > peak_err) peak_err = e;
    }
    topeRealExec(&framework, &plan, back, INVERSE);
    $CHECKERROR
    double maxerr = 0;
    for (i = 0; i < n; i++) {
        double e = fabs(back[i] - x[i]);
        if (e > maxerr) maxerr = e;
    }
    printf("%d\t%d\t%d\t%.9f\t%.3e\t%.3e\n", NX, NY, NZ, plan.totalKernel * 1e-9, peak_err, maxerr);
    topeRealDestroy(&framework, &plan);
    free(x);
    free(spec);
    free(back);
    return (NX < 8 || peak_err < 1e-12) && maxerr < 1e-12 ? 0 : 2;
}
