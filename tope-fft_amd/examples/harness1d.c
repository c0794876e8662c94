/*
 * Drop-in check: the reference's 1D call sequence (test/1d.c:60-79) against
 * include/topefft.h and libtopefft.so.
 *   usage: harness1d N
 * Input x[i] = i + 1 (the reference's ramp).  Prints
 * "log2(N) plan+kernel_s memory_s dc_err maxerr": dc_err compares X[0] with
 * the closed form N(N+1)/2, maxerr is the largest deviation after a FORWARD
 * then INVERSE round trip.  Any N >= 1 (non-powers of two included).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "topefft.h"

int main(int argc, char *argv[])
{
    if (argc != 2) {
        printf("<exec> N\n");
        return 1;
    }
    int N = atoi(argv[1]);
    int i;
    double *data = calloc((size_t)N * 2, sizeof(double));
    for (i = 0; i < N; i++) data[2 * i] = i + 1;

    struct topeFFT framework;
    topeFFTInit(&framework);
    struct topeFFT *f = &framework;
    $CHECKERROR

    struct topePlan1D plan;
    tope1DPlanInit(&framework, &plan, N, C2C, data);
    $CHECKERROR
    tope1DExec(&framework, &plan, data, FORWARD);
    $CHECKERROR
    const double dc = 0.5 * (double)N * (double)(N + 1);
    const double dc_err = fabs(data[0] - dc) / dc + fabs(data[1]) / dc;
    tope1DExec(&framework, &plan, data, INVERSE);
    $CHECKERROR

    double maxerr = 0;
    for (i = 0; i < N; i++) {
        double e = fabs(data[2 * i] - (i + 1)) + fabs(data[2 * i + 1]);
        if (e > maxerr) maxerr = e;
    }
    printf("%.3f\t%.9f\t%.9f\t%.3e\t%.3e\n", log2((double)N), 1.0e-9 * (plan.totalKernel + plan.totalPreKernel),
           plan.totalMemory * 1e-9, dc_err, maxerr / N);
    tope1DDestroy(&framework, &plan);
    free(data);
    return (dc_err < 1e-12 && maxerr / N < 1e-12) ? 0 : 2;
}
