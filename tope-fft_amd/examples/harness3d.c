/*
 * Drop-in check: the reference's 3D call sequence (test/3d.c:60-117)
 * against include/topefft.h and libtopefft.so.
 *   usage: harness3d NX NY NZ
 * Input: the reference's sin(2 pi count / (NX NY NZ)) in the real parts.
 * Prints "log2(NX) log2(NY) log2(NZ) kernel_s memory_s maxerr" where maxerr
 * is the largest deviation after a FORWARD then INVERSE round trip.
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
    long long i, n = (long long)NX * NY * NZ;
    double PI = acos(-1);
    double *data = calloc((size_t)n * 2, sizeof(double));
    double *orig = calloc((size_t)n * 2, sizeof(double));
    for (i = 0; i < n; i++) {
        data[2 * i] = sin(2 * PI * (double)i / (double)n);
        orig[2 * i] = data[2 * i];
    }

    struct topeFFT framework;
    topeFFTInit(&framework);
    struct topeFFT *f = &framework;
    $CHECKERROR

    struct topePlan3D plan;
    tope3DPlanInit(&framework, &plan, NX, NY, NZ, C2C, data);
    $CHECKERROR
    tope3DExec(&framework, &plan, data, FORWARD);
    $CHECKERROR
    tope3DExec(&framework, &plan, data, INVERSE);
    $CHECKERROR

    double maxerr = 0;
    for (i = 0; i < 2 * n; i++) {
        double e = fabs(data[i] - orig[i]);
        if (e > maxerr) maxerr = e;
    }
    printf("%d\t%d\t%d\t%.9f\t%.9f\t%.3e\n", (int)log2(NX), (int)log2(NY), (int)log2(NZ),
           plan.totalKernel * 1e-9, plan.totalMemory * 1e-9, maxerr);
    tope3DDestroy(&framework, &plan);
    free(data);
    free(orig);
    return maxerr < 1e-9 ? 0 : 2;
}
