/*
 * Drop-in check: the reference's call sequence (test/2d.c:73-114) compiled
 * unchanged in shape against include/topefft.h and libtopefft.so.
 *   usage: harness2d NX NY
 * Prints "log2(NX) log2(NY) kernel_s memory_s maxerr" where maxerr is the
 * largest deviation after a FORWARD then INVERSE round trip.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "topefft.h"

int main(int argc, char *argv[])
{
    if (argc != 3) {
        printf("<exec> NX NY\n");
        return 1;
    }
    int NX = atoi(argv[1]);
    int NY = atoi(argv[2]);
    int i, j, count = 0;
    double PI = acos(-1);
    double *data = calloc((size_t)NX * NY * 2, sizeof(double));
    double *orig = calloc((size_t)NX * NY * 2, sizeof(double));
    for (j = 0; j < NY; j++) {
        for (i = 0; i < NX; i++) {
            data[2 * (j * NX + i)] = sin(2 * PI * count / (NX * NY));
            orig[2 * (j * NX + i)] = data[2 * (j * NX + i)];
            count++;
        }
    }

    struct topeFFT framework;
    topeFFTInit(&framework);
    struct topeFFT *f = &framework;
    $CHECKERROR

    struct topePlan2D plan;
    tope2DPlanInit(&framework, &plan, NX, NY, C2C, data);
    $CHECKERROR
    tope2DExec(&framework, &plan, data, FORWARD);
    $CHECKERROR
    tope2DExec(&framework, &plan, data, INVERSE);
    $CHECKERROR

    double maxerr = 0;
    for (i = 0; i < 2 * NX * NY; i++) {
        double e = fabs(data[i] - orig[i]);
        if (e > maxerr) maxerr = e;
    }
    printf("%d\t%d\t%.9f\t%.9f\t%.3e\n", (int)log2(NX), (int)log2(NY), plan.totalKernel * 1e-9,
           plan.totalMemory * 1e-9, maxerr);
    tope2DDestroy(&framework, &plan);
    free(data);
    free(orig);
    return maxerr < 1e-9 ? 0 : 2;
}
