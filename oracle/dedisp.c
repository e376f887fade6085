/*
 * dedisp.c -- the oracle's restatement of prepsubband's GPU dedispersion
 * (prepsubband_kernels.cu:26-77).  TEST INFRASTRUCTURE ONLY.
 *
 * oracle_downsamp   DownsampDataArray_kernel: ds[t][s] = (sum_{j < bs}
 *                   in[t bs + j][s]) / bs, float, j ascending
 * oracle_dedisp     getDedispGPU: out[dm][x] = sum_s (x < worklen - off ?
 *                   last[x + off][s] : cur[x - (worklen - off)][s]),
 *                   off = offsets[dm][s], float, s ascending
 */
#include "oracle.h"

void oracle_downsamp(float *out, const float *in, int worklen, int bs, int nsub)
{
    for (long t = 0; t < worklen; t++)
        for (int s = 0; s < nsub; s++) {
            float v = 0.0f;
            for (int j = 0; j < bs; j++) v += in[(t * bs + j) * nsub + s];
            out[t * nsub + s] = v / (float)bs;
        }
}

void oracle_dedisp(const float *cur, const float *last, float *out, int worklen, int nsub, int numdms,
                   const int *offsets)
{
    for (int dm = 0; dm < numdms; dm++)
        for (long x = 0; x < worklen; x++) {
            float v = 0.0f;
            for (int s = 0; s < nsub; s++) {
                const int off = offsets[dm * nsub + s];
                if (x < worklen - off) v += last[(x + off) * nsub + s];
                else v += cur[(x - (worklen - off)) * nsub + s];
            }
            out[(long)dm * worklen + x] = v;
        }
}
