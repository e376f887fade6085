// pzl_kernels.h -- argument blocks of the search kernels (host and device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pzl_geom.h"

// Per-(rstep in batch, harmonic) r geometry written by k_segments.
struct PzlHrTab {
    long long rlo;
    int numrs;
    int pad;
};

// A threshold crossing: key orders hits exactly as PRESTO C inserts them
// (rstep, stage, w, z, r) -- accelsearch.c:183-223 with accel_utils.c:1507-1555.
struct PzlHit {
    long long key;
    float power;
    int pad;
};

struct PzlBatchArgs {
    const PzlHarm *harms;
    int nh;                // harmonics (fundamental first, then stage order)
    long long rstep0;      // global index of the batch's first rstep
    int nb;                // rsteps in this batch
    const double *startr;  // fundamental startr of every rstep of the search
    int rstep_bins;        // corr_uselen / 2
    int corr_uselen;
    const float2 *fft;     // the spectrum (complex64), resident in HBM
    long long numbins, file_lobin;
    double nph;            // >0: photon normalisation
    float2 *segs;          // [nb][seg_stride] segment spectra
    long long seg_stride;
    uint16_t *rinds;       // [nb][nh][corr_uselen]
    PzlHrTab *hr_tab;      // [nb][nh]
};

struct PzlCorrArgs {
    const PzlHarm *harms;
    int nh, nb;
    const int *class_harms;  // harmonics of one FFT length
    int nclass;
    long long total_items;   // sum over the class of nrows * nb
    const float2 *segs;
    long long seg_stride;
    const float2 *bank;
    float *planes;
    long long plane_stride;  // floats per rstep
    const PzlHrTab *hr_tab;
};

struct PzlSumArgs {
    const PzlHarm *harms;
    int nh, nb, nstages;
    long long rstep0;
    int numz, numw, corr_uselen;
    const float *planes;
    long long plane_stride;
    const uint16_t *rinds;   // [nb][nh][corr_uselen]
    const uint16_t *zinds;   // [nh][numz]
    const int *winds;        // [nh][numw]
    float powcut[6];
    PzlHit *hits;
    unsigned long long *nhits;
    unsigned long long maxhits;
    float *sums_out;         // optional [nb][numw][numz][corr_uselen]: final summed powers
    unsigned int *map_errors;  // count of index-map entries outside a plane (must stay 0)
};

#ifdef __cplusplus
extern "C" {
#endif
hipError_t pzl_launch_bank_build(const PzlHarm *d_harms, int nh, int total_rows, int max_fftlen, float2 *d_bank,
                                 hipStream_t st);
hipError_t pzl_launch_segments(const PzlBatchArgs *a, int max_fftlen, hipStream_t st);
hipError_t pzl_launch_corr(int N, const PzlCorrArgs *a, int nblocks, hipStream_t st);
hipError_t pzl_launch_harmsum(const PzlSumArgs *a, hipStream_t st);
int pzl_corr_supported(int n);
#ifdef __cplusplus
}
#endif
