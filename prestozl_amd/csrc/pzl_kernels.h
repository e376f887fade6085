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

// A threshold crossing after ordering, as the host consumes it: key orders
// hits exactly as PRESTO C inserts them (rstep, stage, w, z, r) --
// accelsearch.c:183-223 with accel_utils.c:1507-1555; rr is the candidate's
// fundamental r (accel_utils.c:1531) and sigma its float significance.
struct PzlHostHit {
    double rr;
    long long key;
    float sigma;
    float power;
    short zi, wi;  // fundamental z and w indices
    int stage;
};

struct PzlBatchArgs {
    const PzlHarm *harms;
    int nh;                // harmonics (fundamental first, then stage order)
    long long rstep0;      // global index of the batch's first rstep ...
    const long long *rstep0_dev;  // ... or, when set, read from device memory (graph replay)
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
    long long total_items;   // work units: per harmonic of the class, ceil(nrows / rows per workgroup),
                             // times nblk
    int rb, nblk;            // a unit's rows take rsteps [k rb, min(nb, (k + 1) rb)) of the batch,
                             // k < nblk (set by the launcher)
    const float2 *segs;
    long long seg_stride;
    const float2 *bank;
    float *planes;
    long long plane_stride;  // floats per rstep
    const PzlHrTab *hr_tab;
    const float2 *twtab;     // the variant's twiddle table (set by the launcher)
};

struct PzlSumArgs {
    const PzlHarm *harms;
    int nh, nb, nstages;
    long long rstep0;
    const long long *rstep0_dev;  // when set, the batch's first rstep is read from here
    int numz, numw, corr_uselen;
    const float *planes;
    long long plane_stride;
    const uint16_t *rinds;   // [nb][nh][corr_uselen]
    const uint16_t *zinds;   // [nh][numz]
    const int *winds;        // [nh][numw]
    const int *zoff_tab;     // [numw][numz][nh - 1]: byte offset of each subharmonic's
                             // (winds, zinds) row inside one rstep's planes
    float powcut[6];
    long long *hit_keys;     // unordered hits of this batch: keys ...
    float *hit_pows;         // ... and summed powers
    unsigned long long *nhits;
    unsigned long long maxhits;
    float *sums_out;         // optional [nb][numw][numz][corr_uselen]: final summed powers
    const double *startr;    // fundamental startr of every rstep
    const PzlHrTab *hr_tab;  // [nb][nh] subharmonic rlo of the batch
    int int_rlo;             // startr are integers: r maps computed in integer arithmetic
    int order;               // workgroup sweep: 0 (r chunk, z block, w fastest), 1 (z block, w, r chunk fastest)
    int rpt;                 // consecutive fundamental r per thread: 1, 2 or 4
};

// Floats allocated behind the last plane set of a batch buffer, so that the
// harmonic sum's multi-float loads never reach past the buffer range.
#define PZL_PLANE_SLACK 1024

#ifdef __cplusplus
extern "C" {
#endif
hipError_t pzl_launch_bank_build(const PzlHarm *d_harms, int nh, int total_rows, int max_fftlen, float2 *d_bank,
                                 hipStream_t st);
hipError_t pzl_launch_segments(const PzlBatchArgs *a, int max_fftlen, hipStream_t st);
// Correlation of one FFT-length class with variant 'variant' (0 = default);
// class_nrows: host copy of the kernel-row counts of the class's harmonics.
hipError_t pzl_launch_corr(int N, int variant, const PzlCorrArgs *a, const int *class_nrows, hipStream_t st);
int pzl_corr_num_variants(int n);
const char *pzl_corr_variant_name(int n, int v);
// Correlation of several FFT-length classes in one launch of group 'group':
// class c has length fftlens[c], arguments *args[c] and host row counts
// class_nrows[c]; a length the group does not hold, or one given twice, is an
// error.
hipError_t pzl_launch_corr_group(int group, int ncls, const int *fftlens, const PzlCorrArgs *const *args,
                                 const int *const *class_nrows, hipStream_t st);
int pzl_corr_num_groups(void);
const char *pzl_corr_group_name(int group);
int pzl_corr_group_has(int group, int n);
hipError_t pzl_launch_harmsum(const PzlSumArgs *a, hipStream_t st);
// Radix sort of n (key, power) pairs on 'end_bit' key bits (pzl_sort.hip):
// from (keys, pows) into (keys_alt, pows_alt), both pairs overwritten;
// temp == nullptr queries the scratch size into *temp_bytes (host only).
hipError_t pzl_sort_hits(void *temp, size_t *temp_bytes, long long *keys, float *pows, long long *keys_alt,
                         float *pows_alt, long long n, int end_bit, hipStream_t st);
// Sorted (key, power) -> host records with rr and sigma (double precision).
hipError_t pzl_launch_hitprep(const long long *keys, const float *pows, long long n, const double *startr,
                              int corr_uselen, int numz, int numw, int nstages, const double *numindep,
                              PzlHostHit *out, hipStream_t st);
int pzl_corr_supported(int n);
hipError_t pzl_launch_add_plane(float *fund, int numws, int numzs, int numrs, const float *sub, int sub_numzs,
                                int sub_numrs, const int *winds, const uint16_t *zinds, const uint16_t *rinds,
                                hipStream_t st);
hipError_t pzl_launch_threshold(const float *plane, long long n, float powcut, long long *idx,
                                unsigned long long *count, long long cap, hipStream_t st);
#ifdef __cplusplus
}
#endif
