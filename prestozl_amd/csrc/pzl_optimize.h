// pzl_optimize.h -- work descriptions of the candidate optimisation kernels
// (max_rzw_arr_harmonics, maximize_rzw.c:85-191).
#pragma once

#include <hip/hip_runtime.h>

#define PZL_OPT_INTERPFAC 4  // points beyond ACCEL_DR/DZ/DW (maximize_rzw.c:97)
#define PZL_OPT_NUMWID 2     // full peak widths searched in z and w (:98)
#define PZL_OPT_NB 8         // numbetween = interpfac * ACCEL_RDR (:99)
#define PZL_OPT_NUMR 24      // 3 full Fourier bins (:100)
#define PZL_OPT_NUMZ 17      // 2 numwid interpfac + 1 (:101)
#define PZL_OPT_NUMW 17
#define PZL_OPT_MAX_FFTLEN 4096
#define PZL_OPT_TAIL 8       // last powers of each volume row kept for negative r indices

// One Fourier interpolation at (r, z, w) with kernel half-width khw.
struct PzlInterpJob {
    double r, z, w;
    int khw, pad;
};

// One harmonic of one candidate: its oversampled volume.
struct PzlVolJob {
    int lobin;        // first spectrum bin of the volume (corr_rzw_vol startbin)
    int fftlen;
    int khw;          // kernel half-width from the volume's extreme z and w
    int numgoodbins;
    double zlo, dz, wlo, dw;
    int rind[PZL_OPT_NUMR];  // volume-row index of each fine r step of the sum (>= -PZL_OPT_TAIL)
};

#ifdef __cplusplus
extern "C" {
#endif
hipError_t pzl_launch_opt_interp(const float2 *data, int numdata, const PzlInterpJob *jobs, int njobs,
                                 float2 *ans, hipStream_t st);
hipError_t pzl_launch_opt_volume(const float2 *data, int numdata, const PzlVolJob *jobs, int njobs,
                                 int max_fftlen, double2 *dfft, float *powv, float *tail, hipStream_t st);
hipError_t pzl_launch_opt_sum(const float *powv, const float *tail, const PzlVolJob *jobs, const float *locpow,
                              const int *job0, const int *nharm, int ncand, int *best, hipStream_t st);
#ifdef __cplusplus
}
#endif
