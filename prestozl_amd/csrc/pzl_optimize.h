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
#define PZL_OPT_RESP_FFT_MAX 4096  // chirp samples of a volume row's w response (and numkern)
#define PZL_OPT_TWN 65536    // twiddle table length (multiple of 8 PZL_OPT_RESP_FFT_MAX)

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
// tw[j] = exp(-2 pi i j / PZL_OPT_TWN), j < PZL_OPT_TWN
hipError_t pzl_launch_opt_twiddles(double2 *tw, hipStream_t st);
// ldsn (<= PZL_OPT_RESP_FFT_MAX): at least the largest numkern (16 khw) of the
// jobs and the largest chirp length of their rows' w responses; a row that
// needs more ORs 1 into *err.  Each job's grid has nz z by nw w rows
// (PZL_OPT_NUMZ x PZL_OPT_NUMW for the harmonic sum, 9 x 9 for one harmonic's
// max_rzw_arr); powv holds [job][w][z][PZL_OPT_NUMR], tail [job][w][z][PZL_OPT_TAIL].
hipError_t pzl_launch_opt_volume(const float2 *data, int numdata, const PzlVolJob *jobs, int njobs, int ldsn,
                                 int nz, int nw, const double2 *tw, float *powv, float *tail, int *err,
                                 hipStream_t st);
// best[c]: the flat (w, z, r) index of candidate c's maximum; bestv[c] (bestv
// may be null) its value
hipError_t pzl_launch_opt_sum(const float *powv, const float *tail, const PzlVolJob *jobs, const float *locpow,
                              const int *job0, const int *nharm, int ncand, int nzw, int *best, float *bestv,
                              hipStream_t st);
#ifdef __cplusplus
}
#endif
