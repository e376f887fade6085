/*
 * oracle.h -- CPU restatement of the PRESTO C jerk-search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in prestozl_amd/: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it.  The product never links it.
 *
 * Every function restates the reference file:line given beside it
 * (paths relative to the PrestoZL source tree).  Two arithmetic modes:
 *   ORACLE_FP32 -- float data, float FFTs (double twiddles rounded to float),
 *                  double where the reference uses double: the reference's
 *                  own arithmetic with our FFT in place of FFTW.
 *   ORACLE_FP64 -- the same algorithm with every float replaced by double:
 *                  the "exact" correlation-and-sum oracle.
 *
 * Parity status per function is recorded in oracle/README.md.
 */
#ifndef PZL_ORACLE_H
#define PZL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_FP32 = 0, ORACLE_FP64 = 1 };

/* ---------------- FFT (stands where FFTW stands in fftcalls.c:71) ------- */
/* sign -1 = forward (e^{-2 pi i jk/n}), +1 = unnormalised inverse. */
void oracle_cfft_f(float *data, int n, int sign);
void oracle_cfft_d(double *data, int n, int sign);
/* realfft, fastffts.c:163-235, NR packing: forward (isign -1) and inverse
 * (isign +1, scaled by 2/n). */
void oracle_realfft_fwd_f(float *data, long n);
void oracle_realfft_fwd_d(double *data, long n);
void oracle_realfft_inv_f(float *data, long n);
void oracle_realfft_inv_d(double *data, long n);

/* ---------------- dedispersion (prepsubband_kernels.cu:26-77) ----------- */
void oracle_downsamp(float *out, const float *in, int worklen, int bs, int nsub);
void oracle_dedisp(const float *cur, const float *last, float *out, int worklen, int nsub, int numdms,
                   const int *offsets);

/* ---------------- sizing helpers ---------------------------------------- */
int oracle_next2_to_n(int x);
int oracle_next_good_fftlen(int n);                 /* corr_prep.c:3-16 */
int oracle_fftlen_from_kernwidth(int kernwidth);    /* corr_prep.c:18-35 */
int oracle_z_resp_halfwidth(double z);              /* responses.c:27-47 (LOWACC) */
int oracle_w_resp_halfwidth(double z, double w);    /* responses.c:49-84 (LOWACC) */

/* ---------------- responses (interleaved re,im) -------------------------- */
/* responses.c:159-225, 228-316, 333-418.  Output numkern complex values.  */
void oracle_gen_r_response(double roffset, int numbetween, int numkern, int mode, void *out);
void oracle_gen_z_response(double roffset, int numbetween, double z, int numkern, int mode, void *out);
void oracle_gen_w_response(double roffset, int numbetween, double z, double w, int numkern,
                           int mode, void *out);
/* Fresnel integrals S(x), C(x) with the pi/2 convention (fresnl.c:176). */
void oracle_fresnel(double x, double *s, double *c);

/* ---------------- candidate optimisation (optimize.c) -------------------- */
/* rderivs and fourierprops, presto.h:161-189 (same layout). */
typedef struct {
    double pow, phs, dpow, dphs, d2pow, d2phs, locpow;
} OracleRDerivs;
typedef struct {
    double r;
    float rerr;
    double z;
    float zerr;
    double w;
    float werr;
    float pow, powerr, sig, rawpow, phs, phserr, cen, cenerr, pur, purerr, locpow;
} OracleFourierProps;
int oracle_z_resp_halfwidth_acc(double z, int highacc);            /* responses.c:27-47 */
int oracle_w_resp_halfwidth_acc(double z, double w, int highacc);  /* responses.c:49-84 */
/* data: interleaved float complex bins; ans: re, im */
void oracle_rz_interp(const float *data, int numdata, double r, double z, int kern_half_width, float *ans);
void oracle_rzw_interp(const float *data, int numdata, double r, double z, double w, int kern_half_width,
                       float *ans);                                 /* rzwinterp.c:145-229 */
double oracle_get_localpower3d(const float *data, int numdata, double r, double z, double w); /* :75-136 */
void oracle_get_derivs3d(const float *data, int numdata, double r, double z, double w, double localpower,
                         OracleRDerivs *result);                    /* characteristics.c:138-191 */
int oracle_corr_rzw_vol(const float *data, int numdata, int numbetween, int startbin, double zlo, double zhi,
                        int numz, double wlo, double whi, int numw, int fftlen, int *nextbin,
                        float *out);                                /* rzwinterp.c:3-140 */
void oracle_max_rzw_arr_harmonics(const float *data, int numdata, int num_harmonics, double rin, double zin,
                                  double win, double *rout, double *zout, double *wout, OracleRDerivs *derivs,
                                  double *powers);                 /* maximize_rzw.c:85-191 */
void oracle_max_rz_arr_harmonics(const float *data, int numdata, int num_harmonics, double rin, double zin,
                                 double *rout, double *zout, OracleRDerivs *derivs,
                                 double *powers);
/* max_rzw_arr (maximize_rzw.c:3-50) and max_rz_arr (maximize_rz.c:25-85): one
 * harmonic each; both return the maximum's raw power */
double oracle_max_rzw_arr(const float *data, int numdata, double rin, double zin, double win, double *rout,
                          double *zout, double *wout, OracleRDerivs *derivs);
double oracle_max_rz_arr(const float *data, int numdata, double rin, double zin, double *rout, double *zout,
                         OracleRDerivs *derivs);                  /* maximize_rz.c:141-220, amoeba.c */
int oracle_calc_props(const OracleRDerivs *d, double r, double z, double w,
                      OracleFourierProps *res);                     /* characteristics.c:193-265 */

/* ---------------- statistics -------------------------------------------- */
double oracle_power_for_sigma(double sigma, int numsum, double numtrials); /* characteristics.c:670 */
double oracle_candidate_sigma(double power, int numsum, double numtrials); /* characteristics.c:647 */
double oracle_chi2_logp2(double chi2, double dof);                          /* characteristics.c:588 */
double oracle_equivalent_gaussian_sigma2(double logp);                      /* characteristics.c:501 */
double oracle_normal_sf(double x);
float oracle_median(const float *arr, int n);                               /* median.c:12 */

/* ---------------- the search -------------------------------------------- */
typedef struct oracle_params {
    double N;           /* samples in the time series (.inf N) */
    double dt;          /* sample time (.inf dt) */
    long long numbins;  /* complex bins in the .fft */
    long long lobin;    /* -lobin (0) */
    int zmax, wmax;     /* as given on the command line */
    int wmaxP;          /* 1 if -wmax was given */
    int numharm;        /* 1,2,4,8,16,32 */
    float sigma;
    int floP, fhiP, rloP, rhiP;
    double flo, fhi, rlo, rhi;
    double nph;         /* >0: photon normalisation */
} oracle_params;

typedef struct oracle_obs {
    long long N, numbins, lobin, highestbin;
    int maxkernlen, corr_uselen, fftlen, numharmstages, numz, numw;
    double dt, T, rlo, rhi, zlo, zhi, wlo, whi;
    float sigma;
    float powcut[6];
    long long numindep[6];
    double nph;
    int zmax_cmd, wmax_cmd;
} oracle_obs;

/* accel_utils.c:1636-2037 (geometry and thresholds only) */
int oracle_create_accelobs(const oracle_params *p, oracle_obs *obs);

/* One harmonic's shape (accel_utils.c:137-149, 199-234). */
typedef struct oracle_harm_info {
    int numharm, harmnum;
    int zmax, wmax, numz, numw, fftlen, halfwidth;
} oracle_harm_info;
int oracle_harm_info_get(const oracle_obs *obs, int numharm, int harmnum, oracle_harm_info *out);

/* The FFT'd kernel of one (z,w) for a harmonic, accel_utils.c:151-168.
 * out: fftlen complex (float or double per mode). */
int oracle_kernel(const oracle_obs *obs, int numharm, int harmnum, int zi, int wi, int mode, void *out);

/* Search context: owns the kernel banks. */
typedef struct oracle_ctx oracle_ctx;
oracle_ctx *oracle_ctx_create(const oracle_obs *obs, int mode, int nthreads);
void oracle_ctx_free(oracle_ctx *ctx);

/* subharm_fderivs_vol, accel_utils.c:1094-1293.  Writes numws*numzs*numrs
 * powers ([w][z][r]) into pows (float or double per mode) and the index
 * tables (rinds: corr_uselen, zinds: numz) when numharm > 1.  Returns numrs,
 * <0 on error.  meta[0]=rlo, meta[1]=zlo, meta[2]=wlo, meta[3]=numzs, meta[4]=numws. */
int oracle_subharm_fderivs_vol(oracle_ctx *ctx, const float *fft, int numharm, int harmnum,
                               double fullrlo, double fullrhi, void *pows,
                               uint16_t *rinds, uint16_t *zinds, long long *meta);

/* Candidate record (accel.h:84-97 fields that the search fills). */
typedef struct oracle_cand {
    float power, sigma;
    int numharm;
    double r, z, w;
} oracle_cand;

/* Hit record: every cell over powcut in scan order
 * (rstep, stage, w, z, r), before candidate insertion. */
typedef struct oracle_hit {
    long long rstep;
    int stage, wi, zi, ri;
    float power;
    float sigma;
} oracle_hit;

/* accelsearch.c:183-223: the rstep loop for rsteps [rstep_lo, rstep_hi)
 * (rstep_hi < 0: to the end).  Candidates are inserted with
 * insert_new_accelcand (accel_utils.c:382-480).  Returns the number of
 * candidates (<= maxcands written) or <0 on error.  Hits are written to
 * hits (<= maxhits) and their total count to *nhits.
 * If sums != NULL the final (all-stage) summed fundamental plane of rstep
 * 'sum_rstep' is written there ([w][z][r], float or double per mode). */
long long oracle_search(oracle_ctx *ctx, const float *fft, long long rstep_lo, long long rstep_hi,
                        oracle_cand *cands, long long maxcands,
                        oracle_hit *hits, long long maxhits, long long *nhits,
                        long long sum_rstep, void *sums, int *stage_sums_numrs);

long long oracle_num_rsteps(const oracle_obs *obs);

/* insert_new_accelcand (accel_utils.c:382-480) applied to n hits in order. */
long long oracle_insert_cands(const double *rr, const float *sigma, const float *power, const int *numharm,
                              const double *z, const double *w, long long n, oracle_cand *out, long long maxout);

#ifdef __cplusplus
}
#endif
#endif
