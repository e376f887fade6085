/*
 * accel.h -- the PRESTO/PrestoZL jerk-search API, served by the MI355X
 * engine in libaccelsearch_lib.so.
 *
 * The structs are layout-identical to the reference so that existing callers
 * (accelsearch_cu.c, the ctypes pipeline bin/accelsearch_pipeline_cu.py
 * :51-233) link or bind unchanged:
 *   accelobs, accelcand, kernel, subharminfo, ffdotpows, ffdotpows_cu
 *                                  include/accel.h:37-150
 *   infodata                       include/makeinf.h:23-56
 *   Cmdline                        include/accelsearch_cmd.h:11-87
 *   fcomplex, rderivs              include/ransomfft.h:14-19, include/presto.h:161-169
 * Candidate lists are glib GSLists as in the reference.
 *
 * Behavioural notes:
 *   - kernel.data is NULL: the FFT'd response kernels live in GPU memory
 *     (built there once per search shape and cached across calls).
 *   - accelsearch_GPU searches the fundamental stage also when numharm == 1
 *     (the reference GPU path returns no candidates then,
 *     accelsearch_lib.c:648); results follow PRESTO C accelsearch.
 *   - accelsearch_CPU2 runs the reference's sequence
 *     (accelsearch_lib.c:903-980): sort, eliminate_harmonics, the
 *     optimisation of every candidate on the GPU (max_rzw_arr_harmonics /
 *     max_rz_arr_harmonics, or per harmonic with -noharmpolish), calc_props,
 *     the ACCEL text tables and the binary .cand file, then frees.
 *   - CPU1, GPU and CPU2 of different files may run at once on different
 *     threads, as the reference pipeline does (no shared mutable state
 *     between the stages of different files).
 */
#ifndef PZL_ACCEL_API_H
#define PZL_ACCEL_API_H

#include <stdio.h>
#include <glib.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ACCEL_USELEN 7470
#define ACCEL_NUMBETWEEN 2
#define ACCEL_DR 0.5
#define ACCEL_RDR 2
#define ACCEL_DZ 2
#define ACCEL_RDZ 0.5
#define ACCEL_DW 20
#define ACCEL_RDW 0.05
#define ACCEL_CLOSEST_R 15.0
#define ACCEL_PADDING 2000

#ifndef _FCOMPLEX_DECLARED_
typedef struct FCOMPLEX {
    float r, i;
} fcomplex;
#define _FCOMPLEX_DECLARED_
#endif

#ifndef RDERIVS_DECLARED
#define RDERIVS_DECLARED
typedef struct RDERIVS {
    double pow, phs, dpow, dphs, d2pow, d2phs, locpow;
} rderivs;
#endif

#ifndef INFODATA_DEFINED
#define INFODATA_DEFINED
#define MAXNUMONOFF 40
/* presto.h:171-189 (the records of the binary .cand file) */
typedef struct FOURIERPROPS {
    double r;
    float rerr;
    double z;
    float zerr;
    double w;
    float werr;
    float pow;
    float powerr;
    float sig;
    float rawpow;
    float phs;
    float phserr;
    float cen;
    float cenerr;
    float pur;
    float purerr;
    float locpow;
} fourierprops;

typedef struct INFODATA {
    double ra_s, dec_s, N, dt, fov, mjd_f, dm, freq, freqband, chan_wid, wavelen, waveband, energy, energyband;
    double onoff[MAXNUMONOFF * 2];
    int num_chan, mjd_i, ra_h, ra_m, dec_d, dec_m, bary, numonoff;
    char notes[500];
    char name[200];
    char object[100];
    char instrument[100];
    char observer[100];
    char analyzer[100];
    char telescope[40];
    char band[40];
    char filt[7];
} infodata;
#endif

#ifndef __accelsearch_cmd__
#define __accelsearch_cmd__
typedef struct s_Cmdline {
    char batchsizeP;
    int batchsize;
    int batchsizeC;
    char ncpusP;
    int ncpus;
    int ncpusC;
    char lobinP;
    int lobin;
    int lobinC;
    char numharmP;
    int numharm;
    int numharmC;
    char zmaxP;
    int zmax;
    int zmaxC;
    char wmaxP;
    int wmax;
    int wmaxC;
    char sigmaP;
    float sigma;
    int sigmaC;
    char rloP;
    double rlo;
    int rloC;
    char rhiP;
    double rhi;
    int rhiC;
    char floP;
    double flo;
    int floC;
    char fhiP;
    double fhi;
    int fhiC;
    char inmemP;
    char photonP;
    char medianP;
    char locpowP;
    char zaplistP;
    char *zaplist;
    int zaplistC;
    char baryvP;
    double baryv;
    int baryvC;
    char otheroptP;
    char noharmpolishP;
    char noharmremoveP;
    int argc;
    char **argv;
    char *full_cmd_line;
    char cudaP;
    char cuda;
} Cmdline;
#endif

typedef struct accelobs {
    long long N;
    long long numbins;
    long long lobin;
    long long highestbin;
    int maxkernlen;
    int corr_uselen;
    int fftlen;
    int numharmstages;
    int numz;
    int numw;
    int numbetween;
    int numzap;
    int dat_input;
    int mmap_file;
    int inmem;
    int norm_type;
    double dt;
    double T;
    double rlo;
    double rhi;
    double dr;
    double zlo;
    double zhi;
    double dz;
    double wlo;
    double whi;
    double dw;
    double baryv;
    float nph;
    float sigma;
    float *powcut;
    float *ffdotplane;
    double *lobins;
    double *hibins;
    long long *numindep;
    FILE *fftfile;
    FILE *workfile;
    fcomplex *fft;
    char *rootfilenm;
    char *candnm;
    char *accelnm;
    char *workfilenm;
    int use_harmonic_polishing;
} accelobs;

typedef struct accelcand {
    float power;
    float sigma;
    int numharm;
    double r;
    double z;
    double w;
    double *pows;
    double *hirs;
    double *hizs;
    double *hiws;
    rderivs *derivs;
} accelcand;

typedef struct kernel {
    int z;
    int w;
    int fftlen;
    int numgoodbins;
    int numbetween;
    int kern_half_width;
    fcomplex *data;
} kernel;

typedef struct subharminfo {
    int numharm;
    int harmnum;
    int zmax;
    int wmax;
    int numkern_zdim;
    int numkern_wdim;
    int numkern;
    kernel **kern;
    unsigned short *rinds;
    unsigned short *zinds;
} subharminfo;

typedef struct ffdotpows {
    long long rlo;
    int zlo;
    int wlo;
    int numrs;
    int numzs;
    int numws;
    float ***powers;
    unsigned short *rinds;
    unsigned short *zinds;
} ffdotpows;

#define matrix_3d_index(x, y, z, y_len, z_len) ((x * y_len + y) * z_len + z)

typedef struct ffdotpows_cu {
    long long rlo;
    int zlo;
    int wlo;
    int numrs;
    int numzs;
    int numws;
    float *powers;
    unsigned short *rinds;
    unsigned short *zinds;
} ffdotpows_cu;

/* accel_utils.c:236-282, :304-325 */
subharminfo **create_subharminfos(accelobs *obs);
void free_subharminfos(accelobs *obs, subharminfo **shis);
/* accel_utils.c:1636-2037, :2039-2065 */
void create_accelobs(accelobs *obs, infodata *idata, Cmdline *cmd, int usemmap);
void free_accelobs(accelobs *obs);
/* accel_utils.c:361-380, :482-571 */
GSList *sort_accelcands(GSList *list);
GSList *eliminate_harmonics(GSList *cands, int *numcands);
void free_accelcand(gpointer data, gpointer user_data);
void print_accelcand(gpointer data, gpointer user_data);
/* accel_utils.c:1058-1092 */
fcomplex *get_fourier_amplitudes(long long lobin, int numbins, accelobs *obs);
/* accel_utils.c:1094-1293: the plane of one (rstep, harmonic), computed on the GPU */
ffdotpows *subharm_fderivs_vol(int numharm, int harmnum, double fullrlo, double fullrhi, subharminfo *shi,
                               accelobs *obs);
void free_ffdotpows(ffdotpows *ffd);
/* accel_utils.c:1356-1377, :1379-1406 */
void add_ffdotpows(ffdotpows *fundamental, ffdotpows *subharmonic, int numharm, int harmnum);
void add_ffdotpows_ptrs(ffdotpows *fundamental, ffdotpows *subharmonic, int numharm, int harmnum);
/* accel_utils.c:1507-1555 */
GSList *search_ffdotpows(ffdotpows *ffdot, int numharm, accelobs *obs, GSList *cands);
/* accelsearch_lib.c:360-471, :473-901, :903-980 */
/* accelsearch_CPU2's candidate stage (accel_utils.c:573-1047) */
void optimize_accelcand(accelcand *cand, accelobs *obs);
void output_fundamentals(fourierprops *props, GSList *list, accelobs *obs, infodata *idata);
void output_harmonics(GSList *list, accelobs *obs, infodata *idata);
void calc_props(rderivs data, double r, double z, double w, fourierprops *result); /* characteristics.c:193 */
/* "value(error)" strings with a one- or two-digit error (output.c:6, :114);
 * len 0: left-justified, else right-justified in len characters (output
 * must hold at least max(len, 64) characters) */
int nice_output_1(char *output, double val, double err, int len);
int nice_output_2(char *output, double val, double err, int len);

void accelsearch_CPU1(int argc, char *argv[], subharminfo ***subharminfs_ptr, accelobs *obs_ptr,
                      infodata *idata_ptr, Cmdline **cmd_ptr);
int accelsearch_GPU(accelobs obs, subharminfo **subharminfs, GSList **cands_ptr, Cmdline *cmd);
void accelsearch_CPU2(GSList **cands, accelobs *obs, infodata *idata, Cmdline *cmd);
/* accelsearch_cmd.c:1259 */
Cmdline *parseCmdline(int argc, char **argv);
/* characteristics.c:647-668, :670-705 */
double candidate_sigma(double power, int numsum, double numtrials);
double power_for_sigma(double sigma, int numsum, double numtrials);

#ifdef __cplusplus
}
#endif
#endif
