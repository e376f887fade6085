/*
 * accel.c -- the PRESTO C jerk search, restated (TEST INFRASTRUCTURE ONLY).
 *
 *   oracle_create_accelobs      accel_utils.c:1636-2037 (geometry, thresholds)
 *   oracle_harm_info_get        accel_utils.c:137-149 (calc_fftlen), :199-234
 *   oracle_kernel               accel_utils.c:151-168 (init_kernel),
 *                               corr_prep.c:92-114 (place_complex_kernel)
 *   oracle_subharm_fderivs_vol  accel_utils.c:1094-1293, :1058-1092
 *   add_ffdotpows               accel_utils.c:1356-1377
 *   search_ffdotpows            accel_utils.c:1507-1555
 *   insert_new_accelcand        accel_utils.c:382-480
 *   oracle_search               accelsearch.c:183-223
 * FP32 mode keeps the reference's float/double split operation for
 * operation; FP64 mode is the same algorithm in double throughout.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#define ACCEL_NUMBETWEEN 2
#define ACCEL_DR 0.5
#define ACCEL_RDR 2
#define ACCEL_DZ 2
#define ACCEL_RDZ 0.5
#define ACCEL_DW 20
#define ACCEL_RDW 0.05
#define ACCEL_CLOSEST_R 15.0
#define DBLCORRECT 1e-14
#define NEAREST_INT(x) (int)((x) < 0 ? (x)-0.5 : (x) + 0.5)

static inline int twon_to_index(int n)
{
    int x = 0;
    while (n > 1) {
        n >>= 1;
        x++;
    }
    return x;
}
static inline double calc_required_r(double hf, double rfull) { return rint(ACCEL_RDR * rfull * hf) * ACCEL_DR; }
static inline int calc_required_z(double hf, double zfull) { return NEAREST_INT(ACCEL_RDZ * zfull * hf) * ACCEL_DZ; }
static inline int calc_required_w(double hf, double wfull) { return NEAREST_INT(ACCEL_RDW * wfull * hf) * ACCEL_DW; }
static inline int index_from_r(double r, double lor) { return (int)((r - lor) * ACCEL_RDR + DBLCORRECT); }
static inline int index_from_z(double z, double loz) { return (int)((z - loz) * ACCEL_RDZ + DBLCORRECT); }
static inline int index_from_w(double w, double low) { return (int)((w - low) * ACCEL_RDW + DBLCORRECT); }

int oracle_create_accelobs(const oracle_params *p, oracle_obs *obs)
{
    int ii, zmax = p->zmax, wmax = p->wmax;
    memset(obs, 0, sizeof(*obs));
    if (zmax % ACCEL_DZ) zmax = (zmax / ACCEL_DZ + 1) * ACCEL_DZ;
    obs->N = (long long)p->N;
    obs->numbins = p->numbins;
    obs->nph = p->nph;
    obs->lobin = p->lobin;
    if (p->numharm != 1 && p->numharm != 2 && p->numharm != 4 && p->numharm != 8 &&
        p->numharm != 16 && p->numharm != 32)
        return -1;
    obs->numharmstages = twon_to_index(p->numharm) + 1;
    obs->numz = (zmax / ACCEL_DZ) * 2 + 1;
    if (p->wmaxP) {
        if (wmax % ACCEL_DW) wmax = (wmax / ACCEL_DW + 1) * ACCEL_DW;
        obs->whi = wmax;
        obs->wlo = -wmax;
        obs->numw = (wmax / ACCEL_DW) * 2 + 1;
        if (wmax == 0) obs->numw = 0;
    } else {
        obs->whi = 0.0;
        obs->wlo = 0.0;
        obs->numw = 0;
        wmax = 0;
    }
    obs->zmax_cmd = zmax;
    obs->wmax_cmd = wmax;
    obs->dt = p->dt;
    obs->T = p->dt * p->N;
    if (p->floP) {
        obs->rlo = floor(p->flo * obs->T);
        if (obs->rlo < obs->lobin) obs->rlo = obs->lobin;
        if (obs->rlo > obs->numbins - 1) return -2;
    } else {
        obs->rlo = p->rloP ? p->rlo : 1.0;
        if (obs->rlo < obs->lobin) obs->rlo = obs->lobin;
        if (obs->rlo > obs->numbins - 1) return -2;
    }
    obs->highestbin = obs->numbins - 1;
    if (p->fhiP) {
        obs->highestbin = (long long)ceil(p->fhi * obs->T);
        if (obs->highestbin > obs->numbins - 1) obs->highestbin = obs->numbins - 1;
        obs->rhi = obs->highestbin;
        if (obs->highestbin < obs->rlo) return -3;
    } else if (p->rhiP) {
        obs->highestbin = (long long)p->rhi;
        if (obs->highestbin > obs->numbins - 1) obs->highestbin = obs->numbins - 1;
        obs->rhi = obs->highestbin;
        if (obs->highestbin < obs->rlo) return -3;
    }
    obs->zhi = zmax;
    obs->zlo = -zmax;
    obs->sigma = p->sigma;
    for (ii = 0; ii < obs->numharmstages; ii++) {
        double dz = ACCEL_DZ, dw = ACCEL_DW;
        if (obs->numz == 1 && obs->numw == 0)
            obs->numindep[ii] = (long long)((obs->rhi - obs->rlo) / (1 << ii));
        else if (obs->numz > 1 && obs->numw == 0)
            obs->numindep[ii] = (long long)((obs->rhi - obs->rlo) * (obs->numz + 1) * (dz / 6.95) / (1 << ii));
        else
            obs->numindep[ii] = (long long)((obs->rhi - obs->rlo) * (obs->numz + 1) * (dz / 6.95) *
                                            (obs->numw + 1) * (dw / 44.2) / (1 << ii));
        obs->powcut[ii] = (float)oracle_power_for_sigma(obs->sigma, 1 << ii, (double)obs->numindep[ii]);
    }
    obs->maxkernlen = 2 * ACCEL_NUMBETWEEN * oracle_w_resp_halfwidth(obs->zhi, obs->whi);
    obs->fftlen = oracle_fftlen_from_kernwidth(obs->maxkernlen);
    if (obs->fftlen < 2048) obs->fftlen = 2048;
    obs->corr_uselen = obs->fftlen - obs->maxkernlen;
    if (obs->corr_uselen % ACCEL_RDR) obs->corr_uselen = obs->corr_uselen / ACCEL_RDR * ACCEL_RDR;
    return 0;
}

long long oracle_num_rsteps(const oracle_obs *obs)
{
    long long n = 0;
    double startr = obs->rlo;
    int rstep = (int)(obs->corr_uselen * ACCEL_DR);
    while (startr + rstep < obs->highestbin) {
        n++;
        startr += rstep;
    }
    return n;
}

static int calc_fftlen(int numharm, int harmnum, int max_zfull, int max_wfull, const oracle_obs *obs)
{
    double hf = (double)harmnum / (double)numharm;
    int bins_needed = (int)ceil(obs->corr_uselen * hf) + 2;
    int end_effects = 2 * ACCEL_NUMBETWEEN *
                      oracle_w_resp_halfwidth(calc_required_z(hf, max_zfull), calc_required_w(hf, max_wfull));
    return oracle_next_good_fftlen(bins_needed + end_effects);
}

int oracle_harm_info_get(const oracle_obs *obs, int numharm, int harmnum, oracle_harm_info *h)
{
    double hf = (double)harmnum / (double)numharm;
    int zmax = (int)obs->zhi, wmax = (int)obs->whi;
    h->numharm = numharm;
    h->harmnum = harmnum;
    h->zmax = calc_required_z(hf, zmax);
    h->wmax = calc_required_w(hf, wmax);
    h->fftlen = (numharm == 1 && harmnum == 1) ? obs->fftlen : calc_fftlen(numharm, harmnum, zmax, wmax, obs);
    h->numz = (h->zmax / ACCEL_DZ) * 2 + 1;
    h->numw = (h->wmax / ACCEL_DW) * 2 + 1;
    /* kern[0][0]: z = -zmax_h, w = -wmax_h */
    h->halfwidth = oracle_w_resp_halfwidth((double)-h->zmax, (double)-h->wmax);
    return 0;
}

int oracle_kernel(const oracle_obs *obs, int numharm, int harmnum, int zi, int wi, int mode, void *out)
{
    oracle_harm_info h;
    int z, w, khw, numkern, half, fftlen;
    oracle_harm_info_get(obs, numharm, harmnum, &h);
    z = -h.zmax + zi * ACCEL_DZ;
    w = -h.wmax + wi * ACCEL_DW;
    fftlen = h.fftlen;
    khw = oracle_w_resp_halfwidth((double)z, (double)w);
    numkern = 2 * ACCEL_NUMBETWEEN * khw;
    half = numkern / 2;
    if (mode == ORACLE_FP64) {
        double *tmp = (double *)malloc(sizeof(double) * 2 * numkern), *o = (double *)out;
        oracle_gen_w_response(0.0, ACCEL_NUMBETWEEN, z, w, numkern, mode, tmp);
        memset(o, 0, sizeof(double) * 2 * fftlen);
        memcpy(o, tmp + 2 * half, sizeof(double) * 2 * half);
        memcpy(o + 2 * (fftlen - half), tmp, sizeof(double) * 2 * half);
        oracle_cfft_d(o, fftlen, -1);
        free(tmp);
    } else {
        float *tmp = (float *)malloc(sizeof(float) * 2 * numkern), *o = (float *)out;
        oracle_gen_w_response(0.0, ACCEL_NUMBETWEEN, z, w, numkern, mode, tmp);
        memset(o, 0, sizeof(float) * 2 * fftlen);
        memcpy(o, tmp + 2 * half, sizeof(float) * 2 * half);
        memcpy(o + 2 * (fftlen - half), tmp, sizeof(float) * 2 * half);
        oracle_cfft_f(o, fftlen, -1);
        free(tmp);
    }
    return 0;
}

typedef struct bank {
    oracle_harm_info h;
    void *kern; /* [w][z][fftlen] complex */
    uint16_t *rinds, *zinds;
} bank;

struct oracle_ctx {
    oracle_obs obs;
    int mode, nthreads;
    bank *banks; /* index: stage 0 -> 0; stage s harm j -> (1<<(s-1)) + (j-1)/2 */
    int nbanks;
};

static int bank_index(int numharm, int harmnum)
{
    if (numharm == 1) return 0;
    return (numharm >> 1) + (harmnum - 1) / 2;
}

oracle_ctx *oracle_ctx_create(const oracle_obs *obs, int mode, int nthreads)
{
    oracle_ctx *ctx = (oracle_ctx *)calloc(1, sizeof(oracle_ctx));
    int stage, nb = 1 << (obs->numharmstages - 1);
    size_t esz = (mode == ORACLE_FP64) ? sizeof(double) : sizeof(float);
    ctx->obs = *obs;
    ctx->mode = mode;
    ctx->nthreads = nthreads > 0 ? nthreads : 1;
    ctx->nbanks = nb;
    ctx->banks = (bank *)calloc(nb, sizeof(bank));
    for (stage = 0; stage < obs->numharmstages; stage++) {
        int numharm = 1 << stage, harm;
        for (harm = 1; harm < (stage ? numharm : 2); harm += 2) {
            bank *b = &ctx->banks[bank_index(numharm, harm)];
            long long nk, k;
            oracle_harm_info_get(obs, numharm, harm, &b->h);
            nk = (long long)b->h.numz * b->h.numw;
            b->kern = malloc(esz * 2 * (size_t)b->h.fftlen * nk);
            b->rinds = (uint16_t *)calloc(obs->corr_uselen, sizeof(uint16_t));
            b->zinds = (uint16_t *)calloc(obs->corr_uselen > obs->numz ? obs->corr_uselen : obs->numz,
                                          sizeof(uint16_t));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(ctx->nthreads)
#endif
            for (k = 0; k < nk; k++) {
                int wi = (int)(k / b->h.numz), zi = (int)(k % b->h.numz);
                oracle_kernel(obs, numharm, harm, zi, wi, mode,
                              (char *)b->kern + esz * 2 * (size_t)b->h.fftlen * k);
            }
        }
    }
    return ctx;
}

void oracle_ctx_free(oracle_ctx *ctx)
{
    int i;
    if (!ctx) return;
    for (i = 0; i < ctx->nbanks; i++) {
        free(ctx->banks[i].kern);
        free(ctx->banks[i].rinds);
        free(ctx->banks[i].zinds);
    }
    free(ctx->banks);
    free(ctx);
}

static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* accel_utils.c:1058-1092 (mmap branch) */
static void get_fourier_amplitudes(const float *fft, long long lobin, int numbins, const oracle_obs *obs,
                                   float *out)
{
    long long ii, offset = 0, firstbin, newnumbins;
    memset(out, 0, sizeof(float) * 2 * numbins);
    if (lobin - obs->lobin < 0) offset = llabs(lobin - obs->lobin);
    firstbin = (lobin - obs->lobin) + offset;
    newnumbins = numbins - offset;
    if (firstbin + newnumbins > obs->numbins) newnumbins -= firstbin + newnumbins - obs->numbins;
    if (newnumbins > 0) memcpy(out + 2 * offset, fft + 2 * firstbin, sizeof(float) * 2 * newnumbins);
    (void)ii;
}

int oracle_subharm_fderivs_vol(oracle_ctx *ctx, const float *fft, int numharm, int harmnum,
                               double fullrlo, double fullrhi, void *pows, uint16_t *rinds_out,
                               uint16_t *zinds_out, long long *meta)
{
    const oracle_obs *obs = &ctx->obs;
    bank *b = &ctx->banks[bank_index(numharm, harmnum)];
    double hf = (double)harmnum / (double)numharm;
    double drlo = calc_required_r(hf, fullrlo), drhi = calc_required_r(hf, fullrhi);
    long long rlo = (long long)floor(drlo), lobin;
    int zlo = calc_required_z(hf, obs->zlo), wlo = calc_required_w(hf, obs->wlo);
    int ii, numrs, numdata, fftlen = b->h.fftlen, binoffset = b->h.halfwidth, offset;
    int numzs = b->h.numz, numws = b->h.numw;
    float *data;

    if (numharm > 1) {
        for (ii = 0; ii < obs->corr_uselen; ii++) {
            double rr = fullrlo + ii * ACCEL_DR;
            b->rinds[ii] = (uint16_t)index_from_r(calc_required_r(hf, rr), (double)rlo);
        }
        for (ii = 0; ii < obs->numz; ii++) {
            double zz = obs->zlo + ii * ACCEL_DZ;
            b->zinds[ii] = (uint16_t)index_from_z(calc_required_z(hf, zz), zlo);
        }
        if (rinds_out) memcpy(rinds_out, b->rinds, sizeof(uint16_t) * obs->corr_uselen);
        if (zinds_out) memcpy(zinds_out, b->zinds, sizeof(uint16_t) * obs->numz);
    }
    numrs = (int)((ceil(drhi) - floor(drlo)) * ACCEL_RDR + DBLCORRECT) + 1;
    if (numharm == 1 && harmnum == 1) numrs = obs->corr_uselen;
    else if (numrs % ACCEL_RDR) numrs = (numrs / ACCEL_RDR + 1) * ACCEL_RDR;
    if (meta) {
        meta[0] = rlo;
        meta[1] = zlo;
        meta[2] = wlo;
        meta[3] = numzs;
        meta[4] = numws;
    }
    lobin = rlo - binoffset;
    numdata = fftlen / ACCEL_NUMBETWEEN;
    offset = binoffset * ACCEL_NUMBETWEEN;
    data = (float *)malloc(sizeof(float) * 2 * numdata);
    get_fourier_amplitudes(fft, lobin, numdata, obs, data);

    if (ctx->mode == ORACLE_FP32) {
        float *pdata = (float *)calloc(2 * (size_t)fftlen, sizeof(float));
        const float norm2 = (float)(1.0 / ((double)fftlen * fftlen));
        long long row, nrows = (long long)numws * numzs;
        if (obs->nph > 0.0) {
            double norm = 1.0 / sqrt(obs->nph);
            for (ii = 0; ii < 2 * numdata; ii++) data[ii] = (float)(data[ii] * norm);
        } else {
            float *powers = (float *)malloc(sizeof(float) * numdata);
            double norm;
            for (ii = 0; ii < numdata; ii++)
                powers[ii] = data[2 * ii] * data[2 * ii] + data[2 * ii + 1] * data[2 * ii + 1];
            norm = 1.0 / sqrt(oracle_median(powers, numdata) / log(2.0));
            free(powers);
            for (ii = 0; ii < 2 * numdata; ii++) data[ii] = (float)(data[ii] * norm);
        }
        for (ii = 0; ii < numdata; ii++) {
            pdata[4 * ii] = data[2 * ii];
            pdata[4 * ii + 1] = data[2 * ii + 1];
        }
        oracle_cfft_f(pdata, fftlen, -1);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 8) num_threads(ctx->nthreads)
#endif
        for (row = 0; row < nrows; row++) {
            const float *k = (const float *)b->kern + 2 * (size_t)fftlen * row;
            float *tmp = (float *)malloc(sizeof(float) * 2 * fftlen);
            float *outp = (float *)pows + (size_t)numrs * row;
            int kk;
            for (kk = 0; kk < 2 * fftlen; kk += 2) {
                const float dr = pdata[kk], di = pdata[kk + 1], kr = k[kk], ki = k[kk + 1];
                tmp[kk] = dr * kr + di * ki;
                tmp[kk + 1] = di * kr - dr * ki;
            }
            oracle_cfft_f(tmp, fftlen, +1);
            for (kk = 0; kk < numrs; kk++) {
                const int ind = 2 * (kk + offset);
                outp[kk] = (tmp[ind] * tmp[ind] + tmp[ind + 1] * tmp[ind + 1]) * norm2;
            }
            free(tmp);
        }
        free(pdata);
    } else {
        double *d = (double *)malloc(sizeof(double) * 2 * numdata);
        double *pdata = (double *)calloc(2 * (size_t)fftlen, sizeof(double));
        const double norm2 = 1.0 / ((double)fftlen * fftlen);
        long long row, nrows = (long long)numws * numzs;
        for (ii = 0; ii < 2 * numdata; ii++) d[ii] = data[ii];
        if (obs->nph > 0.0) {
            double norm = 1.0 / sqrt(obs->nph);
            for (ii = 0; ii < 2 * numdata; ii++) d[ii] *= norm;
        } else {
            /* exact order statistic of the double powers */
            double *powers = (double *)malloc(sizeof(double) * numdata), norm, med;
            for (ii = 0; ii < numdata; ii++) powers[ii] = d[2 * ii] * d[2 * ii] + d[2 * ii + 1] * d[2 * ii + 1];
            qsort(powers, (size_t)numdata, sizeof(double), cmp_double);
            med = powers[(numdata - 1) / 2];
            norm = 1.0 / sqrt(med / log(2.0));
            free(powers);
            for (ii = 0; ii < 2 * numdata; ii++) d[ii] *= norm;
        }
        for (ii = 0; ii < numdata; ii++) {
            pdata[4 * ii] = d[2 * ii];
            pdata[4 * ii + 1] = d[2 * ii + 1];
        }
        oracle_cfft_d(pdata, fftlen, -1);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 8) num_threads(ctx->nthreads)
#endif
        for (row = 0; row < nrows; row++) {
            const double *k = (const double *)b->kern + 2 * (size_t)fftlen * row;
            double *tmp = (double *)malloc(sizeof(double) * 2 * fftlen);
            double *outp = (double *)pows + (size_t)numrs * row;
            int kk;
            for (kk = 0; kk < 2 * fftlen; kk += 2) {
                const double dr = pdata[kk], di = pdata[kk + 1], kr = k[kk], ki = k[kk + 1];
                tmp[kk] = dr * kr + di * ki;
                tmp[kk + 1] = di * kr - dr * ki;
            }
            oracle_cfft_d(tmp, fftlen, +1);
            for (kk = 0; kk < numrs; kk++) {
                const int ind = 2 * (kk + offset);
                outp[kk] = (tmp[ind] * tmp[ind] + tmp[ind + 1] * tmp[ind + 1]) * norm2;
            }
            free(tmp);
        }
        free(pdata);
        free(d);
    }
    free(data);
    return numrs;
}

/* ---------------- candidate list: accel_utils.c:382-480 ---------------- */
typedef struct node {
    oracle_cand c;
    struct node *next;
} node;

static void set_cand(node *n, float power, float sigma, int numharm, double r, double z, double w)
{
    n->c.power = power;
    n->c.sigma = sigma;
    n->c.numharm = numharm;
    n->c.r = r;
    n->c.z = z;
    n->c.w = w;
}

static node *new_node(float power, float sigma, int numharm, double r, double z, double w)
{
    node *n = (node *)malloc(sizeof(node));
    set_cand(n, power, sigma, numharm, r, z, w);
    n->next = NULL;
    return n;
}

static node *insert_new_accelcand(node *list, float power, float sigma, int numharm, double rr, double zz,
                                  double ww, int *added)
{
    node *tmp_list = list, *prev_list = NULL, *nl;
    double prev_diff_r = ACCEL_CLOSEST_R + 1.0, next_diff_r;
    *added = 0;
    if (!list) {
        *added = 1;
        return new_node(power, sigma, numharm, rr, zz, ww);
    }
    while (tmp_list->next && tmp_list->c.r < rr) {
        prev_list = tmp_list;
        tmp_list = tmp_list->next;
    }
    next_diff_r = fabs(rr - tmp_list->c.r);
    if (prev_list) prev_diff_r = fabs(rr - prev_list->c.r);
    if (prev_diff_r < ACCEL_CLOSEST_R) {
        if (prev_list->c.sigma < sigma) {
            set_cand(prev_list, power, sigma, numharm, rr, zz, ww);
            *added = 1;
        }
        if (next_diff_r < ACCEL_CLOSEST_R) {
            if (tmp_list->c.sigma < sigma) {
                if (*added) {
                    /* remove tmp_list (g_slist_remove_link) */
                    node *p = list, *pp = NULL;
                    while (p && p != tmp_list) {
                        pp = p;
                        p = p->next;
                    }
                    if (pp) pp->next = tmp_list->next;
                    else list = tmp_list->next;
                    free(tmp_list);
                } else {
                    set_cand(tmp_list, power, sigma, numharm, rr, zz, ww);
                    *added = 1;
                }
            }
        }
    } else if (next_diff_r < ACCEL_CLOSEST_R) {
        if (tmp_list->c.sigma < sigma) {
            set_cand(tmp_list, power, sigma, numharm, rr, zz, ww);
            *added = 1;
        }
    } else {
        nl = new_node(power, sigma, numharm, rr, zz, ww);
        *added = 1;
        if (!tmp_list->next && tmp_list->c.r < (rr - ACCEL_CLOSEST_R)) {
            tmp_list->next = nl;
            return list;
        }
        if (prev_list) {
            prev_list->next = nl;
            nl->next = tmp_list;
        } else {
            nl->next = list;
            return nl;
        }
    }
    return list;
}

/* The list logic alone: apply n hits in order (for pinning the product's
 * host-side candidate list against the literal linked-list walk). */
long long oracle_insert_cands(const double *rr, const float *sigma, const float *power, const int *numharm,
                              const double *z, const double *w, long long n, oracle_cand *out, long long maxout)
{
    node *list = NULL, *nd;
    long long i, cnt = 0;
    int added;
    for (i = 0; i < n; i++)
        list = insert_new_accelcand(list, power[i], sigma[i], numharm[i], rr[i], z[i], w[i], &added);
    for (nd = list; nd;) {
        node *nx = nd->next;
        if (cnt < maxout) out[cnt] = nd->c;
        cnt++;
        free(nd);
        nd = nx;
    }
    return cnt;
}

typedef struct search_state {
    node *cands;
    oracle_hit *hits;
    long long maxhits, nhits;
} search_state;

/* accel_utils.c:1507-1555 over a plane of numws x numzs x numrs (float or double). */
static void search_plane(search_state *st, const oracle_obs *obs, int mode, const void *pows, int numws,
                         int numzs, int numrs, long long rlo, int zlo, int wlo, int numharm, long long rstep_idx,
                         int stage)
{
    float powcut = obs->powcut[twon_to_index(numharm)];
    long long numindep = obs->numindep[twon_to_index(numharm)];
    int ii, jj, kk;
    for (ii = 0; ii < numws; ii++)
        for (jj = 0; jj < numzs; jj++)
            for (kk = 0; kk < numrs; kk++) {
                size_t idx = ((size_t)ii * numzs + jj) * numrs + kk;
                double pw = (mode == ORACLE_FP64) ? ((const double *)pows)[idx] : ((const float *)pows)[idx];
                if (pw > powcut) {
                    float pow = (float)pw, sig;
                    double rr, zz, ww;
                    int added;
                    sig = (float)oracle_candidate_sigma(pow, numharm, (double)numindep);
                    rr = (rlo + kk * (double)ACCEL_DR) / (double)numharm;
                    zz = (zlo + jj * (double)ACCEL_DZ) / (double)numharm;
                    ww = (wlo + ii * (double)ACCEL_DW) / (double)numharm;
                    if (st->nhits < st->maxhits && st->hits) {
                        oracle_hit *h = &st->hits[st->nhits];
                        h->rstep = rstep_idx;
                        h->stage = stage;
                        h->wi = ii;
                        h->zi = jj;
                        h->ri = kk;
                        h->power = pow;
                        h->sigma = sig;
                    }
                    st->nhits++;
                    st->cands = insert_new_accelcand(st->cands, pow, sig, numharm, rr, zz, ww, &added);
                }
            }
}

/* accel_utils.c:1356-1377 */
static void add_plane(int mode, void *fund, int numws, int numzs, int numrs, int fwlo, const void *sub,
                      int sub_numzs, int sub_numrs, int sub_wlo, const uint16_t *rinds, const uint16_t *zinds,
                      int numharm, int harmnum)
{
    const double hf = (double)harmnum / (double)numharm;
    int ii, jj, kk;
    for (ii = 0; ii < numws; ii++) {
        int ww = fwlo + ii * ACCEL_DW;
        int subw = calc_required_w(hf, ww);
        int wind = index_from_w(subw, sub_wlo);
        for (jj = 0; jj < numzs; jj++) {
            int zind = zinds[jj];
            for (kk = 0; kk < numrs; kk++) {
                size_t fi = ((size_t)ii * numzs + jj) * numrs + kk;
                size_t si = ((size_t)wind * sub_numzs + zind) * sub_numrs + rinds[kk];
                if (mode == ORACLE_FP64) ((double *)fund)[fi] += ((const double *)sub)[si];
                else ((float *)fund)[fi] += ((const float *)sub)[si];
            }
        }
    }
}

long long oracle_search(oracle_ctx *ctx, const float *fft, long long rstep_lo, long long rstep_hi,
                        oracle_cand *cands, long long maxcands, oracle_hit *hits, long long maxhits,
                        long long *nhits, long long sum_rstep, void *sums, int *sums_numrs)
{
    const oracle_obs *obs = &ctx->obs;
    int rstep = (int)(obs->corr_uselen * ACCEL_DR);
    size_t esz = (ctx->mode == ORACLE_FP64) ? sizeof(double) : sizeof(float);
    double startr = obs->rlo, lastr, nextr;
    long long idx = 0, ncand = 0;
    search_state st;
    bank *fb = &ctx->banks[0];
    size_t fsize = (size_t)fb->h.numw * fb->h.numz * obs->corr_uselen;
    void *fund = malloc(esz * fsize);
    void *sub = NULL;
    size_t subcap = 0;
    node *n;

    st.cands = NULL;
    st.hits = hits;
    st.maxhits = maxhits;
    st.nhits = 0;
    while (startr + rstep < obs->highestbin) {
        nextr = startr + rstep;
        lastr = nextr - ACCEL_DR;
        if (idx >= rstep_lo && (rstep_hi < 0 || idx < rstep_hi)) {
            long long meta[5];
            int numrs = oracle_subharm_fderivs_vol(ctx, fft, 1, 1, startr, lastr, fund, NULL, NULL, meta);
            int stage;
            long long frlo = meta[0];
            int fzlo = (int)meta[1], fwlo = (int)meta[2];
            search_plane(&st, obs, ctx->mode, fund, fb->h.numw, fb->h.numz, numrs, frlo, fzlo, fwlo, 1, idx, 0);
            for (stage = 1; stage < obs->numharmstages; stage++) {
                int harmtosum = 1 << stage, harm;
                for (harm = 1; harm < harmtosum; harm += 2) {
                    bank *sb = &ctx->banks[bank_index(harmtosum, harm)];
                    size_t need = (size_t)sb->h.numw * sb->h.numz * (size_t)(obs->corr_uselen + 8) * esz;
                    long long smeta[5];
                    int snumrs;
                    if (need > subcap) {
                        free(sub);
                        sub = malloc(need);
                        subcap = need;
                    }
                    snumrs = oracle_subharm_fderivs_vol(ctx, fft, harmtosum, harm, startr, lastr, sub, NULL, NULL,
                                                        smeta);
                    add_plane(ctx->mode, fund, fb->h.numw, fb->h.numz, numrs, fwlo, sub, sb->h.numz, snumrs,
                              (int)smeta[2], sb->rinds, sb->zinds, harmtosum, harm);
                }
                search_plane(&st, obs, ctx->mode, fund, fb->h.numw, fb->h.numz, numrs, frlo, fzlo, fwlo, harmtosum,
                             idx, stage);
            }
            if (sums && idx == sum_rstep) {
                memcpy(sums, fund, esz * fsize);
                if (sums_numrs) *sums_numrs = numrs;
            }
        }
        startr = nextr;
        idx++;
    }
    free(fund);
    free(sub);
    if (nhits) *nhits = st.nhits;
    for (n = st.cands; n;) {
        node *nx = n->next;
        if (ncand < maxcands && cands) cands[ncand] = n->c;
        ncand++;
        free(n);
        n = nx;
    }
    return ncand;
}
