/*
 * optimize.c -- oracle restatement of the candidate optimisation of
 * accelsearch_CPU2 (TEST INFRASTRUCTURE ONLY): Fourier interpolation at a
 * point, local power, derivatives, the oversampled F/Fdot/Fdotdot volume by
 * correlation, the harmonic-summed volume maximum and the Fourier properties.
 *
 * Reference arithmetic: float data, float responses and FFTs (this oracle's
 * FFT in place of FFTW), double where the reference computes in double.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NUMFINTBINS 16
#define NUMLOCPOWAVG 20
#define DELTAAVGBINS 5
#define ACCEL_DR 0.5
#define ACCEL_RDR 2
#define ACCEL_DZ 2
#define ACCEL_DW 20

/* responses.c:27-47 and 49-84 with the accuracy switch (LOWACC 0, HIGHACC 1) */
int oracle_z_resp_halfwidth_acc(double z, int highacc)
{
    int m = (int)(0.5 * 1.1 * fabs(z));
    return m + (highacc ? NUMFINTBINS * 3 : NUMFINTBINS);
}

int oracle_w_resp_halfwidth_acc(double z, double w, int highacc)
{
    if (fabs(w) < 1.0e-7) return oracle_z_resp_halfwidth_acc(z, highacc);
    double r0 = 0.5 * (w / 6.0 - z), r1 = 0.5 * (w / 6.0 + z);
    double maxdev = fabs(r0) > fabs(r1) ? fabs(r0) : fabs(r1);
    double u_ext = 0.5 - z / w;
    if (u_ext > 0.0 && u_ext < 1.0) {
        double z0 = z - w / 2.0;
        double r_ext = 0.5 * w * u_ext * u_ext + z0 * u_ext + r0;
        maxdev = fabs(r_ext) > maxdev ? fabs(r_ext) : maxdev;
    }
    return (int)(1.1 * maxdev) + (highacc ? NUMFINTBINS * 3 : NUMFINTBINS);
}

/* The manual correlation sum of rz_interp / rzw_interp (rzinterp.c:196-229,
 * rzwinterp.c:185-228) over float data and response. */
static void interp_sum(const float *data, int numdata, int intfreq, int kern_half_width,
                       const float *response, float *ans)
{
    int numkern = 2 * kern_half_width, lodata, hidata, loresp, hiresp, nsum, ii;
    lodata = intfreq - kern_half_width;
    if (lodata < 0) {
        loresp = abs(lodata);
        lodata = 0;
    } else {
        loresp = 0;
    }
    hidata = intfreq + kern_half_width - 1;
    hiresp = (hidata > numdata - 1) ? numkern - hidata + numdata - 1 : numkern;
    nsum = hiresp - loresp;
    const float *dp = data + 2 * lodata, *rp = response + 2 * loresp;
    /* the answer is the float fcomplex itself; each term is formed in double
     * (tmpd, tmpr are double; the product of the two imaginary parts is float) */
    float fr = 0.0f, fi = 0.0f;
    for (ii = 0; ii < nsum; ii++) {
        double tmpd = dp[0], tmpr = rp[0];
        const float im_im = dp[1] * rp[1];
        fr = (float)((double)fr + (tmpd * tmpr + (double)im_im));
        fi = (float)((double)fi + ((double)dp[1] * tmpr - (double)rp[1] * tmpd));
        dp += 2;
        rp += 2;
    }
    ans[0] = fr;
    ans[1] = fi;
}

/* rzinterp.c:144-230 */
void oracle_rz_interp(const float *data, int numdata, double r, double z, int kern_half_width, float *ans)
{
    double fracfreq, dintfreq;
    int intfreq;
    if (r > numdata - 1.0 || r < 0.0) {
        ans[0] = ans[1] = 0.0f;
        return;
    }
    fracfreq = modf(r, &dintfreq);
    intfreq = (int)dintfreq;
    if (fabs(z) < 1E-4) {
        if (fracfreq < 1E-5) {
            ans[0] = data[2 * intfreq];
            ans[1] = data[2 * intfreq + 1];
            return;
        }
        if ((1.0 - fracfreq) < 1E-5) {
            ans[0] = data[2 * (intfreq + 1)];
            ans[1] = data[2 * (intfreq + 1) + 1];
            return;
        }
    }
    float *resp = (float *)malloc(sizeof(float) * 4 * kern_half_width);
    oracle_gen_z_response(fracfreq, 1, z, 2 * kern_half_width, ORACLE_FP32, resp);
    interp_sum(data, numdata, intfreq, kern_half_width, resp, ans);
    free(resp);
}

/* rzwinterp.c:145-229 */
void oracle_rzw_interp(const float *data, int numdata, double r, double z, double w, int kern_half_width,
                       float *ans)
{
    double fracfreq, dintfreq;
    int intfreq;
    if (r > numdata - 1.0 || r < 0.0) {
        ans[0] = ans[1] = 0.0f;
        return;
    }
    fracfreq = modf(r, &dintfreq);
    intfreq = (int)dintfreq;
    if (fabs(w) < 1E-4) {
        oracle_rz_interp(data, numdata, r, z, kern_half_width, ans);
        return;
    }
    float *resp = (float *)malloc(sizeof(float) * 4 * kern_half_width);
    oracle_gen_w_response(fracfreq, 1, z, w, 2 * kern_half_width, ORACLE_FP32, resp);
    interp_sum(data, numdata, intfreq, kern_half_width, resp, ans);
    free(resp);
}

/* POWER (presto.h:72-73) with double argument variables, and with float ones
 * (maximize_rzw.c:92 declares them float) */
static double power_d(const float *a) { return (double)a[0] * a[0] + (double)a[1] * a[1]; }
static float power_f(const float *a)
{
    const float x = a[0] * a[0], y = a[1] * a[1];
    return x + y;
}

/* characteristics.c:75-136 */
double oracle_get_localpower3d(const float *data, int numdata, double r, double z, double w)
{
    double sum = 0.0, lo1, lo2, hi1, hi2, freq;
    int binsperside = NUMLOCPOWAVG / 2;
    int khw = oracle_w_resp_halfwidth_acc(z, w, 0);
    float ans[2];
    lo1 = r - DELTAAVGBINS - binsperside;
    hi1 = lo1 + binsperside;
    lo2 = r + DELTAAVGBINS;
    hi2 = lo2 + binsperside;
    if (lo1 < 0.0) lo1 = 0.0;
    if (lo2 < 0.0) lo2 = 0.0;
    if (hi1 < 0.0) hi1 = 0.0;
    if (hi2 < 0.0) hi2 = 0.0;
    if (lo1 > numdata) lo1 = (double)numdata;
    if (lo2 > numdata) lo2 = (double)numdata;
    if (hi1 > numdata) hi1 = (double)numdata;
    if (hi2 > numdata) hi2 = (double)numdata;
    for (freq = lo1; freq < hi1; freq += 1.0) {
        oracle_rzw_interp(data, numdata, freq, z, w, khw, ans);
        sum += power_d(ans);
    }
    for (freq = lo2; freq < hi2; freq += 1.0) {
        oracle_rzw_interp(data, numdata, freq, z, w, khw, ans);
        sum += power_d(ans);
    }
    return sum / (double)NUMLOCPOWAVG;
}

/* RADIAN_PHASE (presto.h:89-91): atan2 in (0, 2 pi] */
static double radian_phase(double re, double im)
{
    double t = atan2(im, re);
    return t > 0.0 ? t : t + 2.0 * M_PI;
}

/* characteristics.c:138-191 */
void oracle_get_derivs3d(const float *data, int numdata, double r, double z, double w, double localpower,
                         OracleRDerivs *result)
{
    double h = 0.003, twoh = 0.006, twohsqrd = 0.000036, f, pwr[5], phs[5];
    int ii, khw = oracle_w_resp_halfwidth_acc(z, w, 1);
    float ans[2];
    for (ii = 0, f = r - twoh; ii < 5; ii++, f += h) {
        oracle_rzw_interp(data, numdata, f, z, w, khw, ans);
        pwr[ii] = power_d(ans);
        phs[ii] = radian_phase(ans[0], ans[1]);
    }
    for (ii = 0; ii < 4; ii++)
        if (fabs(phs[ii + 1] - phs[ii]) > M_PI) phs[ii + 1] -= 2.0 * M_PI;
    result->pow = pwr[2];
    result->phs = phs[2];
    result->dpow = (pwr[3] - pwr[1]) / twoh;
    result->dphs = (phs[3] - phs[1]) / twoh;
    result->d2pow = (pwr[4] - 2.0 * pwr[2] + pwr[0]) / twohsqrd;
    result->d2phs = (phs[4] - 2.0 * phs[2] + phs[0]) / twohsqrd;
    result->locpow = localpower;
}

/* place_complex_kernel (corr_prep.c:92-114) */
static void place_kernel(const float *kern, int numkern, float *out, int fftlen)
{
    int halfwidth = numkern / 2;
    memset(out, 0, sizeof(float) * 2 * fftlen);
    memcpy(out, kern + 2 * halfwidth, sizeof(float) * 2 * halfwidth);
    memcpy(out + 2 * (fftlen - halfwidth), kern, sizeof(float) * 2 * halfwidth);
}

/* corr_rzw_vol (rzwinterp.c:3-140) with corr_complex (corr_routines.c:5-269,
 * RAW data, FFT'd kernel, CORR): out[(i * numz + j) * numgoodbins + k],
 * interleaved re,im; returns numgoodbins. */
int oracle_corr_rzw_vol(const float *data, int numdata, int numbetween, int startbin, double zlo, double zhi,
                        int numz, double wlo, double whi, int numw, int fftlen, int *nextbin, float *out)
{
    double maxz, dz, maxw, dw, t;
    int ii, jj, k, khw, numkern, numgoodbins, numbins, beginbin;
    if (zlo > zhi) {
        t = zlo;
        zlo = zhi;
        zhi = t;
    }
    if (wlo > whi) {
        t = wlo;
        wlo = whi;
        whi = t;
    }
    dz = numz == 1 ? 0.0 : (zhi - zlo) / (numz - 1);
    maxz = (fabs(zlo) < fabs(zhi)) ? zhi : zlo;
    dw = numw == 1 ? 0.0 : (whi - wlo) / (numw - 1);
    maxw = (fabs(wlo) < fabs(whi)) ? whi : wlo;
    khw = oracle_w_resp_halfwidth_acc(maxz, maxw, 0);
    numkern = 2 * numbetween * khw;
    numgoodbins = fftlen - 2 * khw * numbetween;
    while (numgoodbins % numbetween) numgoodbins--;
    *nextbin = startbin + numgoodbins / numbetween;

    /* data: spread and FFT (corr_complex, datainf RAW) */
    float *darr = (float *)calloc((size_t)2 * fftlen, sizeof(float));
    numbins = fftlen / numbetween;
    beginbin = startbin - khw;
    for (k = 0; k < numbins; k++) {
        const int b = beginbin + k;
        if (b >= 0 && b < numdata) {
            darr[2 * k * numbetween] = data[2 * b];
            darr[2 * k * numbetween + 1] = data[2 * b + 1];
        }
    }
    oracle_cfft_f(darr, fftlen, -1);

    float *resp = (float *)malloc(sizeof(float) * 2 * numkern);
    float *kern = (float *)malloc(sizeof(float) * 2 * fftlen);
    float *prod = (float *)malloc(sizeof(float) * 2 * fftlen);
    const float normal = (float)(1.0 / fftlen);
    for (ii = 0; ii < numw; ii++) {
        for (jj = 0; jj < numz; jj++) {
            oracle_gen_w_response(0.0, numbetween, zlo + jj * dz, wlo + ii * dw, numkern, ORACLE_FP32, resp);
            place_kernel(resp, numkern, kern, fftlen);
            oracle_cfft_f(kern, fftlen, -1);
            /* complex_corr_conv (correlations.c:56-64): data * conj(kernel) / n */
            for (k = 0; k < fftlen; k++) {
                const float dr = darr[2 * k], di = darr[2 * k + 1], kr = kern[2 * k], ki = kern[2 * k + 1];
                prod[2 * k] = (float)(dr * kr + di * ki) * normal;
                prod[2 * k + 1] = (float)(di * kr - ki * dr) * normal;
            }
            oracle_cfft_f(prod, fftlen, +1);
            /* chop_complex_ends (corr_prep.c:140-165) */
            float *o = out + (size_t)2 * ((size_t)ii * numz + jj) * numgoodbins;
            memcpy(o, prod + 2 * khw * numbetween, sizeof(float) * 2 * numgoodbins);
        }
    }
    free(darr);
    free(resp);
    free(kern);
    free(prod);
    return numgoodbins;
}

/* max_rzw_arr_harmonics (maximize_rzw.c:85-191) */
void oracle_max_rzw_arr_harmonics(const float *data, int numdata, int num_harmonics, double rin, double zin,
                                  double win, double *rout, double *zout, double *wout, OracleRDerivs *derivs,
                                  double *powers)
{
    float maxpow = 0.0f, pw, locpow;
    long lobin, hhlobin = 0;
    int hind, ii, jj, kk, extra = 10, nextbin, fftlen;
    const int interpfac = 4, numwid = 2, numbetween = interpfac * ACCEL_RDR;
    const int numr = 3 * numbetween, numz = 2 * numwid * interpfac + 1, numw = numz;
    int rind = 0, zind = 0, wind = 0;
    const double dr = 1.0 / (double)numbetween;
    float *powsum = (float *)calloc((size_t)numw * numz * numr, sizeof(float));

    for (hind = 0; hind < num_harmonics; hind++) {
        const int n = num_harmonics - hind;
        const double rh = rin * n, zh = zin * n, wh = win * n;
        const int khw = oracle_w_resp_halfwidth_acc(fabs(zh) + 2, fabs(wh) + 20, 0);
        double rh_int, hfrac = n / (double)num_harmonics;
        locpow = (float)oracle_get_localpower3d(data, numdata, rh, zh, wh);
        modf(rh, &rh_int);
        lobin = (long)rh_int - 1;
        if (hind == 0) hhlobin = lobin;
        fftlen = oracle_next2_to_n(numbetween * (2 * khw + extra));
        const double zlo = zh - numwid * hfrac * ACCEL_DZ, zhi = zh + numwid * hfrac * ACCEL_DZ;
        const double wlo = wh - numwid * hfrac * ACCEL_DW, whi = wh + numwid * hfrac * ACCEL_DW;
        float *vol = (float *)malloc(sizeof(float) * 2 * (size_t)numw * numz * fftlen);
        const int ngood = oracle_corr_rzw_vol(data, numdata, numbetween, (int)lobin, zlo, zhi, numz, wlo, whi,
                                              numw, fftlen, &nextbin, vol);
        for (ii = 0; ii < numw; ii++)
            for (jj = 0; jj < numz; jj++)
                for (kk = 0; kk < numr; kk++) {
                    const long ri = (long)round(((hhlobin + kk * dr) * hfrac - lobin) * numbetween);
                    /* vol is one contiguous [w][z][r] block: a negative index
                     * reads the end of the preceding row; before the block the
                     * reference reads outside its allocation, taken here as 0 */
                    const long flat = ((long)ii * numz + jj) * ngood + ri;
                    const float pw0 = flat >= 0 ? power_f(vol + 2 * flat) : 0.0f;
                    powsum[((size_t)ii * numz + jj) * numr + kk] += pw0 / locpow;
                }
        free(vol);
    }
    for (ii = 0; ii < numw; ii++)
        for (jj = 0; jj < numz; jj++)
            for (kk = 0; kk < numr; kk++) {
                pw = powsum[((size_t)ii * numz + jj) * numr + kk];
                if (pw > maxpow) {
                    maxpow = pw;
                    wind = ii;
                    zind = jj;
                    rind = kk;
                }
            }
    {
        const double nh = num_harmonics;
        *rout = (hhlobin + rind * dr) / nh;
        *zout = ((zin * nh) - numwid * ACCEL_DZ + zind * ACCEL_DZ / (double)interpfac) / nh;
        *wout = ((win * nh) - numwid * ACCEL_DW + wind * ACCEL_DW / (double)interpfac) / nh;
    }
    for (ii = 0; ii < num_harmonics; ii++) {
        const int hh = ii + 1;
        const double rr = *rout * hh, zz = *zout * hh, ww = *wout * hh;
        const double lp = (float)oracle_get_localpower3d(data, numdata, rr, zz, ww);
        oracle_get_derivs3d(data, numdata, rr, zz, ww, lp, &derivs[ii]);
        powers[ii] = derivs[ii].pow;
    }
    free(powsum);
}

/* calc_props (characteristics.c:193-265) */
int oracle_calc_props(const OracleRDerivs *d, double r, double z, double w, OracleFourierProps *res)
{
    double tmppow, tmpd2pow;
    if (fabs(d->pow) < 1e-14 || d->locpow < 1e-14) return -1;
    tmppow = d->pow / d->locpow;
    tmpd2pow = d->d2pow / d->locpow;
    res->pur = (float)(sqrt(1.5 * fabs(tmpd2pow) / tmppow) / M_PI);
    if (fabs(tmppow) < 1e-14) res->pur = 1.0f;
    res->purerr = (float)(1.0 / (res->pur * sqrt(10.0 * tmppow)));
    res->r = r;
    res->rerr = (float)(3.0 / (M_PI * res->pur * sqrt(6.0 * tmppow)));
    res->z = z;
    res->zerr = (float)(3.0 * sqrt(10.0) / (M_PI * res->pur * res->pur * sqrt(tmppow)));
    res->w = w;
    res->werr = (float)(6.0 * sqrt(105.0) / (M_PI * res->pur * res->pur * res->pur * sqrt(tmppow)));
    res->pow = (float)tmppow;
    res->powerr = (float)sqrt(2.0 * tmppow);
    res->sig = (float)sqrt(2.0 * tmppow - log(M_PI * tmppow));
    res->rawpow = (float)d->pow;
    res->phs = (float)d->phs;
    res->phserr = (float)(1.0 / res->powerr);
    res->cen = (float)(-d->dphs / (2.0 * M_PI));
    res->cenerr = (float)(1.0 / sqrt(24.0 * tmppow));
    res->locpow = (float)d->locpow;
    return 0;
}

/* ---------------- w = 0 searches: simplex in (r, z / 4) ------------------ */
#define ZSCALE 4.0

typedef struct {
    const float *data;
    int numdata, nh, khw;
    const float *locpows;
} RzFunc;

/* power_call_rz_harmonics (maximize_rz.c:99-113): minus the locpow-normalised
 * harmonic power sum at (r, z / ZSCALE) */
static double rz_func(const RzFunc *F, const double x[2])
{
    double total = 0.0;
    float ans[2];
    for (int ii = 0; ii < F->nh; ii++) {
        const int n = ii + 1;
        oracle_rz_interp(F->data, F->numdata, x[0] * n, x[1] * ZSCALE * n, F->khw, ans);
        total += power_d(ans) / F->locpows[ii];
    }
    return -total;
}

/* amotry (amoeba.c:90-106) */
static double amotry(const RzFunc *F, double p[3][2], double *y, double *psum, int ihi, double fac)
{
    const double fac1 = 0.5 * (1.0 - fac), fac2 = fac1 - fac;
    double ptry[2];
    ptry[0] = psum[0] * fac1 - p[ihi][0] * fac2;
    ptry[1] = psum[1] * fac1 - p[ihi][1] * fac2;
    const double ytry = rz_func(F, ptry);
    if (ytry < y[ihi]) {
        y[ihi] = ytry;
        psum[0] += ptry[0] - p[ihi][0];
        p[ihi][0] = ptry[0];
        psum[1] += ptry[1] - p[ihi][1];
        p[ihi][1] = ptry[1];
    }
    return ytry;
}

/* amoeba (amoeba.c:21-88): the 2-D downhill simplex, at most 5000 evaluations */
static void amoeba(const RzFunc *F, double p[3][2], double *y, double ftol)
{
    int ii, ihi, ilo, inhi, nfunk = 0;
    double rtol, ysave, ytry, psum[2], t;
    psum[0] = p[0][0] + p[1][0] + p[2][0];
    psum[1] = p[0][1] + p[1][1] + p[2][1];
    for (;;) {
        ilo = 0;
        if (y[0] > y[1]) {
            inhi = 1;
            ihi = 0;
        } else {
            inhi = 0;
            ihi = 1;
        }
        for (ii = 0; ii <= 2; ii++) {
            if (y[ii] <= y[ilo]) ilo = ii;
            if (y[ii] > y[ihi]) {
                inhi = ihi;
                ihi = ii;
            } else if (y[ii] > y[inhi] && ii != ihi) {
                inhi = ii;
            }
        }
        rtol = 2.0 * fabs(y[ihi] - y[ilo]) / (fabs(y[ihi]) + fabs(y[ilo]) + 1.0e-15);
        if (rtol < ftol) {
            t = y[0];
            y[0] = y[ilo];
            y[ilo] = t;
            for (ii = 0; ii < 2; ii++) {
                t = p[0][ii];
                p[0][ii] = p[ilo][ii];
                p[ilo][ii] = t;
            }
            break;
        }
        if (nfunk >= 5000) return;
        nfunk += 2;
        ytry = amotry(F, p, y, psum, ihi, -1.0);
        if (ytry <= y[ilo]) {
            ytry = amotry(F, p, y, psum, ihi, 2.0);
        } else if (ytry >= y[inhi]) {
            ysave = y[ihi];
            ytry = amotry(F, p, y, psum, ihi, 0.5);
            if (ytry >= ysave) {
                for (ii = 0; ii <= 2; ii++) {
                    if (ii != ilo) {
                        p[ii][0] = psum[0] = 0.5 * (p[ii][0] + p[ilo][0]);
                        p[ii][1] = psum[1] = 0.5 * (p[ii][1] + p[ilo][1]);
                        y[ii] = rz_func(F, psum);
                    }
                }
                nfunk += 2;
                psum[0] = p[0][0] + p[1][0] + p[2][0];
                psum[1] = p[0][1] + p[1][1] + p[2][1];
            }
        } else {
            --nfunk;
        }
    }
}

/* max_rz_arr_harmonics (maximize_rz.c:141-220) */
void oracle_max_rz_arr_harmonics(const float *data, int numdata, int num_harmonics, double rin, double zin,
                                 double *rout, double *zout, OracleRDerivs *derivs, double *powers)
{
    double y[3], x[3][2];
    const double step = 0.4;
    float *locpow = (float *)malloc(sizeof(float) * num_harmonics);
    int ii;
    for (ii = 0; ii < num_harmonics; ii++) {
        const int n = ii + 1;
        locpow[ii] = (float)oracle_get_localpower3d(data, numdata, rin * n, zin * n, 0.0);
    }
    RzFunc F = {data, numdata, num_harmonics, oracle_z_resp_halfwidth_acc(fabs(zin * num_harmonics) + 4.0, 0),
                locpow};
    x[0][0] = rin - step;
    x[0][1] = zin / ZSCALE - step;
    x[1][0] = rin - step;
    x[1][1] = zin / ZSCALE + step;
    x[2][0] = rin + step;
    x[2][1] = zin / ZSCALE;
    for (ii = 0; ii < 3; ii++) y[ii] = rz_func(&F, x[ii]);
    amoeba(&F, x, y, 1.0e-7);
    F.khw = oracle_z_resp_halfwidth_acc(fabs(x[0][1] * num_harmonics) + 4.0, 1);
    x[1][0] = x[0][0] + 0.01;
    x[1][1] = x[0][1];
    x[2][0] = x[0][0];
    x[2][1] = x[0][1] + 0.01;
    for (ii = 0; ii < 3; ii++) y[ii] = rz_func(&F, x[ii]);
    amoeba(&F, x, y, 1.0e-10);
    *rout = x[0][0];
    *zout = x[0][1] * ZSCALE;
    for (ii = 0; ii < num_harmonics; ii++) {
        const int n = ii + 1;
        float ans[2];
        locpow[ii] = (float)oracle_get_localpower3d(data, numdata, *rout * n, *zout * n, 0.0);
        /* power_call_rz (maximize_rz.c:10-18) at (rout n, zout n): the plain power */
        oracle_rz_interp(data, numdata, *rout * n, *zout / ZSCALE * n * ZSCALE, F.khw, ans);
        powers[ii] = power_d(ans);
        oracle_get_derivs3d(data, numdata, *rout * n, *zout * n, 0.0, locpow[ii], &derivs[ii]);
    }
    free(locpow);
}

/* max_rzw_arr (maximize_rzw.c:3-50): one harmonic, the 9 x 9 x 24 volume over
 * z +- ACCEL_DZ, w +- ACCEL_DW and three bins from int(rin) - 1, its first
 * raw-power maximum, and the derivatives there with the local power of the
 * input point.  Returns that (float) power. */
double oracle_max_rzw_arr(const float *data, int numdata, double rin, double zin, double win, double *rout,
                          double *zout, double *wout, OracleRDerivs *derivs)
{
    const float locpow = (float)oracle_get_localpower3d(data, numdata, rin, zin, win);
    float maxpow = 0.0f;
    const int extra = 10, startbin = (int)rin - 1, interpfac = 4;
    const int numbetween = interpfac * ACCEL_RDR, numr = 3 * numbetween, numz = 2 * interpfac + 1, numw = numz;
    const int khw = oracle_w_resp_halfwidth_acc(fabs(zin) + 2, fabs(win) + 20, 0);
    const int fftlen = oracle_next2_to_n(numbetween * (2 * khw + extra));
    int nextbin, ii, jj, kk, wind = 0, zind = 0, rind = 0;
    float *vol = (float *)malloc(sizeof(float) * 2 * (size_t)numw * numz * fftlen);
    const int ngood = oracle_corr_rzw_vol(data, numdata, numbetween, startbin, zin - ACCEL_DZ, zin + ACCEL_DZ, numz,
                                          win - ACCEL_DW, win + ACCEL_DW, numw, fftlen, &nextbin, vol);
    for (ii = 0; ii < numw; ii++)
        for (jj = 0; jj < numz; jj++)
            for (kk = 0; kk < numr && kk < ngood; kk++) {
                const float pw = power_f(vol + 2 * (((size_t)ii * numz + jj) * ngood + kk));
                if (pw > maxpow) {
                    maxpow = pw;
                    wind = ii;
                    zind = jj;
                    rind = kk;
                }
            }
    free(vol);
    *rout = startbin + rind * ACCEL_DR / (double)interpfac;
    *zout = zin - ACCEL_DZ + zind * ACCEL_DZ / (double)interpfac;
    *wout = win - ACCEL_DW + wind * ACCEL_DW / (double)interpfac;
    oracle_get_derivs3d(data, numdata, *rout, *zout, *wout, locpow, derivs);
    return maxpow;
}

/* max_rz_arr (maximize_rz.c:25-85): the simplex of max_rz_arr_harmonics on one
 * harmonic's raw power (power_call_rz, maximize_rz.c:10-23, divides by
 * nothing), the local power and derivatives at the peak.  Returns the peak's
 * raw power. */
double oracle_max_rz_arr(const float *data, int numdata, double rin, double zin, double *rout, double *zout,
                         OracleRDerivs *derivs)
{
    double y[3], x[3][2];
    const double step = 0.4;
    float one = 1.0f;
    RzFunc F = {data, numdata, 1, oracle_z_resp_halfwidth_acc(fabs(zin) + 4.0, 0), &one};
    int ii;
    x[0][0] = rin - step;
    x[0][1] = zin / ZSCALE - step;
    x[1][0] = rin - step;
    x[1][1] = zin / ZSCALE + step;
    x[2][0] = rin + step;
    x[2][1] = zin / ZSCALE;
    for (ii = 0; ii < 3; ii++) y[ii] = rz_func(&F, x[ii]);
    amoeba(&F, x, y, 1.0e-7);
    F.khw = oracle_z_resp_halfwidth_acc(fabs(x[0][1]) + 4.0, 1);
    x[1][0] = x[0][0] + 0.01;
    x[1][1] = x[0][1];
    x[2][0] = x[0][0];
    x[2][1] = x[0][1] + 0.01;
    for (ii = 0; ii < 3; ii++) y[ii] = rz_func(&F, x[ii]);
    amoeba(&F, x, y, 1.0e-10);
    *rout = x[0][0];
    *zout = x[0][1] * ZSCALE;
    const float lp = (float)oracle_get_localpower3d(data, numdata, *rout, *zout, 0.0);
    oracle_get_derivs3d(data, numdata, *rout, *zout, 0.0, lp, derivs);
    return -y[0];
}
