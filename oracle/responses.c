/*
 * responses.c -- Fourier response kernels (TEST INFRASTRUCTURE ONLY).
 *
 * Restates responses.c:27-84 (half widths), :159-225 (gen_r_response),
 * :228-316 (gen_z_response), :333-418 (gen_w_response) and corr_prep.c:3-35
 * (FFT-length tables).  Fresnel integrals follow the textbook method
 * (power series for small x, modified-Lentz continued fraction for the
 * complementary error function otherwise, as in Numerical Recipes 6.9); the
 * reference's Cephes fresnl() (fresnl.c:176-217) is built into oracle/_ref
 * and the two are compared by the tests.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NUMFINTBINS 16

int oracle_next2_to_n(int x)
{
    int i = 1;
    while (i < x) i <<= 1;
    return i;
}

int oracle_next_good_fftlen(int n)
{
    static const int lens[17] = {128, 192, 256, 384, 512, 768, 1024, 1280, 2048, 4096,
                                 5120, 7680, 10240, 12288, 15360, 16384, 25600};
    int ii = 0;
    if (n <= lens[0]) return lens[0];
    if (n > lens[16]) return oracle_next2_to_n(n);
    while (n > lens[ii]) ii++;
    return lens[ii];
}

int oracle_fftlen_from_kernwidth(int kw)
{
    if (kw < 6) return 128;
    if (kw < 52) return 256;
    if (kw < 67) return 512;
    if (kw < 378) return 1024;
    if (kw < 664) return 2048;
    if (kw < 1672) return 4096;
    if (kw < 3015) return 10240;
    if (kw < 3554) return 15360;
    if (kw < 6000) return 25600;
    return oracle_next2_to_n(kw * 5);
}

int oracle_z_resp_halfwidth(double z) { return (int)(0.5 * 1.1 * fabs(z)) + NUMFINTBINS; }

int oracle_w_resp_halfwidth(double z, double w)
{
    double r0, r1, maxdev, u_ext;
    if (fabs(w) < 1.0e-7) return oracle_z_resp_halfwidth(z);
    r0 = 0.5 * (w / 6.0 - z);
    r1 = 0.5 * (w / 6.0 + z);
    maxdev = fabs(r0) > fabs(r1) ? fabs(r0) : fabs(r1);
    u_ext = 0.5 - z / w;
    if (u_ext > 0.0 && u_ext < 1.0) {
        double z0 = z - w / 2.0;
        double r_ext = 0.5 * w * u_ext * u_ext + z0 * u_ext + r0;
        maxdev = fabs(r_ext) > maxdev ? fabs(r_ext) : maxdev;
    }
    return (int)(1.1 * maxdev) + NUMFINTBINS;
}

void oracle_fresnel(double x, double *s, double *c)
{
    double ax = fabs(x);
    if (ax < 1e-150) {
        *s = 0.0;
        *c = ax;
    } else if (ax <= 1.5) {
        /* C(x) = sum_n (-1)^n (pi/2)^(2n) x^(4n+1) / ((2n)! (4n+1))
         * S(x) = sum_n (-1)^n (pi/2)^(2n+1) x^(4n+3) / ((2n+1)! (4n+3)) */
        double fact = 0.5 * M_PI * ax * ax, term = ax, sumc = ax, sums = 0.0;
        int k;
        for (k = 1; k < 200; k++) {
            double contrib;
            term *= fact / k;          /* (pi/2 x^2)^k x / k! */
            contrib = term / (2 * k + 1);
            /* k odd -> S term, k even -> C term; sign (-1)^(floor(k/2)) */
            if (k & 1) sums += ((k >> 1) & 1 ? -contrib : contrib);
            else sumc += ((k >> 1) & 1 ? -contrib : contrib);
            if (term < 1e-17 * (fabs(sumc) + fabs(sums))) break;
        }
        *s = sums;
        *c = sumc;
    } else {
        /* (C + iS) = (1+i)/2 * (1 - e^{i pi x^2/2} * h * (x - i x)) with h from
         * the erfc continued fraction, modified Lentz. */
        double pix2 = M_PI * ax * ax;
        double complex b = 1.0 - I * pix2, cc = 1.0e300, d, h, del;
        int n = -1, k;
        d = h = 1.0 / b;
        for (k = 2; k < 1000; k++) {
            double a;
            n += 2;
            a = -(double)(n * (n + 1));
            b += 4.0;
            d = 1.0 / (a * d + b);
            cc = b + a / cc;
            del = cc * d;
            h *= del;
            if (fabs(creal(del) - 1.0) + fabs(cimag(del)) < 1e-16) break;
        }
        h *= (ax - I * ax);
        {
            double complex cs = (0.5 + 0.5 * I) * (1.0 - (cos(0.5 * pix2) + I * sin(0.5 * pix2)) * h);
            *c = creal(cs);
            *s = cimag(cs);
        }
    }
    if (x < 0.0) {
        *c = -*c;
        *s = -*s;
    }
}

static void store(void *out, int mode, int ii, double re, double im)
{
    if (mode == ORACLE_FP64) {
        ((double *)out)[2 * ii] = re;
        ((double *)out)[2 * ii + 1] = im;
    } else {
        ((float *)out)[2 * ii] = (float)re;
        ((float *)out)[2 * ii + 1] = (float)im;
    }
}

/* responses.c:159-225.  sin/cos are evaluated directly rather than by the
 * reference's rotation recurrence (a ~1e-15 relative difference). */
void oracle_gen_r_response(double roffset, int numbetween, int numkern, int mode, void *out)
{
    double startr = M_PI * (numkern / (double)(2 * numbetween) + roffset);
    double delta = -M_PI / numbetween;
    int ii;
    for (ii = 0; ii < numkern; ii++) {
        double r = startr + ii * delta, s = sin(r), c = cos(r), sinc;
        sinc = (r == 0.0) ? 1.0 : s / r;
        store(out, mode, ii, c * sinc, s * sinc);
    }
    if (roffset < 1E-3) {
        double t = roffset * roffset;
        store(out, mode, numkern / 2, 1 - 6.579736267392905746 * t,
              roffset * (M_PI - 10.335425560099940058 * t));
    }
}

/* responses.c:228-316 */
void oracle_gen_z_response(double roffset, int numbetween, double z, int numkern, int mode, void *out)
{
    double absz = fabs(z), startr, startroffset, zd, cons, pibyz, delta, r, tmprl;
    int ii, signz;
    if (absz < 1E-4) {
        oracle_gen_r_response(roffset, numbetween, numkern, mode, out);
        return;
    }
    startr = roffset - (0.5 * z);
    startroffset = (startr < 0) ? 1.0 + modf(startr, &tmprl) : modf(startr, &tmprl);
    signz = (z < 0.0) ? -1 : 1;
    zd = signz * M_SQRT2 / sqrt(absz);
    cons = zd / 2.0;
    pibyz = M_PI / z;
    startr += numkern / (double)(2 * numbetween);
    delta = -1.0 / numbetween;
    for (ii = 0, r = startr; ii < numkern; ii++, r += delta) {
        double yy = r * zd, zz = yy + z * zd, xx = pibyz * r * r;
        double c = cos(xx), s = sin(xx), fressy, frescy, fressz, frescz, tr, ti;
        oracle_fresnel(yy, &fressy, &frescy);
        oracle_fresnel(zz, &fressz, &frescz);
        tr = signz * (frescz - frescy);
        ti = fressy - fressz;
        store(out, mode, ii, (tr * c - ti * s) * cons, -(tr * s + ti * c) * cons);
    }
    if (startroffset < 1E-3 && absz < 1E-3) {
        double zz2 = z * z, xx = startroffset * startroffset, re, im;
        int h = numkern / 2;
        re = 1.0 - 0.16449340668482264365 * zz2;
        im = -0.5235987755982988731 * z;
        re += startroffset * 1.6449340668482264365 * z;
        im += startroffset * (M_PI - 0.5167712780049970029 * zz2);
        re += xx * (-6.579736267392905746 + 0.9277056288952613070 * zz2);
        im += xx * (3.1006276680299820175 * z);
        store(out, mode, h, re, im);
    }
}

/* responses.c:333-418.  FP32 mode keeps the reference's float data array and
 * float real FFT; FP64 mode runs the same steps in double. */
void oracle_gen_w_response(double roffset, int numbetween, double z, double w, int numkern,
                           int mode, void *out)
{
    int ii, fbar, npts;
    double amp, f, fd, fdd, dt, dfbar;
    if (fabs(w) < 1E-4) {
        oracle_gen_z_response(roffset, numbetween, z, numkern, mode, out);
        return;
    }
    npts = oracle_next2_to_n(6 * oracle_w_resp_halfwidth(z, w) + 200 + numkern / numbetween);
    dt = 1.0 / (double)npts;
    amp = 2.0 * dt;
    fbar = npts / 4;
    dfbar = (double)fbar + roffset;
    f = dfbar - 0.5 * z + w / 12.0;
    fd = (z - 0.5 * w) / 2.0;
    fdd = w / 6.0;
    if (mode == ORACLE_FP64) {
        double *data = (double *)calloc((size_t)npts * numbetween, sizeof(double));
        for (ii = 0; ii < npts; ii++) {
            double t = ii * dt;
            data[ii] = amp * cos(2.0 * M_PI * (t * (t * (t * fdd + fd) + f)));
        }
        oracle_realfft_fwd_d(data, (long)npts * numbetween);
        memcpy(out, data + 2 * (fbar * numbetween - numkern / 2), sizeof(double) * 2 * numkern);
        free(data);
    } else {
        float *data = (float *)calloc((size_t)npts * numbetween, sizeof(float));
        for (ii = 0; ii < npts; ii++) {
            double t = ii * dt;
            data[ii] = (float)(amp * cos(2.0 * M_PI * (t * (t * (t * fdd + fd) + f))));
        }
        oracle_realfft_fwd_f(data, (long)npts * numbetween);
        memcpy(out, data + 2 * (fbar * numbetween - numkern / 2), sizeof(float) * 2 * numkern);
        free(data);
    }
}
