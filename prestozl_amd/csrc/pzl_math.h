// pzl_math.h -- response-kernel mathematics shared by the host library and
// the kernel-bank builder on the GPU.
//   r/z/w response half-widths        responses.c:27-84 (LOWACC)
//   FFT length tables                 corr_prep.c:3-35
//   Fresnel integrals S(x), C(x)      (pi/2 convention of fresnl.c:176)
// The Fresnel integrals use the textbook series / erfc continued-fraction
// pair (Abramowitz & Stegun 7.3.11-7.3.13, modified Lentz evaluation).
#pragma once

#include "pzl_geom.h"

#define PZL_NUMFINTBINS 16

PZL_HD int pzl_next2_to_n(int x)
{
    int i = 1;
    while (i < x) i <<= 1;
    return i;
}

PZL_HD int pzl_next_good_fftlen(int n)
{
    const int lens[17] = {128, 192, 256, 384, 512, 768, 1024, 1280, 2048, 4096,
                          5120, 7680, 10240, 12288, 15360, 16384, 25600};
    int ii = 0;
    if (n <= lens[0]) return lens[0];
    if (n > lens[16]) return pzl_next2_to_n(n);
    while (n > lens[ii]) ii++;
    return lens[ii];
}

PZL_HD int pzl_fftlen_from_kernwidth(int kw)
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
    return pzl_next2_to_n(kw * 5);
}

// responses.c:27-84; highacc: HIGHACC (three times the Fourier-interpolation
// bins) instead of LOWACC
PZL_HD int pzl_z_resp_halfwidth_acc(double z, bool highacc)
{
    return (int)(0.5 * 1.1 * fabs(z)) + (highacc ? 3 * PZL_NUMFINTBINS : PZL_NUMFINTBINS);
}

PZL_HD int pzl_w_resp_halfwidth_acc(double z, double w, bool highacc)
{
    if (fabs(w) < 1.0e-7) return pzl_z_resp_halfwidth_acc(z, highacc);
    double r0 = 0.5 * (w / 6.0 - z);
    double r1 = 0.5 * (w / 6.0 + z);
    double maxdev = fabs(r0) > fabs(r1) ? fabs(r0) : fabs(r1);
    double u_ext = 0.5 - z / w;
    if (u_ext > 0.0 && u_ext < 1.0) {
        double z0 = z - w / 2.0;
        double r_ext = 0.5 * w * u_ext * u_ext + z0 * u_ext + r0;
        maxdev = fabs(r_ext) > maxdev ? fabs(r_ext) : maxdev;
    }
    return (int)(1.1 * maxdev) + (highacc ? 3 * PZL_NUMFINTBINS : PZL_NUMFINTBINS);
}

PZL_HD int pzl_z_resp_halfwidth(double z) { return (int)(0.5 * 1.1 * fabs(z)) + PZL_NUMFINTBINS; }

PZL_HD int pzl_w_resp_halfwidth(double z, double w)
{
    if (fabs(w) < 1.0e-7) return pzl_z_resp_halfwidth(z);
    double r0 = 0.5 * (w / 6.0 - z);
    double r1 = 0.5 * (w / 6.0 + z);
    double maxdev = fabs(r0) > fabs(r1) ? fabs(r0) : fabs(r1);
    double u_ext = 0.5 - z / w;
    if (u_ext > 0.0 && u_ext < 1.0) {
        double z0 = z - w / 2.0;
        double r_ext = 0.5 * w * u_ext * u_ext + z0 * u_ext + r0;
        maxdev = fabs(r_ext) > maxdev ? fabs(r_ext) : maxdev;
    }
    return (int)(1.1 * maxdev) + PZL_NUMFINTBINS;
}

// Fresnel integrals: C(x) = int_0^x cos(pi t^2 / 2) dt, S(x) likewise with sin.
PZL_HD void pzl_fresnel(double x, double *s, double *c)
{
    const double PI = 3.14159265358979323846;
    double ax = fabs(x);
    if (ax < 1e-150) {
        *s = 0.0;
        *c = ax;
    } else if (ax <= 1.5) {
        double fact = 0.5 * PI * ax * ax, term = ax, sc = ax, ss = 0.0;
        for (int k = 1; k < 200; k++) {
            term *= fact / k;
            double contrib = term / (2 * k + 1);
            double sgn = ((k >> 1) & 1) ? -1.0 : 1.0;
            if (k & 1) ss += sgn * contrib;
            else sc += sgn * contrib;
            if (term < 1e-17 * (fabs(sc) + fabs(ss))) break;
        }
        *s = ss;
        *c = sc;
    } else {
        // h = continued fraction for the erfc-based auxiliary; complex in (re, im)
        double pix2 = PI * ax * ax;
        double br = 1.0, bi = -pix2;
        double ccr = 1.0e30, cci = 0.0;
        double den = br * br + bi * bi;
        double dr = br / den, di = -bi / den;
        double hr = dr, hi = di;
        int n = -1;
        for (int k = 2; k < 1000; k++) {
            n += 2;
            double a = -(double)(n * (n + 1));
            br += 4.0;
            // d = 1 / (a d + b)
            double tr = a * dr + br, ti = a * di + bi;
            den = tr * tr + ti * ti;
            dr = tr / den;
            di = -ti / den;
            // cc = b + a / cc
            den = ccr * ccr + cci * cci;
            double qr = a * ccr / den, qi = -a * cci / den;
            ccr = br + qr;
            cci = bi + qi;
            // del = cc * d ; h *= del
            double delr = ccr * dr - cci * di, deli = ccr * di + cci * dr;
            double nhr = hr * delr - hi * deli, nhi = hr * deli + hi * delr;
            hr = nhr;
            hi = nhi;
            if (fabs(delr - 1.0) + fabs(deli) < 1e-16) break;
        }
        // h *= (ax - i ax)
        double h2r = hr * ax + hi * ax, h2i = hi * ax - hr * ax;
        double cs = cos(0.5 * pix2), sn = sin(0.5 * pix2);
        // e = (cs + i sn) * h2 ; t = 1 - e ; out = (0.5 + 0.5 i) t
        double er = cs * h2r - sn * h2i, ei = cs * h2i + sn * h2r;
        double tr = 1.0 - er, ti = -ei;
        *c = 0.5 * tr - 0.5 * ti;
        *s = 0.5 * tr + 0.5 * ti;
    }
    if (x < 0.0) {
        *c = -*c;
        *s = -*s;
    }
}
