// pzl_response.h -- Fourier-interpolation responses for one kernel row,
// computed by the threads of a workgroup in double precision:
// gen_r_response / gen_z_response / gen_w_response (responses.c:159-418)
// with their fractional offset 'roffset' and 'numbetween' points per bin.
//
// Thread tid holds the values m = tid + i * NT (i < MAXV, m < numkern).  The w
// response samples a cubic-phase chirp of npts points (float, as the
// reference) into 'samp' (LDS, >= npts floats) and evaluates the numkern bins
// of its numbetween-times zero-padded real FFT that the reference keeps, each
// as a direct DFT over the npts nonzero samples.  Every thread of the
// workgroup must call; the function ends with a barrier, after which 'samp'
// may be reused.
#pragma once

#include <hip/hip_runtime.h>

#include "pzl_math.h"

// Value m (< numkern) of gen_z_response (responses.c:228-316) for
// |z| >= 1e-4, else of gen_r_response (responses.c:159-225).
__device__ inline double2 pzl_zr_response_at(double roffset, int numbetween, double z, int numkern, int m)
{
    const double PI = 3.14159265358979323846;
    if (fabs(z) >= 1e-4) {
        const double absz = fabs(z);
        const int signz = (z < 0.0) ? -1 : 1;
        const double zd = signz * 1.41421356237309504880 / sqrt(absz);
        const double cons = zd / 2.0, pibyz = PI / z;
        double startr = roffset - 0.5 * z;
        double ip;
        const double startroffset = (startr < 0) ? 1.0 + modf(startr, &ip) : modf(startr, &ip);
        startr += numkern / (double)(2 * numbetween);
        const double delta = -1.0 / numbetween;
        const double r = startr + m * delta;
        const double yy = r * zd, zz = yy + z * zd, xx = pibyz * r * r;
        double fsy, fcy, fsz, fcz;
        pzl_fresnel(yy, &fsy, &fcy);
        pzl_fresnel(zz, &fsz, &fcz);
        const double c = cos(xx), s = sin(xx);
        const double tr = signz * (fcz - fcy), ti = fsy - fsz;
        double re = (tr * c - ti * s) * cons;
        double im = -(tr * s + ti * c) * cons;
        if (m == numkern / 2 && startroffset < 1E-3 && absz < 1E-3) {
            const double z2 = z * z, x2 = startroffset * startroffset;
            re = 1.0 - 0.16449340668482264365 * z2 + startroffset * 1.6449340668482264365 * z +
                 x2 * (-6.579736267392905746 + 0.9277056288952613070 * z2);
            im = -0.5235987755982988731 * z + startroffset * (PI - 0.5167712780049970029 * z2) +
                 x2 * (3.1006276680299820175 * z);
        }
        return double2{re, im};
    }
    // sinc(r) e^{i r}, r = pi (numkern / (2 nb) + roffset - m / nb)
    const double x = numkern / (double)(2 * numbetween) + roffset - m / (double)numbetween;  // r / pi
    if (m == numkern / 2 && roffset < 1E-3) {
        const double r2 = roffset * roffset;
        return double2{1.0 - 6.579736267392905746 * r2, roffset * (PI - 10.335425560099940058 * r2)};
    }
    if (x == 0.0) return double2{1.0, 0.0};
    double s, c;
    sincospi(x, &s, &c);
    const double sinc = s / (PI * x);
    return double2{c * sinc, s * sinc};
}

template <int NT, int MAXV>
__device__ void pzl_gen_response(double roffset, int numbetween, double z, double w, int numkern, int tid,
                                 float *samp, double2 (&resp)[MAXV])
{
    if (fabs(w) >= 1e-4) {
        // gen_w_response (responses.c:333-418)
        const int npts = pzl_next2_to_n(6 * pzl_w_resp_halfwidth(z, w) + 200 + numkern / numbetween);
        const double dt = 1.0 / (double)npts, amp = 2.0 * dt;
        const int fbar = npts / 4;
        const double f = (double)fbar + roffset - 0.5 * z + w / 12.0, fd = (z - 0.5 * w) / 2.0, fdd = w / 6.0;
        for (int t = tid; t < npts; t += NT) {
            const double tt = t * dt;
            samp[t] = (float)(amp * cos(6.283185307179586477 * (tt * (tt * (tt * fdd + fd) + f))));
        }
        __syncthreads();
        const long long nfft = (long long)npts * numbetween;  // real FFT length
        const long long k0 = (long long)fbar * numbetween - numkern / 2;
#pragma unroll
        for (int i = 0; i < MAXV; i++) {
            const int m = tid + i * NT;
            double ar = 0.0, ai = 0.0;
            if (m < numkern) {
                const long long k = k0 + m;
                // X[k] = sum_t samp[t] exp(-2 pi i t k / nfft); the rotation is
                // re-anchored exactly every 64 samples
                double sr, cr;
                sincospi(-2.0 * (double)k / (double)nfft, &sr, &cr);
                for (int t0 = 0; t0 < npts; t0 += 64) {
                    double s0, c0;
                    const long long ph = ((long long)t0 * k) % nfft;
                    sincospi(-2.0 * (double)ph / (double)nfft, &s0, &c0);
                    const int tend = t0 + 64 < npts ? t0 + 64 : npts;
                    for (int t = t0; t < tend; t++) {
                        const double d = (double)samp[t];
                        ar += d * c0;
                        ai += d * s0;
                        const double nc = c0 * cr - s0 * sr;
                        s0 = c0 * sr + s0 * cr;
                        c0 = nc;
                    }
                }
            }
            resp[i] = double2{ar, ai};
        }
        __syncthreads();
    } else {
#pragma unroll
        for (int i = 0; i < MAXV; i++) {
            const int m = tid + i * NT;
            resp[i] = (m < numkern) ? pzl_zr_response_at(roffset, numbetween, z, numkern, m) : double2{0.0, 0.0};
        }
    }
}
