/*
 * fft.c -- the oracle's CPU FFT (TEST INFRASTRUCTURE ONLY, see oracle.h).
 *
 * The reference calls FFTW through COMPLEXFFT = fftwcall (fftcalls.c:71-135,
 * include/ransomfft.h:37-39) for the segment and kernel FFTs and
 * fftwf_execute_dft (accel_utils.c:1222-1269) for the correlation IFFTs.
 * FFTW3f is not available here, so this file provides the same transform:
 *   X[k] = sum_j x[j] exp(sign * 2 pi i j k / n), unnormalised,
 * as a self-sorting (Stockham) mixed-radix DIF: radix 4, 2, 3 and a direct
 * DFT for any other prime factor.  Twiddles are computed in double and
 * rounded to the working precision, as FFTW does.
 *
 * realfft() follows fastffts.c:163-235 (the Numerical Recipes real-FFT
 * unpacking with the corrected exponent sign), forward and inverse.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define MAXSTAGES 40

typedef struct fft_plan {
    int n, sign, nstages;
    int radix[MAXSTAGES];
    int len[MAXSTAGES];    /* transform length handled at this stage */
    int stride[MAXSTAGES]; /* s at this stage */
    size_t twoff[MAXSTAGES];
    double *twd;           /* twiddles, interleaved re/im, double */
    float *twf;            /* same rounded to float */
    double *rootd[MAXSTAGES]; /* p-th roots for generic radix */
    float *rootf[MAXSTAGES];
} fft_plan;

static pthread_mutex_t plan_lock = PTHREAD_MUTEX_INITIALIZER;
static fft_plan *plans[256];
static int nplans = 0;

static fft_plan *make_plan(int n, int sign)
{
    fft_plan *pl = (fft_plan *)calloc(1, sizeof(fft_plan));
    int nn = n, st = 0, s = 1;
    size_t ntw = 0;
    pl->n = n;
    pl->sign = sign;
    while (nn > 1) {
        int p;
        if (nn % 4 == 0) p = 4;
        else if (nn % 2 == 0) p = 2;
        else if (nn % 3 == 0) p = 3;
        else {
            p = 5;
            while (nn % p) p += 2;
        }
        pl->radix[st] = p;
        pl->len[st] = nn;
        pl->stride[st] = s;
        pl->twoff[st] = ntw;
        ntw += (size_t)(nn / p) * (p - 1);
        nn /= p;
        s *= p;
        st++;
    }
    pl->nstages = st;
    pl->twd = (double *)malloc(sizeof(double) * 2 * (ntw + 1));
    pl->twf = (float *)malloc(sizeof(float) * 2 * (ntw + 1));
    for (st = 0; st < pl->nstages; st++) {
        int p = pl->radix[st], L = pl->len[st], m = L / p, k, t;
        for (k = 0; k < m; k++)
            for (t = 1; t < p; t++) {
                size_t idx = pl->twoff[st] + (size_t)k * (p - 1) + (t - 1);
                /* exact angle reduction: (k*t) mod L in integers */
                long long kt = ((long long)k * t) % L;
                double ang = sign * 2.0 * M_PI * (double)kt / (double)L;
                pl->twd[2 * idx] = cos(ang);
                pl->twd[2 * idx + 1] = sin(ang);
                pl->twf[2 * idx] = (float)pl->twd[2 * idx];
                pl->twf[2 * idx + 1] = (float)pl->twd[2 * idx + 1];
            }
        if (p != 2 && p != 4) {
            int r;
            pl->rootd[st] = (double *)malloc(sizeof(double) * 2 * p);
            pl->rootf[st] = (float *)malloc(sizeof(float) * 2 * p);
            for (r = 0; r < p; r++) {
                double ang = sign * 2.0 * M_PI * (double)r / (double)p;
                pl->rootd[st][2 * r] = cos(ang);
                pl->rootd[st][2 * r + 1] = sin(ang);
                pl->rootf[st][2 * r] = (float)cos(ang);
                pl->rootf[st][2 * r + 1] = (float)sin(ang);
            }
        }
    }
    return pl;
}

static fft_plan *get_plan(int n, int sign)
{
    int i;
    fft_plan *pl = NULL;
    pthread_mutex_lock(&plan_lock);
    for (i = 0; i < nplans; i++)
        if (plans[i]->n == n && plans[i]->sign == sign) {
            pl = plans[i];
            break;
        }
    if (!pl) {
        pl = make_plan(n, sign);
        if (nplans < 256) plans[nplans++] = pl;
    }
    pthread_mutex_unlock(&plan_lock);
    return pl;
}

/* One generic implementation, instantiated for float and double. */
#define DEFINE_CFFT(SUFFIX, REAL, TW, ROOT)                                                  \
    void oracle_cfft_##SUFFIX(REAL *x, int n, int sign)                                       \
    {                                                                                         \
        fft_plan *pl;                                                                         \
        REAL *y, *src, *dst;                                                                  \
        int st;                                                                               \
        if (n <= 1) return;                                                                   \
        pl = get_plan(n, sign);                                                               \
        y = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)n);                                     \
        src = x;                                                                              \
        dst = y;                                                                              \
        for (st = 0; st < pl->nstages; st++) {                                                \
            const int p = pl->radix[st], L = pl->len[st], m = L / p, s = pl->stride[st];     \
            const REAL *tw = pl->TW + 2 * pl->twoff[st];                                      \
            int k, q;                                                                         \
            for (k = 0; k < m; k++) {                                                         \
                for (q = 0; q < s; q++) {                                                     \
                    REAL ar[64], ai[64], br[64], bi[64];                                      \
                    int r, t;                                                                 \
                    for (r = 0; r < p; r++) {                                                 \
                        size_t ii = (size_t)q + (size_t)s * (k + (size_t)r * m);              \
                        ar[r] = src[2 * ii];                                                  \
                        ai[r] = src[2 * ii + 1];                                              \
                    }                                                                         \
                    if (p == 2) {                                                             \
                        br[0] = ar[0] + ar[1];                                                \
                        bi[0] = ai[0] + ai[1];                                                \
                        br[1] = ar[0] - ar[1];                                                \
                        bi[1] = ai[0] - ai[1];                                                \
                    } else if (p == 4) {                                                      \
                        REAL s0r = ar[0] + ar[2], s0i = ai[0] + ai[2];                        \
                        REAL d0r = ar[0] - ar[2], d0i = ai[0] - ai[2];                        \
                        REAL s1r = ar[1] + ar[3], s1i = ai[1] + ai[3];                        \
                        REAL d1r = ar[1] - ar[3], d1i = ai[1] - ai[3];                        \
                        /* J = sign*i ; J*(x+iy) = sign*(-y + i x) */                         \
                        REAL jr = (REAL)(-sign) * d1i, ji = (REAL)sign * d1r;                 \
                        br[0] = s0r + s1r;                                                    \
                        bi[0] = s0i + s1i;                                                    \
                        br[2] = s0r - s1r;                                                    \
                        bi[2] = s0i - s1i;                                                    \
                        br[1] = d0r + jr;                                                     \
                        bi[1] = d0i + ji;                                                     \
                        br[3] = d0r - jr;                                                     \
                        bi[3] = d0i - ji;                                                     \
                    } else {                                                                  \
                        const REAL *rt = pl->ROOT[st];                                        \
                        for (t = 0; t < p; t++) {                                             \
                            REAL sr = 0, si = 0;                                              \
                            for (r = 0; r < p; r++) {                                         \
                                int e = (r * t) % p;                                          \
                                sr += ar[r] * rt[2 * e] - ai[r] * rt[2 * e + 1];              \
                                si += ar[r] * rt[2 * e + 1] + ai[r] * rt[2 * e];              \
                            }                                                                 \
                            br[t] = sr;                                                       \
                            bi[t] = si;                                                       \
                        }                                                                     \
                    }                                                                         \
                    for (t = 0; t < p; t++) {                                                 \
                        size_t oo = (size_t)q + (size_t)s * ((size_t)p * k + t);              \
                        if (t == 0) {                                                         \
                            dst[2 * oo] = br[0];                                              \
                            dst[2 * oo + 1] = bi[0];                                          \
                        } else {                                                              \
                            const REAL wr = tw[2 * ((size_t)k * (p - 1) + t - 1)];            \
                            const REAL wi = tw[2 * ((size_t)k * (p - 1) + t - 1) + 1];        \
                            dst[2 * oo] = br[t] * wr - bi[t] * wi;                            \
                            dst[2 * oo + 1] = br[t] * wi + bi[t] * wr;                        \
                        }                                                                     \
                    }                                                                         \
                }                                                                             \
            }                                                                                 \
            {                                                                                 \
                REAL *tmp = src;                                                              \
                src = dst;                                                                    \
                dst = tmp;                                                                    \
            }                                                                                 \
        }                                                                                     \
        if (src != x) memcpy(x, src, sizeof(REAL) * 2 * (size_t)n);                           \
        free(y);                                                                              \
    }

DEFINE_CFFT(f, float, twf, rootf)
DEFINE_CFFT(d, double, twd, rootd)

/* fastffts.c:163-235, isign = -1 branch. */
#define DEFINE_REALFFT(SUFFIX, REAL)                                                         \
    void oracle_realfft_fwd_##SUFFIX(REAL *idata, long n)                                     \
    {                                                                                         \
        long nby2 = n >> 1, il, ih;                                                           \
        double cc = -0.5, h1r, h1i, h2r, h2i, h2rwr, h2iwr, h2rwi, h2iwi;                     \
        double wr, wi, wpr, wpi, tmp1, theta = -2.0 * M_PI / (double)n;                      \
        REAL *d = idata;                                                                      \
        oracle_cfft_##SUFFIX(d, (int)nby2, -1);                                               \
        wr = cos(theta);                                                                      \
        wi = sin(theta);                                                                      \
        tmp1 = sin(0.5 * theta);                                                              \
        wpr = -2.0 * tmp1 * tmp1;                                                             \
        wpi = wi;                                                                             \
        il = 1;                                                                               \
        ih = nby2 - il;                                                                       \
        for (; il <= (n >> 2); il++, ih--) {                                                  \
            h1r = 0.5 * (d[2 * il] + d[2 * ih]);                                              \
            h1i = 0.5 * (d[2 * il + 1] - d[2 * ih + 1]);                                      \
            h2r = -cc * (d[2 * il + 1] + d[2 * ih + 1]);                                      \
            h2i = cc * (d[2 * il] - d[2 * ih]);                                               \
            h2rwr = h2r * wr;                                                                 \
            h2rwi = h2r * wi;                                                                 \
            h2iwr = h2i * wr;                                                                 \
            h2iwi = h2i * wi;                                                                 \
            d[2 * il] = (REAL)(h1r + h2rwr - h2iwi);                                          \
            d[2 * il + 1] = (REAL)(h1i + h2iwr + h2rwi);                                      \
            d[2 * ih] = (REAL)(h1r - h2rwr + h2iwi);                                          \
            d[2 * ih + 1] = (REAL)(-h1i + h2iwr + h2rwi);                                     \
            tmp1 = wr;                                                                        \
            wr = tmp1 * wpr - wi * wpi + wr;                                                  \
            wi = wi * wpr + tmp1 * wpi + wi;                                                  \
        }                                                                                     \
        tmp1 = d[0];                                                                          \
        d[0] = (REAL)(tmp1 + d[1]);                                                           \
        d[1] = (REAL)(tmp1 - d[1]);                                                           \
    }

DEFINE_REALFFT(f, float)
DEFINE_REALFFT(d, double)

/* fastffts.c:163-235, isign = +1 branch: the NR packing undone (with the
 * sign flips of bin n/4 before and after the unpacking loop, :186-188 and
 * :224-226), the unnormalised inverse FFT of n/2 points, then every value
 * times 2/n. */
#define DEFINE_REALFFT_INV(SUFFIX, REAL)                                                      \
    void oracle_realfft_inv_##SUFFIX(REAL *idata, long n)                                     \
    {                                                                                         \
        long nby2 = n >> 1, il, ih;                                                           \
        double cc = 0.5, h1r, h1i, h2r, h2i, h2rwr, h2iwr, h2rwi, h2iwi;                      \
        double wr, wi, wpr, wpi, tmp1, theta = 2.0 * M_PI / (double)n;                        \
        REAL *d = idata;                                                                      \
        if ((n + 2) % 4) d[2 * (n >> 2) + 1] = -d[2 * (n >> 2) + 1];                          \
        wr = cos(theta);                                                                      \
        wi = sin(theta);                                                                      \
        tmp1 = sin(0.5 * theta);                                                              \
        wpr = -2.0 * tmp1 * tmp1;                                                             \
        wpi = wi;                                                                             \
        il = 1;                                                                               \
        ih = nby2 - il;                                                                       \
        for (; il <= (n >> 2); il++, ih--) {                                                  \
            h1r = 0.5 * (d[2 * il] + d[2 * ih]);                                              \
            h1i = 0.5 * (d[2 * il + 1] - d[2 * ih + 1]);                                      \
            h2r = -cc * (d[2 * il + 1] + d[2 * ih + 1]);                                      \
            h2i = cc * (d[2 * il] - d[2 * ih]);                                               \
            h2rwr = h2r * wr;                                                                 \
            h2rwi = h2r * wi;                                                                 \
            h2iwr = h2i * wr;                                                                 \
            h2iwi = h2i * wi;                                                                 \
            d[2 * il] = (REAL)(h1r + h2rwr - h2iwi);                                          \
            d[2 * il + 1] = (REAL)(h1i + h2iwr + h2rwi);                                      \
            d[2 * ih] = (REAL)(h1r - h2rwr + h2iwi);                                          \
            d[2 * ih + 1] = (REAL)(-h1i + h2iwr + h2rwi);                                     \
            tmp1 = wr;                                                                        \
            wr = tmp1 * wpr - wi * wpi + wr;                                                  \
            wi = wi * wpr + tmp1 * wpi + wi;                                                  \
        }                                                                                     \
        if ((n + 2) % 4) d[2 * (n >> 2) + 1] = -d[2 * (n >> 2) + 1];                          \
        tmp1 = d[0];                                                                          \
        d[0] = (REAL)(0.5 * (tmp1 + d[1]));                                                   \
        d[1] = (REAL)(0.5 * (tmp1 - d[1]));                                                   \
        oracle_cfft_##SUFFIX(d, (int)nby2, 1);                                                \
        tmp1 = 2.0 / (double)n;                                                               \
        for (il = 0; il < n; il++) d[il] = (REAL)(d[il] * tmp1);                              \
    }

DEFINE_REALFFT_INV(f, float)
DEFINE_REALFFT_INV(d, double)
