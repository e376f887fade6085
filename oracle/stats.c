/*
 * stats.c -- thresholds and significances (TEST INFRASTRUCTURE ONLY).
 *
 * Restates characteristics.c:400-705 and the parts of dcdflib.c the search
 * reaches:
 *   power_for_sigma   characteristics.c:670-705  (cdfnor which=1, cdfchi which=2)
 *   candidate_sigma   characteristics.c:647-668  -> chi2_logp2 :588-628
 *                     -> cdfchi2 / gratio2 dcdflib.c:2336-2475
 *                     -> equivalent_gaussian_sigma2 :501-544 -> cdfnor2 dcdflib.c:4753
 *   median            median.c:12-69 (returns sorted element (n-1)/2)
 * The normal tail and its inverse are computed from libm's erfc to double
 * precision (dcdflib's cumnor/dinvnr are accurate to ~1e-13 relative); the
 * inverse chi-square is solved by bisection on the exact Poisson-sum tail
 * (dcdflib's dinvr stops at ~1e-8 relative).  Both are checked against the
 * reference's own dcdflib built into oracle/_ref.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Q(x) = P(X > x) for X ~ N(0,1) */
double oracle_normal_sf(double x) { return 0.5 * erfc(x / M_SQRT2); }

/* x such that Q(x) = q, 0 < q < 1 (Newton on erfc, bisection-safeguarded). */
static double normal_isf(double q)
{
    double lo = -40.0, hi = 40.0, x;
    int it;
    if (q <= 0.0) return 40.0;
    if (q >= 1.0) return -40.0;
    /* start: Acklam-free simple guess from the tail asymptote */
    if (q < 0.5) {
        double t = sqrt(-2.0 * log(q));
        x = t - (2.515517 + t * (0.802853 + t * 0.010328)) /
                    (1.0 + t * (1.432788 + t * (0.189269 + t * 0.001308)));
    } else {
        double t = sqrt(-2.0 * log(1.0 - q));
        x = -(t - (2.515517 + t * (0.802853 + t * 0.010328)) /
                      (1.0 + t * (1.432788 + t * (0.189269 + t * 0.001308))));
    }
    for (it = 0; it < 200; it++) {
        double f = oracle_normal_sf(x) - q;       /* decreasing in x */
        double dens = exp(-0.5 * x * x) / sqrt(2.0 * M_PI);
        double xn;
        if (f > 0) lo = x; else hi = x;
        if (dens <= 0.0) break;
        xn = x + f / dens;
        if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
        if (fabs(xn - x) <= 1e-15 * fabs(xn) + 1e-300) {
            x = xn;
            break;
        }
        x = xn;
    }
    return x;
}

/* A&S 26.2.23 on log(p): characteristics.c:400-417 */
static double extended_equiv_gaussian_sigma(double logp)
{
    double t = sqrt(-2.0 * logp);
    double num = 2.515517 + t * (0.802853 + t * 0.010328);
    double denom = 1.0 + t * (1.432788 + t * (0.189269 + t * 0.001308));
    return t - num / denom;
}

/* A&S 6.5.32: characteristics.c:419-440 */
static double log_asymtotic_incomplete_gamma(double a, double z)
{
    double x = 1.0, newxpart = 1.0, term = 1.0;
    int ii = 1;
    while (fabs(newxpart) > 1e-15) {
        term *= (a - ii);
        newxpart = term / pow(z, ii);
        x += newxpart;
        ii += 1;
    }
    return (a - 1.0) * log(z) - z + log(x);
}

/* A&S 6.1.41 (Stirling series to 1/z^7): characteristics.c:442-455 */
static double log_asymtotic_gamma(double z)
{
    double x, y;
    x = (z - 0.5) * log(z) - z + 0.5 * log(2.0 * M_PI);
    y = 1.0 / (z * z);
    x += (((-1.0 / 1680.0 * y + 1.0 / 1260.0) * y - 1.0 / 360.0) * y + 1.0 / 12.0) / z;
    return x;
}

/* Upper regularised incomplete gamma Q(a, x) as gratio2 computes it for
 * the a = numharm values the search uses (dcdflib.c:2366-2475). */
static double gratio2_q(double a, double x)
{
    if (a * x == 0.0) return 0.0;
    if (a >= 20.0 && (a > x || x >= 31.0)) {
        /* series branch (reached only for numharm = 32) */
        double wk[20], r, amn, t, sum;
        int n, m, maxn;
        r = exp(a * log(x) - x) / tgamma(a);
        if (r == 0.0) return 0.0;
        amn = a - 1.0;
        t = amn / x;
        wk[0] = t;
        for (n = 2; n <= 20; n++) {
            amn -= 1.0;
            t *= (amn / x);
            if (fabs(t) <= 1.e-3) break;
            wk[n - 1] = t;
        }
        sum = t;
        while (fabs(t) > 5e-15) {
            amn -= 1.0;
            t *= (amn / x);
            sum += t;
        }
        maxn = n - 1;
        for (m = 1; m <= maxn; m++) {
            n -= 1;
            sum += wk[n - 1];
        }
        return r / x * (1.0 + sum);
    }
    {
        /* integer a: finite Poisson sum in the order gratio2 adds it */
        int i = (int)a, n = 1;
        double sum = exp(-x), t = sum, c = 0.0;
        while (n != i) {
            n += 1;
            c += 1.0;
            t = x * t / c;
            sum += t;
        }
        return sum;
    }
}

double oracle_chi2_logp2(double chi2, double dof)
{
    if (chi2 <= 0.0) return -INFINITY;
    if (chi2 / dof > 15.0 || (dof > 150 && chi2 / dof > 6.0))
        return log_asymtotic_incomplete_gamma(0.5 * dof, 0.5 * chi2) -
               log_asymtotic_gamma(0.5 * dof);
    return log(gratio2_q(0.5 * dof, 0.5 * chi2));
}

double oracle_equivalent_gaussian_sigma2(double logp)
{
    double x;
    if (logp < -600.0) {
        x = extended_equiv_gaussian_sigma(logp);
    } else {
        double q = exp(logp);
        x = normal_isf(q);
    }
    return (x < 0.0) ? 0.0 : x;
}

double oracle_candidate_sigma(double power, int numsum, double numtrials)
{
    double logp;
    if (power <= 0.0) return 0.0;
    logp = oracle_chi2_logp2(2.0 * power, 2.0 * numsum);
    logp += log(numtrials);
    return oracle_equivalent_gaussian_sigma2(logp);
}

/* Exact chi-square tail for even dof = 2k: Q(k, x/2) as a Poisson sum,
 * evaluated in log space to stay accurate far in the tail. */
static double log_chi2_sf_even(double x, int k)
{
    double hx = 0.5 * x, logterm = -hx, lse = -hx, m;
    int n;
    /* log-sum-exp of log(e^-hx hx^n / n!), n = 0..k-1 */
    double terms[64];
    terms[0] = -hx;
    m = terms[0];
    for (n = 1; n < k; n++) {
        logterm += log(hx) - log((double)n);
        terms[n] = logterm;
        if (logterm > m) m = logterm;
    }
    lse = 0.0;
    for (n = 0; n < k; n++) lse += exp(terms[n] - m);
    return m + log(lse);
}

double oracle_power_for_sigma(double sigma, int numsum, double numtrials)
{
    /* q = Q(sigma) / numtrials; solve chi2_sf(x; 2*numsum) = q; return x/2 */
    double logq = log(oracle_normal_sf(sigma)) - log(numtrials);
    double lo = 0.0, hi = 1.0;
    int it;
    while (log_chi2_sf_even(hi, numsum) > logq) hi *= 2.0;
    for (it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi);
        if (log_chi2_sf_even(mid, numsum) > logq) lo = mid;
        else hi = mid;
        if (hi - lo <= 1e-15 * hi) break;
    }
    return 0.25 * (lo + hi);
}

static int cmp_float(const void *a, const void *b)
{
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* median.c:12-69 returns the element that would sit at index (n-1)/2 of the
 * sorted array (a lower median for even n). */
float oracle_median(const float *arr, int n)
{
    float *tmp = (float *)malloc(sizeof(float) * (size_t)n), med;
    memcpy(tmp, arr, sizeof(float) * (size_t)n);
    qsort(tmp, (size_t)n, sizeof(float), cmp_float);
    med = tmp[(n - 1) / 2];
    free(tmp);
    return med;
}
