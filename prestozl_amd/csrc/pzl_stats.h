// pzl_stats.h -- detection thresholds and candidate significances, usable on
// the host and in HIP kernels (double precision in both).
//
//   pzl_power_for_sigma    characteristics.c:670-705: the summed power whose
//                          chi-square tail (2*numsum dof), times numtrials,
//                          equals the Gaussian tail of 'sigma'.
//   pzl_candidate_sigma    characteristics.c:647-668 via chi2_logp2 :588-628
//                          (A&S 26.4.19 asymptotics for chi2/dof > 15, else the
//                          tail gratio2 evaluates for integer a,
//                          dcdflib.c:2366-2475) and equivalent_gaussian_sigma2
//                          :501-544 (A&S 26.2.23 below log p = -600, else the
//                          inverse normal tail).
#pragma once

#include <math.h>

#include "pzl_geom.h"

PZL_HD double pzl_upper_normal_tail(double x) { return 0.5 * erfc(x * 0.70710678118654752440); }

// rational approximation to the normal quantile in t = sqrt(-2 log p) (A&S 26.2.23)
PZL_HD double pzl_as26223(double t)
{
    return t - (2.515517 + t * (0.802853 + t * 0.010328)) / (1.0 + t * (1.432788 + t * (0.189269 + t * 0.001308)));
}

// x with pzl_upper_normal_tail(x) == q, safeguarded Newton to full precision
PZL_HD double pzl_inverse_upper_normal_tail(double q)
{
    if (q <= 0.0) return 40.0;
    if (q >= 1.0) return -40.0;
    double x = (q < 0.5) ? pzl_as26223(sqrt(-2.0 * log(q))) : -pzl_as26223(sqrt(-2.0 * log1p(-q)));
    double lo = -40.0, hi = 40.0;
    for (int it = 0; it < 100; ++it) {
        const double f = pzl_upper_normal_tail(x) - q;
        if (f > 0) lo = x;
        else hi = x;
        const double dens = exp(-0.5 * x * x) * 0.39894228040143267794;
        if (!(dens > 0.0)) break;
        double xn = x + f / dens;
        if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
        const bool done = fabs(xn - x) <= 1e-15 * fabs(xn);
        x = xn;
        if (done) break;
    }
    return x;
}

// log upper tail of chi-square with 2a dof at 2x (A&S 6.5.32 over 6.1.41)
PZL_HD double pzl_log_q_asymptotic(double a, double x)
{
    double sum = 1.0, part = 1.0, term = 1.0;
    for (int i = 1; fabs(part) > 1e-15; ++i) {
        term *= (a - i);
        part = term / pow(x, (double)i);
        sum += part;
    }
    const double log_upper_gamma = (a - 1.0) * log(x) - x + log(sum);
    const double y = 1.0 / (a * a);
    const double log_gamma = (a - 0.5) * log(a) - a + 0.91893853320467274178 +
                             (((-1.0 / 1680.0 * y + 1.0 / 1260.0) * y - 1.0 / 360.0) * y + 1.0 / 12.0) / a;
    return log_upper_gamma - log_gamma;
}

// Q(a, x) for integer a as gratio2 evaluates it: a Poisson sum for a < 20,
// the terminating large-x series otherwise (numharm = 32).
PZL_HD double pzl_q_integer_a(double a, double x)
{
    if (a * x == 0.0) return 0.0;
    if (a >= 20.0 && (a > x || x >= 31.0)) {
        const double r = exp(a * log(x) - x) / tgamma(a);
        if (r == 0.0) return 0.0;
        // the first terms kept, then added smallest first; both loops fully
        // unrolled so that wk stays in registers
        double wk[20], amn = a - 1.0, t = amn / x, sum;
        int n = 21;  // the term count where the first loop stopped
        wk[0] = t;
#pragma unroll
        for (int k = 2; k <= 20; ++k) {
            if (n == 21) {
                amn -= 1.0;
                t *= amn / x;
                if (fabs(t) <= 1e-3) n = k;
                else wk[k - 1] = t;
            }
        }
        sum = t;
        while (fabs(t) > 5e-15) {
            amn -= 1.0;
            t *= amn / x;
            sum += t;
        }
#pragma unroll
        for (int m = 20; m >= 1; --m)
            if (m <= n - 1) sum += wk[m - 1];
        return r / x * (1.0 + sum);
    }
    const int ia = (int)a;
    double term = exp(-x), sum = term;
    for (int n = 1; n < ia; ++n) {
        term = x * term / (double)n;
        sum += term;
    }
    return sum;
}

PZL_HD double pzl_chi2_logp(double chi2, double dof)
{
    if (chi2 <= 0.0) return -INFINITY;
    if (chi2 / dof > 15.0 || (dof > 150 && chi2 / dof > 6.0)) return pzl_log_q_asymptotic(0.5 * dof, 0.5 * chi2);
    return log(pzl_q_integer_a(0.5 * dof, 0.5 * chi2));
}

PZL_HD double pzl_equivalent_gaussian_sigma(double logp)
{
    const double x = (logp < -600.0) ? pzl_as26223(sqrt(-2.0 * logp)) : pzl_inverse_upper_normal_tail(exp(logp));
    return x < 0.0 ? 0.0 : x;
}

PZL_HD double pzl_candidate_sigma(double power, int numsum, double numtrials)
{
    if (power <= 0.0) return 0.0;
    const double logp = pzl_chi2_logp(2.0 * power, 2.0 * numsum) + log(numtrials);
    return pzl_equivalent_gaussian_sigma(logp);
}

#ifdef __cplusplus
extern "C" {
#endif
double pzl_power_for_sigma(double sigma, int numsum, double numtrials);
#ifdef __cplusplus
}
#endif
