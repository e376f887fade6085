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

// c[0] + c[1] x + ... + c[7] x^7
PZL_HD double pzl_poly7(const double *c, double x)
{
    double s = c[7];
    for (int i = 6; i >= 0; --i) s = s * x + c[i];
    return s;
}

// x with upper normal tail q (0 < q < 1) to about 1e-16 relative: Wichura's
// algorithm AS 241 (PPND16; Applied Statistics 37 (1988) 477-484), rational
// functions of 0.180625 - (1/2 - q)^2 near the centre and of
// sqrt(-log(smaller tail)) beyond.
PZL_HD double pzl_as241_upper(double q)
{
    const double a[8] = {3.3871328727963666080e0,  1.3314166789178437745e+2, 1.9715909503065514427e+3,
                         1.3731693765509461125e+4, 4.5921953931549871457e+4, 6.7265770927008700853e+4,
                         3.3430575583588128105e+4, 2.5090809287301226727e+3};
    const double b[8] = {1.0,
                         4.2313330701600911252e+1, 6.8718700749205790830e+2, 5.3941960214247511077e+3,
                         2.1213794301586595867e+4, 3.9307895800092710610e+4, 2.8729085735721942674e+4,
                         5.2264952788528545610e+3};
    const double c[8] = {1.42343711074968357734e0,  4.63033784615654529590e0,  5.76949722146069140550e0,
                         3.64784832476320460504e0,  1.27045825245236838258e0,  2.41780725177450611770e-1,
                         2.27238449892691845833e-2, 7.74545014278341407640e-4};
    const double d[8] = {1.0,
                         2.05319162663775882187e0,  1.67638483018380384940e0,  6.89767334985100004550e-1,
                         1.48103976427480074590e-1, 1.51986665636164571966e-2, 5.47593808499534494600e-4,
                         1.05075007164441684324e-9};
    const double e[8] = {6.65790464350110377720e0,  5.46378491116411436990e0,  1.78482653991729133580e0,
                         2.96560571828504891230e-1, 2.65321895265761230930e-2, 1.24266094738807843860e-3,
                         2.71155556874348757815e-5, 2.01033439929228813265e-7};
    const double f[8] = {1.0,
                         5.99832206555887937690e-1, 1.36929880922735805310e-1, 1.48753612908506148525e-2,
                         7.86869131145613259100e-4, 1.84631831751005468180e-5, 1.42151175831644588870e-7,
                         2.04426310338993978564e-15};
    const double h = 0.5 - q;  // the lower tail of x minus 1/2
    if (fabs(h) <= 0.425) {
        const double r = 0.180625 - h * h;
        return h * pzl_poly7(a, r) / pzl_poly7(b, r);
    }
    const double r = sqrt(-(h > 0.0 ? log(q) : log1p(-q)));
    const double v = r <= 5.0 ? pzl_poly7(c, r - 1.6) / pzl_poly7(d, r - 1.6) : pzl_poly7(e, r - 5.0) / pzl_poly7(f, r - 5.0);
    return h > 0.0 ? v : -v;
}

// x with pzl_upper_normal_tail(x) == q, safeguarded Newton to full precision
// from AS 241's value (one step, in practice, to meet the tolerance)
PZL_HD double pzl_inverse_upper_normal_tail(double q)
{
    if (q <= 0.0) return 40.0;
    if (q >= 1.0) return -40.0;
    double x = pzl_as241_upper(q);
    double lo = -40.0, hi = 40.0;
    for (int it = 0; it < 100; ++it) {
        const double f = pzl_upper_normal_tail(x) - q;
        if (f > 0) lo = x;
        else hi = x;
        const double dens = exp(-0.5 * x * x) * 0.39894228040143267794;
        if (!(dens > 0.0)) break;
        double xn = x + f / dens;
        // converged (a step of zero included: x is then a bracket end, which
        // the safeguard below would take for a step out of the bracket)
        if (fabs(xn - x) <= 1e-15 * fabs(xn)) {
            x = xn;
            break;
        }
        if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
        x = xn;
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
