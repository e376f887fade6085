// pzl_stats.cpp -- detection thresholds and candidate significances.
//
//   pzl_power_for_sigma   characteristics.c:670-705: the summed power whose
//                         chi-square tail (2*numsum dof), times numtrials,
//                         equals the Gaussian tail of 'sigma'.
//   pzl_candidate_sigma   characteristics.c:647-668 with chi2_logp2 :588-628
//                         (A&S 26.4.19 asymptotics for chi2/dof > 15, else the
//                         exact tail gratio2 evaluates for integer a,
//                         dcdflib.c:2366-2475) and equivalent_gaussian_sigma2
//                         :501-544 (A&S 26.2.23 below log p = -600, else the
//                         inverse normal tail).
// Everything is double precision on the host; hits are rare, so the sigma of
// every hit is evaluated here after the GPU has compacted them.
#include <cmath>

#include "pzl_stats.h"

namespace {

const double kPi = 3.14159265358979323846;

double upper_normal_tail(double x) { return 0.5 * std::erfc(x / std::sqrt(2.0)); }

double as_26_2_23(double t)
{
    // rational approximation to the normal quantile in t = sqrt(-2 log p)
    return t - (2.515517 + t * (0.802853 + t * 0.010328)) /
                   (1.0 + t * (1.432788 + t * (0.189269 + t * 0.001308)));
}

// x with upper_normal_tail(x) == q, refined by safeguarded Newton steps.
double inverse_upper_normal_tail(double q)
{
    if (q <= 0.0) return 40.0;
    if (q >= 1.0) return -40.0;
    double x = (q < 0.5) ? as_26_2_23(std::sqrt(-2.0 * std::log(q))) : -as_26_2_23(std::sqrt(-2.0 * std::log1p(-q)));
    double lo = -40.0, hi = 40.0;
    for (int it = 0; it < 100; ++it) {
        const double f = upper_normal_tail(x) - q;
        if (f > 0) lo = x;
        else hi = x;
        const double dens = std::exp(-0.5 * x * x) / std::sqrt(2.0 * kPi);
        if (!(dens > 0.0)) break;
        double xn = x + f / dens;
        if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
        const bool done = std::fabs(xn - x) <= 1e-15 * std::fabs(xn);
        x = xn;
        if (done) break;
    }
    return x;
}

// log of the upper tail of chi-square with 2a dof at 2x, A&S 6.5.32 / 6.1.41
double log_q_asymptotic(double a, double x)
{
    double sum = 1.0, part = 1.0, term = 1.0;
    for (int i = 1; std::fabs(part) > 1e-15; ++i) {
        term *= (a - i);
        part = term / std::pow(x, i);
        sum += part;
    }
    const double log_upper_gamma = (a - 1.0) * std::log(x) - x + std::log(sum);
    const double y = 1.0 / (a * a);
    const double log_gamma = (a - 0.5) * std::log(a) - a + 0.5 * std::log(2.0 * kPi) +
                             (((-1.0 / 1680.0 * y + 1.0 / 1260.0) * y - 1.0 / 360.0) * y + 1.0 / 12.0) / a;
    return log_upper_gamma - log_gamma;
}

// Q(a, x) for integer a the way gratio2 evaluates it: a Poisson sum for
// a < 20, the terminating large-x series otherwise.
double q_integer_a(double a, double x)
{
    if (a * x == 0.0) return 0.0;
    if (a >= 20.0 && (a > x || x >= 31.0)) {
        const double r = std::exp(a * std::log(x) - x) / std::tgamma(a);
        if (r == 0.0) return 0.0;
        double wk[20], amn = a - 1.0, t = amn / x, sum;
        int n;
        wk[0] = t;
        for (n = 2; n <= 20; ++n) {
            amn -= 1.0;
            t *= amn / x;
            if (std::fabs(t) <= 1e-3) break;
            wk[n - 1] = t;
        }
        sum = t;
        while (std::fabs(t) > 5e-15) {
            amn -= 1.0;
            t *= amn / x;
            sum += t;
        }
        for (int m = n - 1; m >= 1; --m) sum += wk[m - 1];
        return r / x * (1.0 + sum);
    }
    const int ia = (int)a;
    double term = std::exp(-x), sum = term;
    for (int n = 1; n < ia; ++n) {
        term = x * term / (double)n;
        sum += term;
    }
    return sum;
}

// log chi-square survival for even dof 2k, evaluated stably in log space
double log_chi2_sf_even(double chi2, int k)
{
    const double hx = 0.5 * chi2;
    double terms[64], mx = -hx, lt = -hx;
    terms[0] = -hx;
    for (int n = 1; n < k; ++n) {
        lt += std::log(hx) - std::log((double)n);
        terms[n] = lt;
        if (lt > mx) mx = lt;
    }
    double s = 0.0;
    for (int n = 0; n < k; ++n) s += std::exp(terms[n] - mx);
    return mx + std::log(s);
}

}  // namespace

double pzl_chi2_logp(double chi2, double dof)
{
    if (chi2 <= 0.0) return -INFINITY;
    if (chi2 / dof > 15.0 || (dof > 150 && chi2 / dof > 6.0)) return log_q_asymptotic(0.5 * dof, 0.5 * chi2);
    return std::log(q_integer_a(0.5 * dof, 0.5 * chi2));
}

double pzl_equivalent_gaussian_sigma(double logp)
{
    const double x = (logp < -600.0) ? as_26_2_23(std::sqrt(-2.0 * logp)) : inverse_upper_normal_tail(std::exp(logp));
    return x < 0.0 ? 0.0 : x;
}

double pzl_candidate_sigma(double power, int numsum, double numtrials)
{
    if (power <= 0.0) return 0.0;
    const double logp = pzl_chi2_logp(2.0 * power, 2.0 * numsum) + std::log(numtrials);
    return pzl_equivalent_gaussian_sigma(logp);
}

double pzl_power_for_sigma(double sigma, int numsum, double numtrials)
{
    const double logq = std::log(upper_normal_tail(sigma)) - std::log(numtrials);
    double lo = 0.0, hi = 1.0;
    while (log_chi2_sf_even(hi, numsum) > logq) hi *= 2.0;
    for (int it = 0; it < 200 && hi - lo > 1e-15 * hi; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (log_chi2_sf_even(mid, numsum) > logq) lo = mid;
        else hi = mid;
    }
    return 0.25 * (lo + hi);
}
