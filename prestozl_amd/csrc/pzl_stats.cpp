// pzl_stats.cpp -- the host-only threshold solver (power_for_sigma,
// characteristics.c:670-705).  Significances are in pzl_stats.h (host and
// device).
#include <cmath>

#include "pzl_stats.h"

namespace {

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

extern "C" double pzl_power_for_sigma(double sigma, int numsum, double numtrials)
{
    const double logq = std::log(pzl_upper_normal_tail(sigma)) - std::log(numtrials);
    double lo = 0.0, hi = 1.0;
    while (log_chi2_sf_even(hi, numsum) > logq) hi *= 2.0;
    for (int it = 0; it < 200 && hi - lo > 1e-15 * hi; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (log_chi2_sf_even(mid, numsum) > logq) lo = mid;
        else hi = mid;
    }
    return 0.25 * (lo + hi);
}
