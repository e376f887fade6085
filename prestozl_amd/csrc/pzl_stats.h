// pzl_stats.h -- thresholds and significances (host, double precision).
#pragma once

#ifdef __cplusplus
extern "C" {
#endif
double pzl_power_for_sigma(double sigma, int numsum, double numtrials);
double pzl_candidate_sigma(double power, int numsum, double numtrials);
double pzl_chi2_logp(double chi2, double dof);
double pzl_equivalent_gaussian_sigma(double logp);
#ifdef __cplusplus
}
#endif
