// pzl_optimize.cpp -- host side of the candidate optimisation
// (optimize_accelcand with harmonic polishing, accel_utils.c:573-612): the
// per-harmonic geometry of max_rzw_arr_harmonics (maximize_rzw.c:85-191) and
// the scalar arithmetic of get_localpower3d / get_derivs3d
// (characteristics.c:75-191) in the reference's double precision; every
// interpolation, volume and harmonic sum runs on the GPU, all candidates of a
// call together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "pzl_accel.h"
#include "pzl_math.h"
#include "pzl_optimize.h"
#include "pzl_plan.h"

using namespace pzl;

namespace {

constexpr int kLocAvg = 20;   // NUMLOCPOWAVG (presto.h:100)
constexpr int kDeltaAvg = 5;  // DELTAAVGBINS (presto.h:104)
constexpr double kDZ = 2.0, kDW = 20.0;  // ACCEL_DZ, ACCEL_DW (accel.h:25-29)
constexpr int kChunkJobs = 4096;         // candidate harmonics per GPU pass
constexpr int kInterpMaxKern = 1024;     // 2 khw of one interpolation (kernel limit)
constexpr int kInterpMaxSamp = 8192;     // chirp samples of its w response
constexpr int kVolCells = PZL_OPT_NUMZ * PZL_OPT_NUMW * PZL_OPT_NUMR;
constexpr double kTwoPi = 6.283185307179586477;

// Chirp samples gen_w_response takes for (z, w) at numkern points, numbetween
// per bin (responses.c:381-382); 0 for the z and r responses.
int w_response_npts(double z, double w, int numkern, int numbetween)
{
    if (fabs(w) < 1e-4) return 0;
    return pzl_next2_to_n(6 * pzl_w_resp_halfwidth(z, w) + 200 + numkern / numbetween);
}

// The interpolation points of get_localpower3d (characteristics.c:91-133).
int localpower_jobs(double r, double z, double w, int numdata, std::vector<PzlInterpJob> &jobs)
{
    const int binsperside = kLocAvg / 2;
    const int khw = pzl_w_resp_halfwidth_acc(z, w, false);
    double lo1 = r - kDeltaAvg - binsperside, hi1 = lo1 + binsperside;
    double lo2 = r + kDeltaAvg, hi2 = lo2 + binsperside;
    auto clip = [numdata](double &v) {
        if (v < 0.0) v = 0.0;
        if (v > numdata) v = (double)numdata;
    };
    clip(lo1);
    clip(lo2);
    clip(hi1);
    clip(hi2);
    const size_t n0 = jobs.size();
    for (double f = lo1; f < hi1; f += 1.0) jobs.push_back({f, z, w, khw, 0});
    for (double f = lo2; f < hi2; f += 1.0) jobs.push_back({f, z, w, khw, 0});
    return (int)(jobs.size() - n0);
}

// POWER with double argument variables (presto.h:72-73, characteristics.c:88)
double power_d(float2 a) { return (double)a.x * a.x + (double)a.y * a.y; }

double localpower_from(const float2 *ans, int count)
{
    double sum = 0.0;
    for (int i = 0; i < count; i++) sum += power_d(ans[i]);
    return sum / (double)kLocAvg;
}

// get_derivs3d's five points, r - 2h ... r + 2h (characteristics.c:164-170)
void derivs_jobs(double r, double z, double w, std::vector<PzlInterpJob> &jobs)
{
    const double h = 0.003, twoh = 0.006;
    const int khw = pzl_w_resp_halfwidth_acc(z, w, true);
    double f = r - twoh;
    for (int ii = 0; ii < 5; ii++, f += h) jobs.push_back({f, z, w, khw, 0});
}

double radian_phase(double re, double im)
{
    const double t = atan2(im, re);
    return t > 0.0 ? t : t + kTwoPi;
}

// characteristics.c:166-190
PzlRDerivs derivs_from(const float2 *ans, double localpower)
{
    const double twoh = 0.006, twohsqrd = 0.000036, PI = 3.14159265358979323846;
    double pwr[5], phs[5];
    for (int ii = 0; ii < 5; ii++) {
        pwr[ii] = power_d(ans[ii]);
        phs[ii] = radian_phase(ans[ii].x, ans[ii].y);
    }
    for (int ii = 0; ii < 4; ii++)
        if (fabs(phs[ii + 1] - phs[ii]) > PI) phs[ii + 1] -= kTwoPi;
    PzlRDerivs d;
    d.pow = pwr[2];
    d.phs = phs[2];
    d.dpow = (pwr[3] - pwr[1]) / twoh;
    d.dphs = (phs[3] - phs[1]) / twoh;
    d.d2pow = (pwr[4] - 2.0 * pwr[2] + pwr[0]) / twohsqrd;
    d.d2phs = (phs[4] - 2.0 * phs[2] + phs[0]) / twohsqrd;
    d.locpow = localpower;
    return d;
}

int check_interp_jobs(const std::vector<PzlInterpJob> &jobs, size_t from)
{
    for (size_t i = from; i < jobs.size(); i++) {
        const PzlInterpJob &J = jobs[i];
        if (2 * J.khw > kInterpMaxKern || w_response_npts(J.z, J.w, 2 * J.khw, 1) > kInterpMaxSamp)
            return fail(PZL_ERR_UNSUPPORTED, "interpolation at z %.1f, w %.1f needs a longer kernel than supported",
                        J.z, J.w);
    }
    return 0;
}

// Device buffers of one pass, grown as needed.
struct OptBuffers {
    PzlInterpJob *jobs = nullptr;
    float2 *ans = nullptr;
    size_t njobs_cap = 0;
    PzlVolJob *vjobs = nullptr;
    double2 *tw = nullptr;
    int *err = nullptr;
    float *powv = nullptr, *tail = nullptr, *locpow = nullptr;
    int *job0 = nullptr, *nharm = nullptr, *best = nullptr;
    float *bestv = nullptr;
    size_t nv_cap = 0, nc_cap = 0;
    ~OptBuffers()
    {
        (void)hipFree(jobs);
        (void)hipFree(ans);
        (void)hipFree(vjobs);
        (void)hipFree(tw);
        (void)hipFree(err);
        (void)hipFree(powv);
        (void)hipFree(tail);
        (void)hipFree(locpow);
        (void)hipFree(job0);
        (void)hipFree(nharm);
        (void)hipFree(best);
        (void)hipFree(bestv);
    }
    int interp(size_t n)
    {
        if (n <= njobs_cap) return 0;
        (void)hipFree(jobs);
        (void)hipFree(ans);
        jobs = nullptr;
        ans = nullptr;
        njobs_cap = 0;
        PZL_HIPCHK(hipMalloc(&jobs, sizeof(PzlInterpJob) * n));
        PZL_HIPCHK(hipMalloc(&ans, sizeof(float2) * n));
        njobs_cap = n;
        return 0;
    }
    int volume(size_t nv, size_t nc, hipStream_t st)
    {
        if (!tw) {
            PZL_HIPCHK(hipMalloc(&err, sizeof(int)));
            PZL_HIPCHK(hipMalloc(&tw, sizeof(double2) * PZL_OPT_TWN));
            PZL_HIPCHK(pzl_launch_opt_twiddles(tw, st));
        }
        if (nv > nv_cap) {
            (void)hipFree(vjobs);
            (void)hipFree(powv);
            (void)hipFree(tail);
            (void)hipFree(locpow);
            vjobs = nullptr;
            powv = nullptr;
            tail = nullptr;
            locpow = nullptr;
            nv_cap = 0;
            PZL_HIPCHK(hipMalloc(&vjobs, sizeof(PzlVolJob) * nv));
            PZL_HIPCHK(hipMalloc(&powv, sizeof(float) * kVolCells * nv));
            PZL_HIPCHK(hipMalloc(&tail, sizeof(float) * PZL_OPT_NUMZ * PZL_OPT_NUMW * PZL_OPT_TAIL * nv));
            PZL_HIPCHK(hipMalloc(&locpow, sizeof(float) * nv));
            nv_cap = nv;
        }
        if (nc > nc_cap) {
            (void)hipFree(job0);
            (void)hipFree(nharm);
            (void)hipFree(best);
            (void)hipFree(bestv);
            job0 = nharm = best = nullptr;
            bestv = nullptr;
            nc_cap = 0;
            PZL_HIPCHK(hipMalloc(&job0, sizeof(int) * nc));
            PZL_HIPCHK(hipMalloc(&nharm, sizeof(int) * nc));
            PZL_HIPCHK(hipMalloc(&best, sizeof(int) * nc));
            PZL_HIPCHK(hipMalloc(&bestv, sizeof(float) * nc));
            nc_cap = nc;
        }
        return 0;
    }
};

// Run the interpolation jobs; answers into 'ans'.
int run_interp(OptBuffers &B, const float2 *d_data, int numdata, const std::vector<PzlInterpJob> &jobs,
               std::vector<float2> &ans, hipStream_t st)
{
    ans.assign(jobs.size(), float2{0.f, 0.f});
    if (jobs.empty()) return 0;
    int rc = B.interp(jobs.size());
    if (rc) return rc;
    PZL_HIPCHK(hipMemcpyAsync(B.jobs, jobs.data(), sizeof(PzlInterpJob) * jobs.size(), hipMemcpyHostToDevice, st));
    PZL_HIPCHK(pzl_launch_opt_interp(d_data, numdata, B.jobs, (int)jobs.size(), B.ans, st));
    PZL_HIPCHK(hipMemcpyAsync(ans.data(), B.ans, sizeof(float2) * jobs.size(), hipMemcpyDeviceToHost, st));
    PZL_HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// One pass over candidates [c0, c1).  single: every candidate is one harmonic
// searched as max_rzw_arr does (maximize_rzw.c:3-50): a 9 x 9 grid over z +-
// ACCEL_DZ, w +- ACCEL_DW, the raw power's maximum, the input point's local
// power in the derivatives.
int optimize_chunk(OptBuffers &B, const float2 *d_data, int numdata, PzlOptCand *cands, long long c0, long long c1,
                   bool single, hipStream_t st)
{
    const int nz = single ? 2 * PZL_OPT_INTERPFAC + 1 : PZL_OPT_NUMZ;  // maximize_rzw.c:21
    const int nw = single ? 2 * PZL_OPT_INTERPFAC + 1 : PZL_OPT_NUMW;
    const double zwid = single ? 1.0 : PZL_OPT_NUMWID;  // half-range in units of ACCEL_DZ, ACCEL_DW
    const long long nc = c1 - c0;
    std::vector<PzlVolJob> vj;
    std::vector<int> job0((size_t)nc), nharm((size_t)nc);
    std::vector<long> hhlobin((size_t)nc);
    std::vector<PzlInterpJob> lj;
    std::vector<int> lj0, ljn;
    int ldsn = 0;  // LDS of a volume row: its numkern and w-response chirp samples (double2)
    for (long long c = 0; c < nc; c++) {
        const PzlOptCand &C = cands[c0 + c];
        const int H = C.numharm;
        job0[c] = (int)vj.size();
        nharm[c] = H;
        for (int hind = 0; hind < H; hind++) {
            const int n = H - hind;  // harmonic number, highest first (maximize_rzw.c:117)
            const double rh = C.r * n, zh = C.z * n, wh = C.w * n, hfrac = n / (double)H;
            lj0.push_back((int)lj.size());
            ljn.push_back(localpower_jobs(rh, zh, wh, numdata, lj));
            const int khw_n = pzl_w_resp_halfwidth_acc(fabs(zh) + 2, fabs(wh) + 20, false);
            double rh_int;
            modf(rh, &rh_int);
            const long lobin = (long)rh_int - 1;
            if (hind == 0) hhlobin[c] = lobin;
            PzlVolJob J{};
            J.lobin = (int)lobin;
            J.fftlen = pzl_next2_to_n(PZL_OPT_NB * (2 * khw_n + 10));
            // corr_rzw_vol (rzwinterp.c:34-60) on the z and w ranges of :130-135
            // (max_rzw_arr: zin -+ ACCEL_DZ, win -+ ACCEL_DW, maximize_rzw.c:27-30)
            const double zlo = zh - zwid * hfrac * kDZ, zhi = zh + zwid * hfrac * kDZ;
            const double wlo = wh - zwid * hfrac * kDW, whi = wh + zwid * hfrac * kDW;
            J.zlo = zlo;
            J.wlo = wlo;
            J.dz = (zhi - zlo) / (nz - 1);
            J.dw = (whi - wlo) / (nw - 1);
            const double maxz = (fabs(zlo) < fabs(zhi)) ? zhi : zlo;
            const double maxw = (fabs(wlo) < fabs(whi)) ? whi : wlo;
            J.khw = pzl_w_resp_halfwidth_acc(maxz, maxw, false);
            int ngood = J.fftlen - 2 * J.khw * PZL_OPT_NB;
            while (ngood > 0 && ngood % PZL_OPT_NB) ngood--;
            J.numgoodbins = ngood;
            const double dr = 1.0 / PZL_OPT_NB;
            for (int kk = 0; kk < PZL_OPT_NUMR; kk++) {
                const long ri = (long)round(((hhlobin[c] + kk * dr) * hfrac - lobin) * PZL_OPT_NB);
                // a step below the row start reads the end of the preceding
                // row (the reference's contiguous volume); at most 6 back
                if (ri < -PZL_OPT_TAIL || ri >= ngood || ngood < PZL_OPT_TAIL)
                    return fail(PZL_ERR_UNSUPPORTED, "volume of harmonic %d of the candidate at r %.3f misses r step %d",
                                n, C.r, kk);
                J.rind[kk] = (int)ri;
            }
            const int numkern = 2 * PZL_OPT_NB * J.khw;
            if (J.fftlen > PZL_OPT_MAX_FFTLEN || numkern > J.fftlen || lobin > INT_MAX)
                return fail(PZL_ERR_UNSUPPORTED, "candidate at r %.3f z %.2f w %.2f: harmonic %d needs an FFT of %d",
                            C.r, C.z, C.w, n, J.fftlen);
            ldsn = std::max(ldsn, numkern);
            for (int i = 0; i < nw; i++)
                for (int j = 0; j < nz; j++) {
                    const int npts = w_response_npts(zlo + j * J.dz, wlo + i * J.dw, numkern, PZL_OPT_NB);
                    if (npts > PZL_OPT_RESP_FFT_MAX)
                        return fail(PZL_ERR_UNSUPPORTED,
                                    "candidate at r %.3f z %.2f w %.2f: harmonic %d needs a w response of %d samples",
                                    C.r, C.z, C.w, n, npts);
                    ldsn = std::max(ldsn, npts);
                }
            vj.push_back(J);
        }
    }
    int rc = check_interp_jobs(lj, 0);
    if (rc) return rc;
    const size_t nv = vj.size();
    rc = B.volume(nv, (size_t)nc, st);
    if (rc) return rc;
    // volumes (independent of the local powers, which only scale the sum)
    PZL_HIPCHK(hipMemcpyAsync(B.vjobs, vj.data(), sizeof(PzlVolJob) * nv, hipMemcpyHostToDevice, st));
    PZL_HIPCHK(hipMemsetAsync(B.err, 0, sizeof(int), st));
    PZL_HIPCHK(pzl_launch_opt_volume(d_data, numdata, B.vjobs, (int)nv, ldsn, nz, nw, B.tw, B.powv, B.tail, B.err,
                                     st));
    std::vector<float2> ans;
    rc = run_interp(B, d_data, numdata, lj, ans, st);
    if (rc) return rc;
    // the sum's local powers (1: max_rzw_arr compares raw powers)
    std::vector<float> locpow(nv), inpow(nv);
    for (size_t j = 0; j < nv; j++) {
        inpow[j] = (float)localpower_from(ans.data() + lj0[j], ljn[j]);
        locpow[j] = single ? 1.0f : inpow[j];
    }
    std::vector<int> best((size_t)nc);
    std::vector<float> bestv((size_t)nc);
    PZL_HIPCHK(hipMemcpyAsync(B.locpow, locpow.data(), sizeof(float) * nv, hipMemcpyHostToDevice, st));
    PZL_HIPCHK(hipMemcpyAsync(B.job0, job0.data(), sizeof(int) * nc, hipMemcpyHostToDevice, st));
    PZL_HIPCHK(hipMemcpyAsync(B.nharm, nharm.data(), sizeof(int) * nc, hipMemcpyHostToDevice, st));
    PZL_HIPCHK(pzl_launch_opt_sum(B.powv, B.tail, B.vjobs, B.locpow, B.job0, B.nharm, (int)nc, nz * nw, B.best,
                                  B.bestv, st));
    PZL_HIPCHK(hipMemcpyAsync(best.data(), B.best, sizeof(int) * nc, hipMemcpyDeviceToHost, st));
    PZL_HIPCHK(hipMemcpyAsync(bestv.data(), B.bestv, sizeof(float) * nc, hipMemcpyDeviceToHost, st));
    int verr = 0;
    PZL_HIPCHK(hipMemcpyAsync(&verr, B.err, sizeof(int), hipMemcpyDeviceToHost, st));
    PZL_HIPCHK(hipStreamSynchronize(st));
    if (verr) return fail(PZL_ERR_UNSUPPORTED, "a volume row's response exceeded the launch's LDS");

    if (single) {
        // max_rzw_arr's maximum (maximize_rzw.c:46-48) and the derivatives
        // there with the input point's local power (:4, :52)
        std::vector<PzlInterpJob> dj;
        for (long long c = 0; c < nc; c++) {
            PzlOptCand &C = cands[c0 + c];
            const int idx = best[c];
            const int wind = idx / (nz * PZL_OPT_NUMR), zind = (idx / PZL_OPT_NUMR) % nz, rind = idx % PZL_OPT_NUMR;
            C.ropt = hhlobin[c] + rind * 0.5 / (double)PZL_OPT_INTERPFAC;
            C.zopt = C.z - kDZ + zind * kDZ / (double)PZL_OPT_INTERPFAC;
            C.wopt = C.w - kDW + wind * kDW / (double)PZL_OPT_INTERPFAC;
            derivs_jobs(C.ropt, C.zopt, C.wopt, dj);
        }
        rc = check_interp_jobs(dj, 0);
        if (rc) return rc;
        rc = run_interp(B, d_data, numdata, dj, ans, st);
        if (rc) return rc;
        const size_t per = dj.size() / (size_t)(nc ? nc : 1);
        for (long long c = 0; c < nc; c++) {
            PzlOptCand &C = cands[c0 + c];
            C.derivs[0] = derivs_from(ans.data() + per * c, inpow[c]);
            C.pows[0] = bestv[c];
        }
        return 0;
    }
    // the maximum (maximize_rzw.c:177-184), then every harmonic's local power
    // and derivatives there (:186-193)
    std::vector<PzlInterpJob> dj;
    std::vector<int> dl0, dln, dd0;
    for (long long c = 0; c < nc; c++) {
        PzlOptCand &C = cands[c0 + c];
        const int idx = best[c];
        const int wind = idx / (PZL_OPT_NUMZ * PZL_OPT_NUMR), zind = (idx / PZL_OPT_NUMR) % PZL_OPT_NUMZ,
                  rind = idx % PZL_OPT_NUMR;
        const double nh = C.numharm, dr = 1.0 / PZL_OPT_NB;
        C.ropt = (hhlobin[c] + rind * dr) / nh;
        C.zopt = ((C.z * nh) - PZL_OPT_NUMWID * kDZ + zind * kDZ / (double)PZL_OPT_INTERPFAC) / nh;
        C.wopt = ((C.w * nh) - PZL_OPT_NUMWID * kDW + wind * kDW / (double)PZL_OPT_INTERPFAC) / nh;
        for (int ii = 0; ii < C.numharm; ii++) {
            const int hh = ii + 1;
            const double rr = C.ropt * hh, zz = C.zopt * hh, ww = C.wopt * hh;
            dl0.push_back((int)dj.size());
            dln.push_back(localpower_jobs(rr, zz, ww, numdata, dj));
            dd0.push_back((int)dj.size());
            derivs_jobs(rr, zz, ww, dj);
        }
    }
    rc = check_interp_jobs(dj, 0);
    if (rc) return rc;
    rc = run_interp(B, d_data, numdata, dj, ans, st);
    if (rc) return rc;
    size_t k = 0;
    for (long long c = 0; c < nc; c++) {
        PzlOptCand &C = cands[c0 + c];
        for (int ii = 0; ii < C.numharm; ii++, k++) {
            const float lp = (float)localpower_from(ans.data() + dl0[k], dln[k]);
            C.derivs[ii] = derivs_from(ans.data() + dd0[k], lp);
            C.pows[ii] = C.derivs[ii].pow;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// w = 0 searches: max_rz_arr_harmonics (maximize_rz.c:141-220) with the
// downhill simplex of amoeba.c:21-106, one resumable state machine per
// candidate; every round evaluates the pending points of all candidates in
// one batch of interpolations.
// ---------------------------------------------------------------------------

constexpr double kZScale = 4.0;  // maximize_rz.c:2

struct Simplex {
    double p[3][2], y[3], psum[2];
    double ptry[2], ysave = 0.0, ftol = 0.0;
    int nfunk = 0, ihi = 0, ilo = 0, inhi = 0;
    enum { kInit1, kRun1, kInit2, kRun2, kDone } phase = kInit1;
    enum { kTop, kReflect, kExpand, kContract, kShrink } step = kTop;
    int npend = 0;
    double pend[3][2];
    int pend_idx[3];
    int khw = 0, nh = 1;
    std::vector<float> locpow;

    void request(const double *pt, int idx)
    {
        pend[npend][0] = pt[0];
        pend[npend][1] = pt[1];
        pend_idx[npend++] = idx;
    }
    // amotry's trial point (amoeba.c:95-98)
    void try_point(double fac)
    {
        const double fac1 = 0.5 * (1.0 - fac), fac2 = fac1 - fac;
        ptry[0] = psum[0] * fac1 - p[ihi][0] * fac2;
        ptry[1] = psum[1] * fac1 - p[ihi][1] * fac2;
        request(ptry, -1);
    }
    // amotry's update (amoeba.c:99-105)
    void accept(double ytry)
    {
        if (ytry < y[ihi]) {
            y[ihi] = ytry;
            psum[0] += ptry[0] - p[ihi][0];
            p[ihi][0] = ptry[0];
            psum[1] += ptry[1] - p[ihi][1];
            p[ihi][1] = ptry[1];
        }
    }
    void start_amoeba(double tol)
    {
        ftol = tol;
        nfunk = 0;
        psum[0] = p[0][0] + p[1][0] + p[2][0];
        psum[1] = p[0][1] + p[1][1] + p[2][1];
        top();
    }
    // the head of amoeba's loop (amoeba.c:30-62)
    void top()
    {
        ilo = 0;
        if (y[0] > y[1]) {
            inhi = 1;
            ihi = 0;
        } else {
            inhi = 0;
            ihi = 1;
        }
        for (int ii = 0; ii <= 2; ii++) {
            if (y[ii] <= y[ilo]) ilo = ii;
            if (y[ii] > y[ihi]) {
                inhi = ihi;
                ihi = ii;
            } else if (y[ii] > y[inhi] && ii != ihi) {
                inhi = ii;
            }
        }
        const double rtol = 2.0 * fabs(y[ihi] - y[ilo]) / (fabs(y[ihi]) + fabs(y[ilo]) + 1.0e-15);
        if (rtol < ftol) {
            std::swap(y[0], y[ilo]);
            std::swap(p[0][0], p[ilo][0]);
            std::swap(p[0][1], p[ilo][1]);
            end_amoeba();
            return;
        }
        if (nfunk >= 5000) {
            end_amoeba();
            return;
        }
        nfunk += 2;
        step = kReflect;
        try_point(-1.0);
    }
    void end_amoeba()
    {
        if (phase == kRun1) {
            // restart at the minimum at HIGHACC (maximize_rz.c:183-199)
            khw = pzl_z_resp_halfwidth_acc(fabs(p[0][1] * nh) + 4.0, true);
            p[1][0] = p[0][0] + 0.01;
            p[1][1] = p[0][1];
            p[2][0] = p[0][0];
            p[2][1] = p[0][1] + 0.01;
            phase = kInit2;
            for (int ii = 0; ii < 3; ii++) request(p[ii], ii);
        } else {
            phase = kDone;
        }
    }
    // the answers to the pending requests, in request order
    void deliver(const double *vals)
    {
        const int n = npend;
        npend = 0;
        if (phase == kInit1 || phase == kInit2) {
            for (int k = 0; k < n; k++) y[pend_idx[k]] = vals[k];
            phase = phase == kInit1 ? kRun1 : kRun2;
            start_amoeba(phase == kRun1 ? 1.0e-7 : 1.0e-10);
            return;
        }
        switch (step) {
        case kReflect: {
            const double ytry = vals[0];
            accept(ytry);
            if (ytry <= y[ilo]) {
                step = kExpand;
                try_point(2.0);
            } else if (ytry >= y[inhi]) {
                ysave = y[ihi];
                step = kContract;
                try_point(0.5);
            } else {
                --nfunk;
                top();
            }
            break;
        }
        case kExpand:
            accept(vals[0]);
            top();
            break;
        case kContract: {
            const double ytry = vals[0];
            accept(ytry);
            if (ytry >= ysave) {
                // shrink towards the best point (amoeba.c:72-83)
                step = kShrink;
                for (int ii = 0; ii <= 2; ii++)
                    if (ii != ilo) {
                        p[ii][0] = 0.5 * (p[ii][0] + p[ilo][0]);
                        p[ii][1] = 0.5 * (p[ii][1] + p[ilo][1]);
                        request(p[ii], ii);
                    }
            } else {
                top();
            }
            break;
        }
        case kShrink:
            for (int k = 0; k < n; k++) y[pend_idx[k]] = vals[k];
            nfunk += 2;
            psum[0] = p[0][0] + p[1][0] + p[2][0];
            psum[1] = p[0][1] + p[1][1] + p[2][1];
            top();
            break;
        default:
            break;
        }
    }
};

// power_call_rz_harmonics (maximize_rz.c:99-113) as interpolation jobs
void rz_point_jobs(const Simplex &S, const double *pt, std::vector<PzlInterpJob> &jobs)
{
    for (int ii = 0; ii < S.nh; ii++) {
        const int n = ii + 1;
        jobs.push_back({pt[0] * n, pt[1] * kZScale * n, 0.0, S.khw, 0});
    }
}

// single: one harmonic each, on its raw power as max_rz_arr does
// (maximize_rz.c:25-85; power_call_rz, :10-23, divides by nothing).
int optimize_rz(OptBuffers &B, const float2 *d_data, int numdata, PzlOptCand *cands, long long nc, bool single,
                hipStream_t st)
{
    std::vector<Simplex> S((size_t)nc);
    std::vector<PzlInterpJob> jobs;
    std::vector<float2> ans;
    std::vector<int> l0, ln;
    // local powers at the start (maximize_rz.c:157-161)
    for (long long c = 0; c < nc && !single; c++)
        for (int ii = 0; ii < cands[c].numharm; ii++) {
            const int n = ii + 1;
            l0.push_back((int)jobs.size());
            ln.push_back(localpower_jobs(cands[c].r * n, cands[c].z * n, 0.0, numdata, jobs));
        }
    int rc = check_interp_jobs(jobs, 0);
    if (rc) return rc;
    rc = run_interp(B, d_data, numdata, jobs, ans, st);
    if (rc) return rc;
    size_t k = 0;
    for (long long c = 0; c < nc; c++) {
        Simplex &X = S[c];
        const PzlOptCand &C = cands[c];
        X.nh = C.numharm;
        for (int ii = 0; ii < C.numharm; ii++, k++)
            X.locpow.push_back(single ? 1.0f : (float)localpower_from(ans.data() + l0[k], ln[k]));
        X.khw = pzl_z_resp_halfwidth_acc(fabs(C.z * C.numharm) + 4.0, false);
        const double step = 0.4;  // maximize_rz.c:165-172
        const double init[3][2] = {{C.r - step, C.z / kZScale - step},
                                   {C.r - step, C.z / kZScale + step},
                                   {C.r + step, C.z / kZScale}};
        for (int ii = 0; ii < 3; ii++) {
            X.p[ii][0] = init[ii][0];
            X.p[ii][1] = init[ii][1];
            X.request(X.p[ii], ii);
        }
    }
    // rounds: every pending point of every candidate in one batch
    for (;;) {
        jobs.clear();
        std::vector<std::pair<long long, int>> owner;  // (candidate, pending slot) of each point
        for (long long c = 0; c < nc; c++)
            for (int q = 0; q < S[c].npend; q++) {
                owner.push_back({c, q});
                rz_point_jobs(S[c], S[c].pend[q], jobs);
            }
        if (owner.empty()) break;
        rc = check_interp_jobs(jobs, 0);
        if (rc) return rc;
        rc = run_interp(B, d_data, numdata, jobs, ans, st);
        if (rc) return rc;
        std::vector<double> vals(3 * (size_t)nc, 0.0);
        size_t j = 0;
        for (auto &o : owner) {
            const Simplex &X = S[o.first];
            double total = 0.0;
            for (int ii = 0; ii < X.nh; ii++, j++) total += power_d(ans[j]) / X.locpow[ii];
            vals[3 * o.first + o.second] = -total;
        }
        for (long long c = 0; c < nc; c++)
            if (S[c].npend) S[c].deliver(vals.data() + 3 * c);
    }
    // the peak (maximize_rz.c:207-219): local power, power and derivatives per harmonic
    std::vector<PzlInterpJob> dj;
    std::vector<int> dl0, dln, dp0, dd0;
    for (long long c = 0; c < nc; c++) {
        PzlOptCand &C = cands[c];
        const Simplex &X = S[c];
        C.ropt = X.p[0][0];
        C.zopt = X.p[0][1] * kZScale;
        C.wopt = 0.0;
        for (int ii = 0; ii < C.numharm; ii++) {
            const int n = ii + 1;
            dl0.push_back((int)dj.size());
            dln.push_back(localpower_jobs(C.ropt * n, C.zopt * n, 0.0, numdata, dj));
            dp0.push_back((int)dj.size());
            dj.push_back({C.ropt * n, C.zopt / kZScale * n * kZScale, 0.0, X.khw, 0});
            dd0.push_back((int)dj.size());
            derivs_jobs(C.ropt * n, C.zopt * n, 0.0, dj);
        }
    }
    rc = check_interp_jobs(dj, 0);
    if (rc) return rc;
    rc = run_interp(B, d_data, numdata, dj, ans, st);
    if (rc) return rc;
    k = 0;
    for (long long c = 0; c < nc; c++) {
        PzlOptCand &C = cands[c];
        for (int ii = 0; ii < C.numharm; ii++, k++) {
            const float lp = (float)localpower_from(ans.data() + dl0[k], dln[k]);
            C.pows[ii] = power_d(ans[dp0[k]]);
            C.derivs[ii] = derivs_from(ans.data() + dd0[k], lp);
        }
    }
    return 0;
}

int optimize_on(const float2 *d_data, long long numbins, PzlOptCand *cands, long long n, bool single,
                hipStream_t st)
{
    if (numbins <= 0 || numbins > INT_MAX) return fail(PZL_ERR_ARG, "spectrum of %lld bins", numbins);
    for (long long c = 0; c < n; c++) {
        const PzlOptCand &C = cands[c];
        if (C.numharm < 1 || C.numharm > PZL_MAX_NUMHARM || !(C.r >= 0.0) || !std::isfinite(C.z) ||
            !std::isfinite(C.w) || (single && C.numharm != 1))
            return fail(PZL_ERR_ARG, "candidate %lld: r %.3f, numharm %d", c, C.r, C.numharm);
    }
    OptBuffers B;
    // w = 0 candidates (the simplex) and the rest (the rzw volume), each as one batch
    std::vector<PzlOptCand> rz;
    std::vector<long long> rz_at;
    for (long long c = 0; c < n; c++)
        if (cands[c].fdot_only) {
            rz.push_back(cands[c]);
            rz_at.push_back(c);
        }
    if (!rz.empty()) {
        const int rc = optimize_rz(B, d_data, (int)numbins, rz.data(), (long long)rz.size(), single, st);
        if (rc) return rc;
        for (size_t k = 0; k < rz.size(); k++) cands[rz_at[k]] = rz[k];
    }
    std::vector<PzlOptCand> vw;
    std::vector<long long> vw_at;
    for (long long c = 0; c < n; c++)
        if (!cands[c].fdot_only) {
            vw.push_back(cands[c]);
            vw_at.push_back(c);
        }
    const long long nv = (long long)vw.size();
    long long c0 = 0;
    while (c0 < nv) {
        long long c1 = c0, jobs = 0;
        while (c1 < nv && (c1 == c0 || jobs + vw[c1].numharm <= kChunkJobs)) jobs += vw[c1++].numharm;
        const int rc = optimize_chunk(B, d_data, (int)numbins, vw.data(), c0, c1, single, st);
        if (rc) return rc;
        c0 = c1;
    }
    for (long long k = 0; k < nv; k++) cands[vw_at[k]] = vw[k];
    return 0;
}

}  // namespace

extern "C" {

int pzl_optimize(void *plan, PzlOptCand *cands, long long n)
{
    Plan *p = (Plan *)plan;
    if (!p || (n > 0 && !cands)) return fail(PZL_ERR_ARG, "null argument");
    if (!p->d_fft) return fail(PZL_ERR_NOSPECTRUM, "no spectrum set");
    PZL_HIPCHK(hipSetDevice(p->device));
    PZL_HIPCHK(hipStreamSynchronize(p->st));
    return optimize_on(p->d_fft, p->fft_bins, cands, n, false, p->st);
}

namespace {
int on_host_spectrum(int device, const float *host_fft, long long numbins, PzlOptCand *cands, long long n,
                     bool single)
{
    if (!host_fft || numbins <= 0 || (n > 0 && !cands)) return fail(PZL_ERR_ARG, "bad arguments");
    PZL_HIPCHK(hipSetDevice(device));
    float2 *d = nullptr;
    hipStream_t st = nullptr;
    PZL_HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipError_t e = hipMalloc(&d, sizeof(float2) * numbins);
    if (e == hipSuccess) e = hipMemcpyAsync(d, host_fft, sizeof(float2) * numbins, hipMemcpyHostToDevice, st);
    int rc = e == hipSuccess ? optimize_on(d, numbins, cands, n, single, st)
                             : fail(PZL_ERR_HIP, "spectrum upload: %s", hipGetErrorString(e));
    (void)hipStreamSynchronize(st);
    (void)hipFree(d);
    (void)hipStreamDestroy(st);
    return rc;
}
}  // namespace

int pzl_optimize_spectrum(int device, const float *host_fft, long long numbins, PzlOptCand *cands, long long n)
{
    return on_host_spectrum(device, host_fft, numbins, cands, n, false);
}

int pzl_maximize(void *plan, PzlOptCand *cands, long long n)
{
    Plan *p = (Plan *)plan;
    if (!p || (n > 0 && !cands)) return fail(PZL_ERR_ARG, "null argument");
    if (!p->d_fft) return fail(PZL_ERR_NOSPECTRUM, "no spectrum set");
    PZL_HIPCHK(hipSetDevice(p->device));
    PZL_HIPCHK(hipStreamSynchronize(p->st));
    return optimize_on(p->d_fft, p->fft_bins, cands, n, true, p->st);
}

int pzl_maximize_spectrum(int device, const float *host_fft, long long numbins, PzlOptCand *cands, long long n)
{
    return on_host_spectrum(device, host_fft, numbins, cands, n, true);
}

}  // extern "C"
