// pzl_geometry.cpp -- search geometry and thresholds.
//
//   create_accelobs         accel_utils.c:1636-2037 (zmax/wmax rounding, r range,
//                           numindep, powcut, maxkernlen, fftlen, corr_uselen)
//   calc_fftlen             accel_utils.c:137-149
//   init_subharminfo        accel_utils.c:199-234 (per-harmonic zmax/wmax/grid)
//   rstep enumeration       accelsearch.c:183-188 (startr += rstep in double)
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "pzl_math.h"
#include "pzl_plan.h"
#include "pzl_stats.h"

namespace pzl {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

namespace {

int twon_to_index(int n)
{
    int x = 0;
    while (n > 1) {
        n >>= 1;
        x++;
    }
    return x;
}

int calc_fftlen(int numharm, int harmnum, int max_zfull, int max_wfull, int corr_uselen)
{
    const double hf = (double)harmnum / (double)numharm;
    const int bins_needed = (int)ceil(corr_uselen * hf) + 2;
    const int end_effects = 2 * PZL_NUMBETWEEN *
                            pzl_w_resp_halfwidth(pzl_calc_required_z(hf, max_zfull), pzl_calc_required_w(hf, max_wfull));
    return pzl_next_good_fftlen(bins_needed + end_effects);
}

}  // namespace

int compute_geometry(const PzlConfig &c, PzlGeometry &g, std::vector<PzlHarm> *harms_out,
                     std::vector<double> *startr_out)
{
    memset(&g, 0, sizeof g);
    int zmax = c.zmax, wmax = c.wmax;
    if (zmax < 0 || wmax < 0) return fail(PZL_ERR_ARG, "zmax and wmax must be >= 0");
    if (zmax % PZL_DZ) zmax = (zmax / PZL_DZ + 1) * PZL_DZ;
    const int nhm = c.numharm;
    if (nhm != 1 && nhm != 2 && nhm != 4 && nhm != 8 && nhm != 16 && nhm != 32)
        return fail(PZL_ERR_ARG, "'numharm' = %d must be a power-of-two (1..32)", c.numharm);
    if (c.numbins < 2 || c.N <= 0 || c.dt <= 0) return fail(PZL_ERR_ARG, "bad N/dt/numbins");
    g.N = (long long)c.N;
    g.numbins = c.numbins;
    g.lobin = c.lobin;
    g.numharmstages = twon_to_index(c.numharm) + 1;
    g.numz = (zmax / PZL_DZ) * 2 + 1;
    if (c.wmaxP) {
        if (wmax % PZL_DW) wmax = (wmax / PZL_DW + 1) * PZL_DW;
        g.whi = wmax;
        g.wlo = -wmax;
        g.numw = (wmax / PZL_DW) * 2 + 1;
        if (wmax == 0) g.numw = 0;
    } else {
        g.whi = g.wlo = 0.0;
        g.numw = 0;
        wmax = 0;
    }
    g.dt = c.dt;
    g.T = c.dt * c.N;
    g.rlo = c.floP ? floor(c.flo * g.T) : (c.rloP ? c.rlo : 1.0);
    if (g.rlo < g.lobin) g.rlo = (double)g.lobin;
    if (g.rlo > g.numbins - 1) return fail(PZL_ERR_ARG, "low frequency to search is above the highest bin");
    g.highestbin = g.numbins - 1;
    // The reference leaves rhi unset unless -fhi/-rhi is given; the default
    // command line always sets -fhi (accelsearch_cmd.c:59-62).
    g.rhi = (double)g.highestbin;
    if (c.fhiP || c.rhiP) {
        g.highestbin = c.fhiP ? (long long)ceil(c.fhi * g.T) : (long long)c.rhi;
        if (g.highestbin > g.numbins - 1) g.highestbin = g.numbins - 1;
        g.rhi = (double)g.highestbin;
        if (g.highestbin < g.rlo) return fail(PZL_ERR_ARG, "high frequency to search is below the low one");
    }
    g.zhi = zmax;
    g.zlo = -zmax;
    g.sigma = c.sigma;
    for (int ii = 0; ii < g.numharmstages; ii++) {
        const double dz = PZL_DZ, dw = PZL_DW;
        if (g.numz == 1 && g.numw == 0)
            g.numindep[ii] = (long long)((g.rhi - g.rlo) / (1 << ii));
        else if (g.numz > 1 && g.numw == 0)
            g.numindep[ii] = (long long)((g.rhi - g.rlo) * (g.numz + 1) * (dz / 6.95) / (1 << ii));
        else
            g.numindep[ii] =
                (long long)((g.rhi - g.rlo) * (g.numz + 1) * (dz / 6.95) * (g.numw + 1) * (dw / 44.2) / (1 << ii));
        g.powcut[ii] = (float)pzl_power_for_sigma(g.sigma, 1 << ii, (double)g.numindep[ii]);
    }
    g.maxkernlen = 2 * PZL_NUMBETWEEN * pzl_w_resp_halfwidth(g.zhi, g.whi);
    g.fftlen = pzl_fftlen_from_kernwidth(g.maxkernlen);
    if (g.fftlen < 2048) g.fftlen = 2048;
    g.corr_uselen = g.fftlen - g.maxkernlen;
    if (g.corr_uselen % PZL_RDR) g.corr_uselen = g.corr_uselen / PZL_RDR * PZL_RDR;
    if (g.corr_uselen <= 0 || g.corr_uselen > 65535) return fail(PZL_ERR_UNSUPPORTED, "corr_uselen out of range");

    std::vector<double> startr;
    const int rstep = (int)(g.corr_uselen * PZL_DR);
    for (double s = g.rlo; s + rstep < g.highestbin; s += rstep) startr.push_back(s);
    g.nrsteps = (long long)startr.size();

    // harmonics: fundamental, then stage order (accelsearch.c:199-211)
    std::vector<PzlHarm> harms;
    long long bank_off = 0, plane_off = 0, seg_off = 0;
    for (int stage = 0; stage < g.numharmstages; stage++) {
        const int numharm = 1 << stage;
        for (int harm = 1; harm < (stage ? numharm : 2); harm += 2) {
            PzlHarm H{};
            H.numharm = numharm;
            H.harmnum = harm;
            H.stage = stage;
            H.hf = (double)harm / (double)numharm;
            H.zmax = pzl_calc_required_z(H.hf, (int)g.zhi);
            H.wmax = pzl_calc_required_w(H.hf, (int)g.whi);
            H.numz = (H.zmax / PZL_DZ) * 2 + 1;
            H.numw = (H.wmax / PZL_DW) * 2 + 1;
            H.nrows = H.numz * H.numw;
            H.zlo = pzl_calc_required_z(H.hf, g.zlo);
            H.wlo = pzl_calc_required_w(H.hf, g.wlo);
            H.fftlen = (stage == 0) ? g.fftlen : calc_fftlen(numharm, harm, (int)g.zhi, (int)g.whi, g.corr_uselen);
            H.halfwidth = pzl_w_resp_halfwidth((double)-H.zmax, (double)-H.wmax);
            int maxrs = 0;
            for (double s : startr) {
                const PzlHarmR hr = pzl_harm_r(H.hf, stage == 0, s, s + rstep - PZL_DR, g.corr_uselen);
                maxrs = std::max(maxrs, hr.numrs);
            }
            if (startr.empty()) maxrs = g.corr_uselen;
            H.max_numrs = maxrs;
            H.rs_stride = (maxrs + 31) / 32 * 32;
            H.bank_off = bank_off;
            H.plane_off = plane_off;
            H.seg_off = seg_off;
            bank_off += (long long)H.nrows * H.fftlen;
            plane_off += (long long)H.nrows * H.rs_stride;
            seg_off += H.fftlen / 2;
            harms.push_back(H);
        }
    }
    g.numharm_kernels = 0;
    for (auto &H : harms) g.numharm_kernels += H.nrows;
    g.bank_bytes = bank_off * 8;
    g.plane_bytes_per_rstep = plane_off * 4;
    if (harms_out) *harms_out = harms;
    if (startr_out) *startr_out = startr;
    return 0;
}

}  // namespace pzl
