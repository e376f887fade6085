// pzl_engine.cpp -- host side of the MI355X jerk search (C ABI in
// include/pzl_accel.h).
//
// Replaces the reference's GPU driver and host helpers:
//   geometry and thresholds     create_accelobs, accel_utils.c:1636-2037
//   harmonic shapes / bank      create_subharminfos :236-282, init_kernel :151-168,
//                               fkern_host_to_dev accel_utils_cu.c:378-429
//   rstep enumeration, batches  accelsearch_GPU, accelsearch_lib.c:473-901
//   hit drain + insertion       accelsearch_lib.c:784-866, insert_to_cands
//                               accel_utils_cu.c:857-929, insert_new_accelcand
//                               accel_utils.c:382-480
// Everything per rstep runs on the GPU: segment preparation, correlation,
// harmonic summing and compaction.  The host only orders the (rare) hits,
// evaluates their significance in double and maintains the candidate list.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "pzl_accel.h"
#include "pzl_kernels.h"
#include "pzl_math.h"
#include "pzl_stats.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
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

#define HIPCHK(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return fail(PZL_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                         \
    } while (0)

const long long kDefaultMaxHits = 1LL << 22;

struct SizeClass {
    int fftlen;
    std::vector<int> harms;
    int *d_harms = nullptr;
};

struct Plan {
    PzlConfig cfg{};
    PzlGeometry g{};
    std::vector<PzlHarm> harms;
    std::vector<double> startr;
    std::vector<SizeClass> classes;
    std::vector<uint16_t> zinds;  // [nh][numz]
    std::vector<int> winds;       // [nh][numw]
    int fnumw = 1, fnumz = 1, fzlo = 0, fwlo = 0;
    int max_fftlen = 0;
    int B = 8;
    int device = 0;
    hipStream_t st = nullptr;
    // device state
    PzlHarm *d_harms = nullptr;
    float2 *d_bank = nullptr;
    double *d_startr = nullptr;
    uint16_t *d_zinds = nullptr;
    int *d_winds = nullptr;
    const float2 *d_fft = nullptr;
    float2 *d_fft_own = nullptr;
    long long fft_bins = 0;
    float2 *d_segs = nullptr;
    long long seg_stride = 0;
    uint16_t *d_rinds = nullptr;
    PzlHrTab *d_hr = nullptr;
    float *d_planes = nullptr;
    long long plane_stride = 0;
    PzlHit *d_hits = nullptr;
    unsigned long long *d_nhits = nullptr;  // [0] hit count, [1] index-map violations
    unsigned long long maxhits = 0;
    float *d_sums = nullptr;
    std::vector<PzlHit> last_hits;
};

// ---------------------------------------------------------------------------
// Geometry (create_accelobs, accel_utils.c:1636-2037; calc_fftlen :137-149)
// ---------------------------------------------------------------------------

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

int compute_geometry(const PzlConfig &c, PzlGeometry &g, std::vector<PzlHarm> *harms_out,
                     std::vector<double> *startr_out)
{
    memset(&g, 0, sizeof g);
    int zmax = c.zmax, wmax = c.wmax;
    if (zmax < 0 || wmax < 0) return fail(PZL_ERR_ARG, "zmax and wmax must be >= 0");
    if (zmax % PZL_DZ) zmax = (zmax / PZL_DZ + 1) * PZL_DZ;
    const int nh_ok = c.numharm;
    if (nh_ok != 1 && nh_ok != 2 && nh_ok != 4 && nh_ok != 8 && nh_ok != 16 && nh_ok != 32)
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
    if (c.floP) {
        g.rlo = floor(c.flo * g.T);
    } else {
        g.rlo = c.rloP ? c.rlo : 1.0;
    }
    if (g.rlo < g.lobin) g.rlo = (double)g.lobin;
    if (g.rlo > g.numbins - 1) return fail(PZL_ERR_ARG, "low frequency to search is above the highest bin");
    g.highestbin = g.numbins - 1;
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

    // rsteps (accelsearch_lib.c:569-593 / accelsearch.c:183-188)
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

// ---------------------------------------------------------------------------
// Candidate list with insert_new_accelcand semantics (accel_utils.c:382-480)
// ---------------------------------------------------------------------------

class CandList {
  public:
    void insert(float power, float sigma, int numharm, double rr, double zz, double ww)
    {
        const PzlCand nc{power, sigma, numharm, rr, zz, ww};
        const double CLOSE = 15.0;
        if (v_.empty()) {
            v_.push_back(nc);
            return;
        }
        // the reference walks from the head to the first node with r >= rr
        // (or the last node); on a non-decreasing list that is lower_bound.
        size_t tmp;
        if (inversions_ == 0) {
            tmp = std::lower_bound(v_.begin(), v_.end(), rr, [](const PzlCand &a, double x) { return a.r < x; }) -
                  v_.begin();
        } else {
            tmp = 0;
            while (tmp + 1 < v_.size() && v_[tmp].r < rr) tmp++;
        }
        if (tmp == v_.size()) tmp = v_.size() - 1;
        const bool has_prev = tmp > 0;
        const size_t prev = tmp - 1;
        const double next_diff = fabs(rr - v_[tmp].r);
        const double prev_diff = has_prev ? fabs(rr - v_[prev].r) : CLOSE + 1.0;
        if (prev_diff < CLOSE) {
            bool added = false;
            if (v_[prev].sigma < sigma) {
                set(prev, nc);
                added = true;
            }
            if (next_diff < CLOSE && v_[tmp].sigma < sigma) {
                if (added) erase(tmp);
                else set(tmp, nc);
            }
        } else if (next_diff < CLOSE) {
            if (v_[tmp].sigma < sigma) set(tmp, nc);
        } else {
            if (tmp + 1 == v_.size() && v_[tmp].r < rr - CLOSE) insert_at(v_.size(), nc);
            else insert_at(tmp, nc);  // before tmp (or at the head)
        }
    }
    const std::vector<PzlCand> &items() const { return v_; }

  private:
    int inv(size_t i) const { return (i + 1 < v_.size() && v_[i].r > v_[i + 1].r) ? 1 : 0; }
    void touch_before(size_t i)
    {
        if (i > 0) inversions_ -= inv(i - 1);
        inversions_ -= inv(i);
    }
    void touch_after(size_t i)
    {
        if (i > 0) inversions_ += inv(i - 1);
        if (i < v_.size()) inversions_ += inv(i);
    }
    void set(size_t i, const PzlCand &c)
    {
        touch_before(i);
        v_[i] = c;
        touch_after(i);
    }
    void erase(size_t i)
    {
        touch_before(i);
        v_.erase(v_.begin() + i);
        if (i > 0) inversions_ += inv(i - 1);
    }
    void insert_at(size_t i, const PzlCand &c)
    {
        if (i > 0 && i < v_.size()) inversions_ -= inv(i - 1);
        v_.insert(v_.begin() + i, c);
        touch_after(i);
    }
    std::vector<PzlCand> v_;
    long long inversions_ = 0;
};

// ---------------------------------------------------------------------------
// Device buffers and launches
// ---------------------------------------------------------------------------

void free_plan(Plan *p)
{
    if (!p) return;
    hipSetDevice(p->device);
    for (auto &c : p->classes) hipFree(c.d_harms);
    hipFree(p->d_harms);
    hipFree(p->d_bank);
    hipFree(p->d_startr);
    hipFree(p->d_zinds);
    hipFree(p->d_winds);
    hipFree(p->d_fft_own);
    hipFree(p->d_segs);
    hipFree(p->d_rinds);
    hipFree(p->d_hr);
    hipFree(p->d_planes);
    hipFree(p->d_hits);
    hipFree(p->d_nhits);
    hipFree(p->d_sums);
    if (p->st) hipStreamDestroy(p->st);
    delete p;
}

struct Timers {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    struct Rec {
        int kind;  // 0 seg, 1 corr, 2 sum
        int fftlen;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    hipEvent_t get()
    {
        if (used == pool.size()) {
            hipEvent_t e;
            hipEventCreate(&e);
            pool.push_back(e);
        }
        return pool[used++];
    }
    void clear()
    {
        used = 0;
        recs.clear();
    }
    ~Timers()
    {
        for (auto e : pool) hipEventDestroy(e);
    }
};

int launch_batch(Plan &p, long long r0, int nb, float *d_sums, Timers *tm)
{
    const int nh = (int)p.harms.size();
    PzlBatchArgs ba{};
    ba.harms = p.d_harms;
    ba.nh = nh;
    ba.rstep0 = r0;
    ba.nb = nb;
    ba.startr = p.d_startr;
    ba.rstep_bins = (int)(p.g.corr_uselen * PZL_DR);
    ba.corr_uselen = p.g.corr_uselen;
    ba.fft = p.d_fft;
    ba.numbins = p.fft_bins;
    ba.file_lobin = p.g.lobin;
    ba.nph = p.cfg.nph;
    ba.segs = p.d_segs;
    ba.seg_stride = p.seg_stride;
    ba.rinds = p.d_rinds;
    ba.hr_tab = p.d_hr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (tm && tm->on) {
        e0 = tm->get();
        HIPCHK(hipEventRecord(e0, p.st));
    }
    HIPCHK(pzl_launch_segments(&ba, p.max_fftlen, p.st));
    if (tm && tm->on) {
        e1 = tm->get();
        HIPCHK(hipEventRecord(e1, p.st));
        tm->recs.push_back({0, 0, e0, e1});
    }
    for (auto &c : p.classes) {
        PzlCorrArgs ca{};
        ca.harms = p.d_harms;
        ca.nh = nh;
        ca.nb = nb;
        ca.class_harms = c.d_harms;
        ca.nclass = (int)c.harms.size();
        long long items = 0;
        for (int h : c.harms) items += (long long)p.harms[h].nrows * nb;
        ca.total_items = items;
        ca.segs = p.d_segs;
        ca.seg_stride = p.seg_stride;
        ca.bank = p.d_bank;
        ca.planes = p.d_planes;
        ca.plane_stride = p.plane_stride;
        ca.hr_tab = p.d_hr;
        if (tm && tm->on) {
            e0 = tm->get();
            HIPCHK(hipEventRecord(e0, p.st));
        }
        HIPCHK(pzl_launch_corr(c.fftlen, &ca, 256 * 16, p.st));
        if (tm && tm->on) {
            e1 = tm->get();
            HIPCHK(hipEventRecord(e1, p.st));
            tm->recs.push_back({1, c.fftlen, e0, e1});
        }
    }
    if (p.g.numharmstages > 1 || true) {
        PzlSumArgs sa{};
        sa.harms = p.d_harms;
        sa.nh = nh;
        sa.nb = nb;
        sa.nstages = p.g.numharmstages;
        sa.rstep0 = r0;
        sa.numz = p.fnumz;
        sa.numw = p.fnumw;
        sa.corr_uselen = p.g.corr_uselen;
        sa.planes = p.d_planes;
        sa.plane_stride = p.plane_stride;
        sa.rinds = p.d_rinds;
        sa.zinds = p.d_zinds;
        sa.winds = p.d_winds;
        for (int s = 0; s < 6; s++) sa.powcut[s] = p.g.powcut[s];
        sa.hits = p.d_hits;
        sa.nhits = p.d_nhits;
        sa.maxhits = p.maxhits;
        sa.sums_out = d_sums;
        sa.map_errors = (unsigned int *)(p.d_nhits + 1);
        if (tm && tm->on) {
            e0 = tm->get();
            HIPCHK(hipEventRecord(e0, p.st));
        }
        HIPCHK(pzl_launch_harmsum(&sa, p.st));
        if (tm && tm->on) {
            e1 = tm->get();
            HIPCHK(hipEventRecord(e1, p.st));
            tm->recs.push_back({2, 0, e0, e1});
        }
    }
    return 0;
}

double plane_alg_bytes(const Plan &p)
{
    // written once by the correlation and read once by the harmonic sum
    double b = 0;
    for (auto &H : p.harms) b += 2.0 * H.nrows * (double)H.max_numrs * 4.0;
    return b;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char *pzl_last_error(void) { return g_err.c_str(); }

int pzl_device_count(int *count)
{
    HIPCHK(hipGetDeviceCount(count));
    return 0;
}

int pzl_geometry_from_config(const PzlConfig *cfg, PzlGeometry *g)
{
    if (!cfg || !g) return fail(PZL_ERR_ARG, "null argument");
    return compute_geometry(*cfg, *g, nullptr, nullptr);
}

int pzl_plan_create(const PzlConfig *cfg, int device, void **plan_out)
{
    if (!cfg || !plan_out) return fail(PZL_ERR_ARG, "null argument");
    *plan_out = nullptr;
    Plan *p = new Plan();
    p->cfg = *cfg;
    p->device = device;
    int rc = compute_geometry(*cfg, p->g, &p->harms, &p->startr);
    if (rc) {
        delete p;
        return rc;
    }
    for (auto &H : p->harms) {
        if (!pzl_corr_supported(H.fftlen) || H.fftlen > 4096) {
            delete p;
            return fail(PZL_ERR_UNSUPPORTED, "correlation length %d (harmonic %d/%d) is not supported", H.fftlen,
                        H.harmnum, H.numharm);
        }
        const int numkern = 2 * PZL_NUMBETWEEN * pzl_w_resp_halfwidth(-H.zmax, -H.wmax);
        if (numkern > 2048) {
            delete p;
            return fail(PZL_ERR_UNSUPPORTED, "kernel width %d exceeds 2048 points", numkern);
        }
        p->max_fftlen = std::max(p->max_fftlen, H.fftlen);
    }
    p->B = cfg->batchsize > 0 ? cfg->batchsize : 8;
    if (p->g.nrsteps > 0 && p->B > p->g.nrsteps) p->B = (int)p->g.nrsteps;
    if (p->B < 1) p->B = 1;
    const int nh = (int)p->harms.size();
    const PzlHarm &F = p->harms[0];
    p->fnumz = F.numz;
    p->fnumw = F.numw;
    p->fzlo = F.zlo;
    p->fwlo = F.wlo;
    // rstep-independent z and w maps (accel_utils.c:1123-1131, 1362-1366)
    p->zinds.assign((size_t)nh * F.numz, 0);
    p->winds.assign((size_t)nh * F.numw, 0);
    for (int h = 0; h < nh; h++) {
        const PzlHarm &H = p->harms[h];
        for (int jj = 0; jj < F.numz; jj++) {
            const double zz = p->g.zlo + jj * PZL_DZ;
            p->zinds[(size_t)h * F.numz + jj] = (uint16_t)pzl_index_from_z(pzl_calc_required_z(H.hf, zz), H.zlo);
        }
        for (int ii = 0; ii < F.numw; ii++) {
            const int ww = F.wlo + ii * PZL_DW;
            p->winds[(size_t)h * F.numw + ii] = pzl_index_from_w(pzl_calc_required_w(H.hf, ww), H.wlo);
        }
    }
    std::map<int, std::vector<int>> cls;
    for (int h = 0; h < nh; h++) cls[p->harms[h].fftlen].push_back(h);
    // longest first: better tail behaviour inside a batch
    for (auto it = cls.rbegin(); it != cls.rend(); ++it) {
        SizeClass c;
        c.fftlen = it->first;
        c.harms = it->second;
        p->classes.push_back(c);
    }

    auto cleanup = [&](int code) {
        free_plan(p);
        return code;
    };
#define PCHK(call)                                                                                      \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return cleanup(fail(PZL_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)));           \
    } while (0)
    PCHK(hipSetDevice(device));
    PCHK(hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking));
    PCHK(hipMalloc(&p->d_harms, sizeof(PzlHarm) * nh));
    PCHK(hipMemcpy(p->d_harms, p->harms.data(), sizeof(PzlHarm) * nh, hipMemcpyHostToDevice));
    for (auto &c : p->classes) {
        PCHK(hipMalloc(&c.d_harms, sizeof(int) * c.harms.size()));
        PCHK(hipMemcpy(c.d_harms, c.harms.data(), sizeof(int) * c.harms.size(), hipMemcpyHostToDevice));
    }
    const long long bank_elems = p->g.bank_bytes / 8;
    PCHK(hipMalloc(&p->d_bank, sizeof(float2) * bank_elems));
    const size_t nstart = std::max<size_t>(p->startr.size(), 1);
    PCHK(hipMalloc(&p->d_startr, sizeof(double) * nstart));
    if (!p->startr.empty())
        PCHK(hipMemcpy(p->d_startr, p->startr.data(), sizeof(double) * p->startr.size(), hipMemcpyHostToDevice));
    PCHK(hipMalloc(&p->d_zinds, sizeof(uint16_t) * p->zinds.size()));
    PCHK(hipMemcpy(p->d_zinds, p->zinds.data(), sizeof(uint16_t) * p->zinds.size(), hipMemcpyHostToDevice));
    PCHK(hipMalloc(&p->d_winds, sizeof(int) * p->winds.size()));
    PCHK(hipMemcpy(p->d_winds, p->winds.data(), sizeof(int) * p->winds.size(), hipMemcpyHostToDevice));
    p->seg_stride = 0;
    for (auto &H : p->harms) p->seg_stride += H.fftlen / 2;
    p->plane_stride = p->g.plane_bytes_per_rstep / 4;
    PCHK(hipMalloc(&p->d_segs, sizeof(float2) * p->seg_stride * p->B));
    PCHK(hipMalloc(&p->d_rinds, sizeof(uint16_t) * (size_t)p->B * nh * p->g.corr_uselen));
    PCHK(hipMalloc(&p->d_hr, sizeof(PzlHrTab) * (size_t)p->B * nh));
    PCHK(hipMalloc(&p->d_planes, sizeof(float) * (size_t)p->plane_stride * p->B));
    p->maxhits = (unsigned long long)(cfg->maxhits > 0 ? cfg->maxhits : kDefaultMaxHits);
    PCHK(hipMalloc(&p->d_hits, sizeof(PzlHit) * p->maxhits));
    PCHK(hipMalloc(&p->d_nhits, 2 * sizeof(unsigned long long)));
    // The kernel bank, built once and kept resident.
    PCHK(pzl_launch_bank_build(p->d_harms, nh, p->g.numharm_kernels, p->max_fftlen, p->d_bank, p->st));
    PCHK(hipStreamSynchronize(p->st));
    p->g.device_bytes = p->g.bank_bytes + (long long)sizeof(float) * p->plane_stride * p->B +
                        (long long)sizeof(float2) * p->seg_stride * p->B + (long long)sizeof(PzlHit) * p->maxhits +
                        (long long)sizeof(uint16_t) * p->B * nh * p->g.corr_uselen;
#undef PCHK
    *plan_out = p;
    return 0;
}

void pzl_plan_destroy(void *plan) { free_plan((Plan *)plan); }

int pzl_plan_geometry(void *plan, PzlGeometry *g)
{
    if (!plan || !g) return fail(PZL_ERR_ARG, "null argument");
    *g = ((Plan *)plan)->g;
    return 0;
}

int pzl_plan_harm(void *plan, int idx, PzlHarmInfo *h)
{
    Plan *p = (Plan *)plan;
    if (!p || !h || idx < 0 || idx >= (int)p->harms.size()) return fail(PZL_ERR_ARG, "bad harmonic index");
    const PzlHarm &H = p->harms[idx];
    h->numharm = H.numharm;
    h->harmnum = H.harmnum;
    h->stage = H.stage;
    h->zmax = H.zmax;
    h->wmax = H.wmax;
    h->numz = H.numz;
    h->numw = H.numw;
    h->fftlen = H.fftlen;
    h->halfwidth = H.halfwidth;
    h->max_numrs = H.max_numrs;
    return 0;
}

int pzl_set_spectrum(void *plan, const float *host_fft, long long numbins)
{
    Plan *p = (Plan *)plan;
    if (!p || !host_fft || numbins < 1) return fail(PZL_ERR_ARG, "bad spectrum");
    HIPCHK(hipSetDevice(p->device));
    if (p->d_fft_own && p->fft_bins != numbins) {
        hipFree(p->d_fft_own);
        p->d_fft_own = nullptr;
    }
    if (!p->d_fft_own) HIPCHK(hipMalloc(&p->d_fft_own, sizeof(float2) * numbins));
    HIPCHK(hipMemcpy(p->d_fft_own, host_fft, sizeof(float2) * numbins, hipMemcpyHostToDevice));
    p->d_fft = p->d_fft_own;
    p->fft_bins = numbins;
    return 0;
}

int pzl_set_spectrum_device(void *plan, const void *dev_fft, long long numbins)
{
    Plan *p = (Plan *)plan;
    if (!p || !dev_fft || numbins < 1) return fail(PZL_ERR_ARG, "bad spectrum");
    p->d_fft = (const float2 *)dev_fft;
    p->fft_bins = numbins;
    return 0;
}

int pzl_search(void *plan, long long rstep_lo, long long rstep_hi, PzlCand *cands, long long maxcands,
               long long *ncands, PzlStats *stats)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    if (!p->d_fft) return fail(PZL_ERR_NOSPECTRUM, "no spectrum set");
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(p->device));
    const long long nr = p->g.nrsteps;
    if (rstep_lo < 0) rstep_lo = 0;
    if (rstep_hi < 0 || rstep_hi > nr) rstep_hi = nr;
    Timers tm;
    tm.on = (stats != nullptr);
    HIPCHK(hipMemsetAsync(p->d_nhits, 0, 2 * sizeof(unsigned long long), p->st));
    hipEvent_t ev_start = nullptr, ev_end = nullptr;
    if (tm.on) {
        ev_start = tm.get();
        HIPCHK(hipEventRecord(ev_start, p->st));
    }
    for (long long r0 = rstep_lo; r0 < rstep_hi; r0 += p->B) {
        const int nb = (int)std::min<long long>(p->B, rstep_hi - r0);
        int rc = launch_batch(*p, r0, nb, nullptr, &tm);
        if (rc) return rc;
    }
    if (tm.on) {
        ev_end = tm.get();
        HIPCHK(hipEventRecord(ev_end, p->st));
    }
    unsigned long long counters[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(counters, p->d_nhits, sizeof counters, hipMemcpyDeviceToHost, p->st));
    HIPCHK(hipStreamSynchronize(p->st));
    const auto t1 = std::chrono::steady_clock::now();
    const unsigned long long nh = counters[0];
    if (counters[1]) return fail(PZL_ERR_HIP, "index maps left a subharmonic plane %llu times", counters[1]);
    if (nh > p->maxhits) {
        p->last_hits.clear();
        if (ncands) *ncands = 0;
        return fail(PZL_ERR_OVERFLOW, "hit buffer overflow: %llu hits > %llu", nh, p->maxhits);
    }
    p->last_hits.resize(nh);
    if (nh) HIPCHK(hipMemcpy(p->last_hits.data(), p->d_hits, sizeof(PzlHit) * nh, hipMemcpyDeviceToHost));
    std::sort(p->last_hits.begin(), p->last_hits.end(),
              [](const PzlHit &a, const PzlHit &b) { return a.key < b.key; });
    // decode in PRESTO C scan order and insert (accel_utils.c:1518-1549)
    CandList cl;
    const long long S = p->g.numharmstages, W = p->fnumw, Z = p->fnumz, R = p->g.corr_uselen;
    for (const PzlHit &h : p->last_hits) {
        long long k = h.key;
        const int r = (int)(k % R);
        k /= R;
        const int z = (int)(k % Z);
        k /= Z;
        const int w = (int)(k % W);
        k /= W;
        const int stage = (int)(k % S);
        const long long rstep = k / S;
        const int numharm = 1 << stage;
        const long long frlo = (long long)floor(pzl_calc_required_r(1.0, p->startr[rstep]));
        const float pow = h.power;
        const float sig = (float)pzl_candidate_sigma(pow, numharm, (double)p->g.numindep[stage]);
        const double rr = (frlo + r * (double)PZL_DR) / (double)numharm;
        const double zz = (p->fzlo + z * (double)PZL_DZ) / (double)numharm;
        const double ww = (p->fwlo + w * (double)PZL_DW) / (double)numharm;
        cl.insert(pow, sig, numharm, rr, zz, ww);
    }
    const auto &items = cl.items();
    if (ncands) *ncands = (long long)items.size();
    if (cands)
        for (long long i = 0; i < (long long)items.size() && i < maxcands; i++) cands[i] = items[i];
    const auto t2 = std::chrono::steady_clock::now();
    if (stats) {
        memset(stats, 0, sizeof *stats);
        float ms = 0;
        std::map<int, double> per_class;
        for (auto &r : tm.recs) {
            hipEventElapsedTime(&ms, r.a, r.b);
            if (r.kind == 0) stats->ms_segments += ms;
            else if (r.kind == 1) {
                stats->ms_corr += ms;
                per_class[r.fftlen] += ms;
                stats->corr_launches++;
            } else stats->ms_sum += ms;
        }
        for (auto &kv : per_class) stats->ms_corr_max_class = std::max(stats->ms_corr_max_class, kv.second);
        hipEventElapsedTime(&ms, ev_start, ev_end);
        stats->ms_gpu = ms;
        stats->ms_wall = std::chrono::duration<double, std::milli>(t2 - t0).count();
        stats->ms_host = std::chrono::duration<double, std::milli>(t2 - t1).count();
        stats->nhits = (long long)nh;
        stats->nrsteps = rstep_hi - rstep_lo;
        stats->cells = (double)stats->nrsteps * p->g.corr_uselen * p->g.numz * std::max(p->g.numw, 1);
        double seg_bytes = 0;
        for (auto &H : p->harms) seg_bytes += H.fftlen / 2 * 8.0;
        const double planes = plane_alg_bytes(*p);
        stats->bytes_alg = stats->nrsteps * (planes + seg_bytes) + (double)p->g.bank_bytes;
        stats->bytes_alg_corr = stats->nrsteps * (planes / 2 + seg_bytes) +
                                (double)p->g.bank_bytes * std::ceil((double)stats->nrsteps / p->B);
    }
    return 0;
}

int pzl_get_hits(void *plan, long long *keys, float *powers, long long maxhits, long long *nhits)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    if (nhits) *nhits = (long long)p->last_hits.size();
    for (long long i = 0; i < (long long)p->last_hits.size() && i < maxhits; i++) {
        if (keys) keys[i] = p->last_hits[i].key;
        if (powers) powers[i] = p->last_hits[i].power;
    }
    return 0;
}

int pzl_get_kernel(void *plan, int idx, int row, float *out)
{
    Plan *p = (Plan *)plan;
    if (!p || !out || idx < 0 || idx >= (int)p->harms.size() || row < 0 || row >= p->harms[idx].nrows)
        return fail(PZL_ERR_ARG, "bad kernel index");
    const PzlHarm &H = p->harms[idx];
    HIPCHK(hipSetDevice(p->device));
    HIPCHK(hipMemcpy(out, p->d_bank + H.bank_off + (long long)row * H.fftlen, sizeof(float2) * H.fftlen,
                     hipMemcpyDeviceToHost));
    return 0;
}

int pzl_compute_plane(void *plan, long long rstep, int idx, float *out, long long maxfloats, int *numrs,
                      long long *rlo)
{
    Plan *p = (Plan *)plan;
    if (!p || idx < 0 || idx >= (int)p->harms.size() || rstep < 0 || rstep >= p->g.nrsteps)
        return fail(PZL_ERR_ARG, "bad plane request");
    if (!p->d_fft) return fail(PZL_ERR_NOSPECTRUM, "no spectrum set");
    HIPCHK(hipSetDevice(p->device));
    int rc = launch_batch(*p, rstep, 1, nullptr, nullptr);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(p->st));
    PzlHrTab hr;
    HIPCHK(hipMemcpy(&hr, p->d_hr + idx, sizeof hr, hipMemcpyDeviceToHost));
    const PzlHarm &H = p->harms[idx];
    if (numrs) *numrs = hr.numrs;
    if (rlo) *rlo = hr.rlo;
    if (out) {
        const long long need = (long long)H.nrows * hr.numrs;
        if (need > maxfloats) return fail(PZL_ERR_ARG, "output too small (%lld floats needed)", need);
        HIPCHK(hipMemcpy2D(out, sizeof(float) * hr.numrs, p->d_planes + H.plane_off, sizeof(float) * H.rs_stride,
                           sizeof(float) * hr.numrs, H.nrows, hipMemcpyDeviceToHost));
    }
    return 0;
}

int pzl_compute_sums(void *plan, long long rstep, float *out, long long maxfloats)
{
    Plan *p = (Plan *)plan;
    if (!p || !out || rstep < 0 || rstep >= p->g.nrsteps) return fail(PZL_ERR_ARG, "bad sums request");
    if (!p->d_fft) return fail(PZL_ERR_NOSPECTRUM, "no spectrum set");
    const long long need = (long long)p->fnumw * p->fnumz * p->g.corr_uselen;
    if (need > maxfloats) return fail(PZL_ERR_ARG, "output too small");
    HIPCHK(hipSetDevice(p->device));
    if (!p->d_sums) HIPCHK(hipMalloc(&p->d_sums, sizeof(float) * need));
    HIPCHK(hipMemsetAsync(p->d_nhits, 0, 2 * sizeof(unsigned long long), p->st));
    int rc = launch_batch(*p, rstep, 1, p->d_sums, nullptr);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(p->st));
    HIPCHK(hipMemcpy(out, p->d_sums, sizeof(float) * need, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
