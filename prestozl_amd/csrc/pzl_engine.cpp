// pzl_engine.cpp -- host side of the MI355X jerk search (C ABI in
// include/pzl_accel.h).
//
// Replaces the reference's GPU driver and host helpers:
//   kernel bank                 create_subharminfos accel_utils.c:236-282 and
//                               fkern_host_to_dev accel_utils_cu.c:378-429
//                               (built on the GPU here, once per plan)
//   rstep batches               accelsearch_GPU, accelsearch_lib.c:473-901
//   hit drain + ordering        accelsearch_lib.c:784-866, thrust sort
//                               accel_kernels.cu:990-1006
//   candidate insertion         insert_to_cands accel_utils_cu.c:857-929 with
//                               insert_new_accelcand accel_utils.c:382-480
// Per rstep everything runs on the GPU: segment preparation, correlation,
// harmonic summing, compaction, ordering and the significance of each hit.
// The host applies the ordered hits to the candidate list while the GPU is
// already working on the next batch of rsteps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "pzl_accel.h"
#include "pzl_candlist.h"
#include "pzl_kernels.h"
#include "pzl_launch.h"
#include "pzl_math.h"
#include "pzl_plan.h"
#include "pzl_stats.h"

using namespace pzl;

namespace {

const long long kDefaultHitCap = 1LL << 26;  // hits per batch slot (12 bytes each)

// Kernel banks are shared by every plan with the same harmonic shapes on a
// device (files of one (zmax, wmax, numharm) class, DM trials, served
// requests) and stay resident until evicted: unreferenced banks are dropped
// oldest-first once the cache exceeds its byte budget.
struct BankEntry {
    float2 *ptr = nullptr;
    long long bytes = 0;
    int refs = 0;
    unsigned long long last_use = 0;
};
std::mutex g_bank_mu;
std::map<std::vector<long long>, BankEntry> g_banks;
unsigned long long g_bank_clock = 0;
long long g_bank_budget = 64LL << 30;

std::vector<long long> bank_key(const Plan &p)
{
    std::vector<long long> k{p.device};
    for (auto &H : p.harms) {
        k.push_back(H.numharm);
        k.push_back(H.harmnum);
        k.push_back(H.zmax);
        k.push_back(H.wmax);
        k.push_back(H.fftlen);
    }
    return k;
}

void release_bank(Plan *p)
{
    if (!p->d_bank) return;
    std::lock_guard<std::mutex> lk(g_bank_mu);
    auto it = g_banks.find(bank_key(*p));
    if (it != g_banks.end() && it->second.ptr == p->d_bank) it->second.refs--;
    p->d_bank = nullptr;
}

void evict_banks_locked(long long incoming)
{
    long long total = incoming;
    for (auto &kv : g_banks) total += kv.second.bytes;
    while (total > g_bank_budget) {
        auto victim = g_banks.end();
        for (auto it = g_banks.begin(); it != g_banks.end(); ++it)
            if (it->second.refs == 0 && (victim == g_banks.end() || it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim == g_banks.end()) break;
        (void)hipSetDevice((int)victim->first[0]);
        (void)hipFree(victim->second.ptr);
        total -= victim->second.bytes;
        g_banks.erase(victim);
    }
}

// The bank of this plan's shape: from the cache, or built on the GPU now.
int acquire_bank(Plan *p, bool *built)
{
    std::lock_guard<std::mutex> lk(g_bank_mu);
    const auto key = bank_key(*p);
    auto it = g_banks.find(key);
    *built = false;
    if (it != g_banks.end()) {
        it->second.refs++;
        it->second.last_use = ++g_bank_clock;
        p->d_bank = it->second.ptr;
        return 0;
    }
    evict_banks_locked(p->g.bank_bytes);
    BankEntry e;
    e.bytes = p->g.bank_bytes;
    PZL_HIPCHK(hipMalloc(&e.ptr, e.bytes));
    hipError_t err = pzl_launch_bank_build(p->d_harms, (int)p->harms.size(), p->g.numharm_kernels, p->max_fftlen,
                                           e.ptr, p->st);
    if (err == hipSuccess) err = hipStreamSynchronize(p->st);
    if (err != hipSuccess) {
        (void)hipFree(e.ptr);
        return fail(PZL_ERR_HIP, "kernel bank build failed: %s", hipGetErrorString(err));
    }
    e.refs = 1;
    e.last_use = ++g_bank_clock;
    g_banks[key] = e;
    p->d_bank = e.ptr;
    *built = true;
    return 0;
}

// Batch size, hit-slot capacity and slot count of a plan (pzl_plan_create
// uses them; the request loop uses them to size a request before admission).
struct PlanSizes {
    int B;
    long long hitcap;
    int nslots;
    long long seg_stride;
};

PlanSizes plan_sizes(const PzlConfig &cfg, const PzlGeometry &g, const std::vector<PzlHarm> &harms)
{
    PlanSizes z;
    z.B = cfg.batchsize > 0 ? cfg.batchsize : 8;
    if (g.nrsteps > 0 && z.B > g.nrsteps) z.B = (int)g.nrsteps;
    if (z.B < 1) z.B = 1;
    const PzlHarm &F = harms[0];
    // one rstep can never overflow a slot
    const long long per_rstep = (long long)g.corr_uselen * F.numz * F.numw * g.numharmstages;
    long long cap = cfg.maxhits > 0 ? cfg.maxhits : kDefaultHitCap;
    cap = std::min<long long>(cap, per_rstep * z.B);
    z.hitcap = std::max<long long>(cap, per_rstep);
    const long long nbatches = (g.nrsteps + z.B - 1) / z.B;
    z.nslots = (int)std::max<long long>(1, std::min<long long>(Plan::kSlots, nbatches));
    z.seg_stride = 0;
    for (auto &H : harms) z.seg_stride += H.fftlen / 2;
    return z;
}

// The buffers a plan carves from its one device arena, in carving order
// (pzl_plan_create assigns them in this order): harmonic tables, two sets of
// batch buffers (segments, r-index tables, harmonic r maps, planes), hit
// slots and ordering buffers.  The sort scratch is the size the sort
// reports for the slot capacity (a host computation, also before a plan exists).
constexpr long long kArenaAlign = 256;

std::vector<long long> plan_arena_items(const PzlGeometry &g, const PlanSizes &z, const std::vector<PzlHarm> &harms,
                                        long long sort_temp)
{
    const long long nh = (long long)harms.size();
    const PzlHarm &F = harms[0];
    if (sort_temp < 0) {
        // the sort reports its scratch without touching the device
        size_t b = 0;
        (void)pzl_sort_hits(nullptr, &b, nullptr, nullptr, nullptr, nullptr, z.hitcap, 64, nullptr);
        sort_temp = (long long)b;
    }
    std::vector<long long> v{(long long)sizeof(PzlHarm) * nh, (long long)sizeof(int) * nh,
                             (long long)sizeof(double) * (g.nrsteps + 1), (long long)sizeof(uint16_t) * nh * F.numz,
                             (long long)sizeof(int) * nh * F.numw,
                             (long long)sizeof(int) * F.numw * F.numz * std::max<long long>(nh - 1, 1)};
    for (int k = 0; k < Plan::kSets; k++) {
        v.push_back((long long)sizeof(float2) * z.seg_stride * z.B);
        v.push_back((long long)sizeof(uint16_t) * z.B * nh * g.corr_uselen);
        v.push_back((long long)sizeof(PzlHrTab) * z.B * nh);
        v.push_back((long long)sizeof(float) * (g.plane_bytes_per_rstep / 4 * z.B + PZL_PLANE_SLACK));
    }
    for (int s = 0; s < z.nslots; s++) {
        v.push_back((long long)sizeof(long long) * z.hitcap);
        v.push_back((long long)sizeof(float) * z.hitcap);
    }
    v.push_back((long long)sizeof(unsigned long long) * 2 * Plan::kSlots);
    v.push_back((long long)sizeof(long long) * z.hitcap);
    v.push_back((long long)sizeof(float) * z.hitcap);
    v.push_back((long long)sizeof(PzlHostHit) * z.hitcap);
    v.push_back(std::max<long long>(sort_temp, 16));
    return v;
}

long long arena_bytes(const std::vector<long long> &items)
{
    long long t = 0;
    for (long long b : items) t += (b + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
    return t;
}

// HBM held by a plan whose spectrum was set from the host: kernel bank,
// spectrum and arena.
long long plan_device_bytes(const PzlGeometry &g, const PlanSizes &z, const std::vector<PzlHarm> &harms,
                            long long sort_temp)
{
    return g.bank_bytes + (long long)sizeof(float2) * g.numbins + arena_bytes(plan_arena_items(g, z, harms, sort_temp));
}

// hipMalloc of a plan-owned buffer, counted in p->dev_bytes
template <typename T> hipError_t plan_malloc(Plan *p, T **ptr, size_t bytes)
{
    hipError_t e = hipMalloc((void **)ptr, bytes);
    if (e == hipSuccess) p->dev_bytes += (long long)bytes;
    return e;
}

// Correlation variants chosen by pzl_autotune, per (device, FFT length,
// batch size, row counts of the class): later plans of the same shape use them.
std::mutex g_tune_mu;
std::map<std::vector<long long>, int> g_tuned;

std::vector<long long> tune_key(const Plan &p, const SizeClass &c)
{
    std::vector<long long> k{p.device, c.fftlen, p.B};
    for (int n : c.nrows) k.push_back(n);
    return k;
}

int tuned_variant(const Plan &p, const SizeClass &c)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(tune_key(p, c));
    return it == g_tuned.end() ? 0 : it->second;
}

// ... and the grouped launch (-1: one launch per length; g: the shorter
// lengths in group g; g + ngroups: every length in group g), per (device,
// batch size, every class's length and row counts).
std::map<std::vector<long long>, int> g_tuned_group;

std::vector<long long> group_key(const Plan &p)
{
    std::vector<long long> k{p.device, p.B};
    for (auto &c : p.classes) {
        k.push_back(-c.fftlen);
        for (int n : c.nrows) k.push_back(n);
    }
    return k;
}

int tuned_group(const Plan &p)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned_group.find(group_key(p));
    return it == g_tuned_group.end() ? -1 : it->second;
}

// Whether group g holds every FFT length of the plan but the longest (all:
// the longest too).
bool group_covers(const Plan &p, int g, bool all)
{
    if (g < 0 || g >= pzl_corr_num_groups() || p.classes.size() < 2) return false;
    int longest = 0;
    for (auto &c : p.classes) longest = std::max(longest, c.fftlen);
    for (auto &c : p.classes)
        if ((all || c.fftlen != longest) && !pzl_corr_group_has(g, c.fftlen)) return false;
    return true;
}

// Every pinned host allocation of the engine, so that each asynchronous copy
// into or out of pinned memory can be checked to lie inside a live
// allocation before it is enqueued (a copy into freed pinned memory would
// otherwise be turned by the runtime into a host write to unmapped pages).
std::mutex g_pinreg_mu;
std::map<uintptr_t, size_t> g_pinreg;

hipError_t pinned_alloc(void **ptr, size_t bytes)
{
    const hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    if (e == hipSuccess) {
        std::lock_guard<std::mutex> lk(g_pinreg_mu);
        g_pinreg[(uintptr_t)*ptr] = bytes;
    }
    return e;
}

void pinned_free(void *ptr)
{
    if (!ptr) return;
    {
        std::lock_guard<std::mutex> lk(g_pinreg_mu);
        g_pinreg.erase((uintptr_t)ptr);
    }
    (void)hipHostFree(ptr);
}

bool pinned_range_ok(const void *ptr, size_t bytes)
{
    std::lock_guard<std::mutex> lk(g_pinreg_mu);
    auto it = g_pinreg.upper_bound((uintptr_t)ptr);
    if (it == g_pinreg.begin()) return false;
    --it;
    return (uintptr_t)ptr + bytes <= it->first + it->second;
}

#define PZL_PINNED_CHECK(ptr, bytes)                                                                        \
    do {                                                                                                   \
        if (!pinned_range_ok((ptr), (bytes)))                                                              \
            return fail(PZL_ERR_HIP, "copy of %zu bytes at %p is outside every live pinned allocation (%s:%d)", \
                        (size_t)(bytes), (const void *)(ptr), __FILE__, __LINE__);                         \
    } while (0)

// Pinned hit-record buffers outlive their plans: pinning host memory is slow
// (on the order of 0.15 s per GiB) and waits for the device, so a plan takes
// buffers from this cache before pinning new memory, and a destroyed plan
// leaves its buffers here, up to kPinCacheLimit bytes.
std::mutex g_pin_mu;
std::multimap<long long, PzlHostHit *> g_pin_cache;  // capacity in records -> buffer
long long g_pin_cache_bytes = 0;
constexpr long long kPinCacheLimit = 8LL << 30;

// The smallest cached buffer of at least 'want' records and at most four times
// that, or null; *cap gets its capacity.
PzlHostHit *pin_take(long long want, long long *cap)
{
    std::lock_guard<std::mutex> lk(g_pin_mu);
    auto it = g_pin_cache.lower_bound(want);
    if (it == g_pin_cache.end() || it->first > 4 * want) return nullptr;
    *cap = it->first;
    PzlHostHit *ptr = it->second;
    g_pin_cache_bytes -= (long long)sizeof(PzlHostHit) * it->first;
    g_pin_cache.erase(it);
    return ptr;
}

void pin_put(PzlHostHit *ptr, long long cap)
{
    if (!ptr) return;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        const long long bytes = (long long)sizeof(PzlHostHit) * cap;
        if (g_pin_cache_bytes + bytes <= kPinCacheLimit) {
            g_pin_cache.emplace(cap, ptr);
            g_pin_cache_bytes += bytes;
            return;
        }
    }
    pinned_free(ptr);
}

void free_plan(Plan *p)
{
    if (!p) return;
    (void)hipSetDevice(p->device);
    // every stream the plan made, once each (st, st3 and st4 are among the
    // ones created with the plan or the CU-masked ones)
    std::vector<hipStream_t> streams = {p->st,        p->st2,       p->st3,      p->st4,
                                        p->st5,       p->st_all[0], p->st_all[1], p->st_all[2]};
    for (auto &m : p->masked) streams.push_back(m.s);
    std::sort(streams.begin(), streams.end());
    streams.erase(std::unique(streams.begin(), streams.end()), streams.end());
    for (hipStream_t s : streams)
        if (s) (void)hipStreamSynchronize(s);
    release_bank(p);
    (void)hipFree(p->d_arena);
    (void)hipFree(p->d_fft_own);
    for (int k = 0; k < Plan::kSets; k++) {
        if (p->ev_corr[k]) (void)hipEventDestroy(p->ev_corr[k]);
        if (p->ev_sum[k]) (void)hipEventDestroy(p->ev_sum[k]);
        if (p->ev_seg[k]) (void)hipEventDestroy(p->ev_seg[k]);
        if (p->ev_side[k]) (void)hipEventDestroy(p->ev_side[k]);
    }
    for (int s = 0; s < Plan::kSlots; s++) {
        if (p->ev_count[s]) (void)hipEventDestroy(p->ev_count[s]);
        if (p->ev_free[s]) (void)hipEventDestroy(p->ev_free[s]);
    }
    if (p->ev_prep) (void)hipEventDestroy(p->ev_prep);
    if (p->ev_copied) (void)hipEventDestroy(p->ev_copied);
    pinned_free(p->h_counts);
    for (auto &hb : p->hpool) {
        pin_put(hb.ptr, hb.cap);
        if (hb.ev) (void)hipEventDestroy(hb.ev);
    }
    for (int k = 0; k < Plan::kSets; k++)
        if (p->gexec[k]) (void)hipGraphExecDestroy(p->gexec[k]);
    (void)hipFree(p->d_sums);
    for (hipStream_t s : streams)
        if (s) (void)hipStreamDestroy(s);
    delete p;
}

struct Timers {
    bool on = false;
    bool active = true;  // record this batch (searches sample every kSampleEvery-th batch)
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    struct Rec {
        int kind;  // 0 segments, 1 correlation, 2 harmonic sum, 3 graph-replayed batch
        int fftlen;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    hipEvent_t get()
    {
        if (used == pool.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            pool.push_back(e);
        }
        return pool[used++];
    }
    ~Timers()
    {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Records HIP events around one launch when timing is on.
struct Scope {
    Timers *tm;
    int kind, fftlen;
    hipStream_t st;
    hipEvent_t a = nullptr;
    Scope(Timers *t, int k, int f, hipStream_t s) : tm(t), kind(k), fftlen(f), st(s)
    {
        if (tm && tm->on && tm->active) {
            a = tm->get();
            (void)hipEventRecord(a, st);
        }
    }
    ~Scope()
    {
        if (a) {
            hipEvent_t b = tm->get();
            (void)hipEventRecord(b, st);
            tm->recs.push_back({kind, fftlen, a, b});
        }
    }
};

PzlCorrArgs corr_args(const Plan &p, const SizeClass &c, int nb, int set)
{
    PzlCorrArgs ca{};
    ca.harms = p.d_harms;
    ca.nh = (int)p.harms.size();
    ca.nb = nb;
    ca.class_harms = c.d_harms;
    ca.nclass = (int)c.harms.size();
    ca.segs = p.d_segs[set];
    ca.seg_stride = p.seg_stride;
    ca.bank = p.d_bank;
    ca.planes = p.d_planes[set];
    ca.plane_stride = p.plane_stride;
    ca.hr_tab = p.d_hr[set];
    return ca;
}

int launch_corrs(Plan &p, int nb, int set, Timers *tm, bool seg_stream);

// Segments of rsteps [r0, r0 + nb) into buffer set 'set' (r0_dev: read the
// first rstep from there), then, unless segments_only, their correlations on
// the correlation stream.
int launch_corr_part(Plan &p, long long r0, int nb, int set, Timers *tm, const long long *r0_dev,
                     bool seg_stream = false, bool segments_only = false)
{
    PzlBatchArgs ba{};
    ba.rstep0_dev = r0_dev;
    ba.harms = p.d_harms;
    ba.nh = (int)p.harms.size();
    ba.rstep0 = r0;
    ba.nb = nb;
    ba.startr = p.d_startr;
    ba.rstep_bins = (int)(p.g.corr_uselen * PZL_DR);
    ba.corr_uselen = p.g.corr_uselen;
    ba.fft = p.d_fft;
    ba.numbins = p.fft_bins;
    ba.file_lobin = p.g.lobin;
    ba.nph = p.cfg.nph;
    ba.segs = p.d_segs[set];
    ba.seg_stride = p.seg_stride;
    ba.rinds = p.d_rinds[set];
    ba.hr_tab = p.d_hr[set];
    // the segments of a batch can run ahead on their own stream, beside the
    // correlation of the previous batch, once the set's last sum is done
    hipStream_t ss = seg_stream ? p.st4 : p.st;
    if (seg_stream) PZL_HIPCHK(hipStreamWaitEvent(ss, p.ev_sum[set], 0));
    {
        Scope s(tm, 0, 0, ss);
        PZL_HIPCHK(pzl_launch_segments(&ba, p.max_fftlen, ss));
    }
    if (seg_stream) {
        PZL_HIPCHK(hipEventRecord(p.ev_seg[set], ss));
        PZL_HIPCHK(hipStreamWaitEvent(p.st, p.ev_seg[set], 0));
    }
    if (segments_only) return 0;
    return launch_corrs(p, nb, set, tm, seg_stream);
}

// The correlations of a batch whose segments are in buffer set 'set'.
int launch_corrs(Plan &p, int nb, int set, Timers *tm, bool seg_stream)
{
    hipStream_t ss = seg_stream ? p.st4 : p.st;
    // With the segments on their own stream, the shorter FFT lengths (small
    // grids that alone leave most of the chip idle) can follow them there and
    // run beside the longest length's correlation on the main stream.
    const bool side = seg_stream && p.side_classes && p.classes.size() > 1;
    size_t main_k = 0;
    for (size_t k = 1; k < p.classes.size(); k++)
        if (p.classes[k].fftlen > p.classes[main_k].fftlen) main_k = k;
    const bool grouped = p.side_group >= 0 && p.classes.size() > 1;
    if (grouped) {
        // the shorter lengths (with group_all: every length, on the main
        // stream) in one grouped launch, longest first
        hipStream_t cs = side && !p.group_all ? ss : p.st;
        std::vector<PzlCorrArgs> args;
        std::vector<const PzlCorrArgs *> argp;
        std::vector<int> lens;
        std::vector<const int *> rows;
        args.reserve(p.classes.size());
        for (size_t k = 0; k < p.classes.size(); k++) {
            if (k == main_k && !p.group_all) continue;
            const SizeClass &c = p.classes[k];
            args.push_back(corr_args(p, c, nb, set));
            lens.push_back(c.fftlen);
            rows.push_back(c.nrows.data());
        }
        for (auto &a : args) argp.push_back(&a);
        Scope s(tm, 1, 0, cs);
        PZL_HIPCHK(pzl_launch_corr_group(p.side_group, (int)args.size(), lens.data(), argp.data(), rows.data(), cs));
    }
    for (size_t k = 0; k < p.classes.size(); k++) {
        if (grouped && (k != main_k || p.group_all)) continue;
        const SizeClass &c = p.classes[k];
        hipStream_t cs = (side && k != main_k) ? ss : p.st;
        const PzlCorrArgs ca = corr_args(p, c, nb, set);
        Scope s(tm, 1, c.fftlen, cs);
        PZL_HIPCHK(pzl_launch_corr(c.fftlen, c.variant, &ca, c.nrows.data(), cs));
    }
    if (side) {
        PZL_HIPCHK(hipEventRecord(p.ev_side[set], ss));
        PZL_HIPCHK(hipStreamWaitEvent(p.st, p.ev_side[set], 0));
    }
    return 0;
}

// Harmonic sum of the batch held in buffer set 'set' into hit slot 'slot',
// on the summing stream.
int launch_sum_part(Plan &p, long long r0, int nb, int set, int slot, float *d_sums, Timers *tm)
{
    PzlSumArgs sa{};
    sa.harms = p.d_harms;
    sa.nh = (int)p.harms.size();
    sa.nb = nb;
    sa.nstages = p.g.numharmstages;
    sa.rstep0 = r0;
    sa.rstep0_dev = nullptr;
    sa.numz = p.fnumz;
    sa.numw = p.fnumw;
    sa.corr_uselen = p.g.corr_uselen;
    sa.planes = p.d_planes[set];
    sa.plane_stride = p.plane_stride;
    sa.rinds = p.d_rinds[set];
    sa.zinds = p.d_zinds;
    sa.winds = p.d_winds;
    sa.zoff_tab = p.d_zoff;
    for (int s = 0; s < 6; s++) sa.powcut[s] = p.g.powcut[s];
    sa.hit_keys = p.d_hkeys[slot];
    sa.hit_pows = p.d_hpows[slot];
    sa.nhits = p.d_counts + 2 * slot;
    sa.maxhits = (unsigned long long)p.hitcap;
    sa.sums_out = d_sums;
    sa.startr = p.d_startr;
    sa.hr_tab = p.d_hr[set];
    sa.int_rlo = p.int_rlo;
    sa.order = p.sum_order;
    sa.rpt = p.sum_rpt;
    Scope s(tm, 2, 0, p.sum_stream());
    PZL_HIPCHK(pzl_launch_harmsum(&sa, p.sum_stream()));
    return 0;
}

int sync_all(Plan &p);

// Drop the captured correlation graphs (they hold the variants chosen at
// capture time): the next full batch of each set captures anew.
int drop_graphs(Plan &p)
{
    int rc = sync_all(p);
    if (rc) return rc;
    for (int k = 0; k < Plan::kSets; k++)
        if (p.gexec[k]) {
            (void)hipGraphExecDestroy(p.gexec[k]);
            p.gexec[k] = nullptr;
        }
    return 0;
}

// The correlations of a full batch as one graph launch: the graph of a
// buffer set (every correlation launch of a batch; they do not depend on the
// batch's rsteps) is built once and replayed for every later full batch in
// that set, after the batch's segments, which are launched as they are (ahead
// on their own stream when overlapped, beside the previous batch's
// correlation); the first batch of each set runs eagerly (it also builds what
// the kernels read once per device).  The graph is built from the
// recorded launches with hipGraphAddKernelNode, in launch order, each node
// depending on the one before: no stream is captured, so HIP calls made
// meanwhile by other threads are unaffected.  A graph the runtime refuses is
// dropped (the batch has run eagerly) and a later batch of the set builds
// again; after kMaxGraphFailures the plan launches eagerly for good.
int build_batch_graph(Plan &p, int set, hipGraphExec_t *out)
{
    std::vector<PzlKernelRec> recs;
    pzl_recorder = &recs;
    const int rc = launch_corrs(p, p.B, set, nullptr, false);
    pzl_recorder = nullptr;
    if (rc) return rc;
    hipGraph_t g = nullptr;
    PZL_HIPCHK(hipGraphCreate(&g, 0));
    hipGraphNode_t prev = nullptr;
    hipError_t e = hipSuccess;
    for (auto &r : recs) {
        void *args[1] = {r.arg.data()};
        hipKernelNodeParams np{};
        np.func = const_cast<void *>(r.func);
        np.gridDim = r.grid;
        np.blockDim = r.block;
        np.sharedMemBytes = (unsigned)r.smem;
        np.kernelParams = args;
        hipGraphNode_t node = nullptr;
        e = hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &np);
        if (e != hipSuccess) break;
        prev = node;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(PZL_ERR_HIP, "batch graph: %s", hipGetErrorString(e));
    }
    return 0;
}

int launch_corr_graph(Plan &p, long long r0, int set, Timers *tm, bool seg_ahead)
{
    if (!p.gexec[set]) {
        int rc = launch_corr_part(p, r0, p.B, set, tm, nullptr, seg_ahead);
        if (rc) return rc;
        if (build_batch_graph(p, set, &p.gexec[set])) {
            (void)hipGetLastError();
            pzl::g_err.clear();
            if (++p.graph_failures >= Plan::kMaxGraphFailures) p.use_graph = false;
        }
        return 0;
    }
    int rc = launch_corr_part(p, r0, p.B, set, tm, nullptr, seg_ahead, true);
    if (rc) return rc;
    Scope s(tm, 3, 0, p.st);
    PZL_HIPCHK(hipGraphLaunch(p.gexec[set], p.st));
    return 0;
}

// Rsteps [r0, r0 + nb) in buffer set 'set', hits into slot 'slot'.  The
// correlation (stream st) waits until the previous harmonic sum of the set
// has read its planes; the harmonic sum (stream st3) waits for this set's
// correlation, so the correlation of one batch overlaps the sum of the
// previous one.
int launch_batch(Plan &p, long long r0, int nb, int set, int slot, float *d_sums, Timers *tm, bool with_sum = true,
                 bool graph = false, bool seg_ahead = false)
{
    // with the segments ahead, the correlation stream waits for them, and
    // they for the set's last sum (the wait below would repeat theirs)
    if (!seg_ahead) PZL_HIPCHK(hipStreamWaitEvent(p.st, p.ev_sum[set], 0));
    const int rc = (graph && nb == p.B) ? launch_corr_graph(p, r0, set, tm, seg_ahead)
                                        : launch_corr_part(p, r0, nb, set, tm, nullptr, seg_ahead);
    if (rc) return rc;
    PZL_HIPCHK(hipEventRecord(p.ev_corr[set], p.st));
    if (!with_sum) {
        PZL_HIPCHK(hipEventRecord(p.ev_sum[set], p.st));
        return 0;
    }
    PZL_HIPCHK(hipStreamWaitEvent(p.sum_stream(), p.ev_corr[set], 0));
    const int rs = launch_sum_part(p, r0, nb, set, slot, d_sums, tm);
    if (rs) return rs;
    PZL_HIPCHK(hipEventRecord(p.ev_sum[set], p.sum_stream()));
    return 0;
}

// Wait for every stream of the plan.
int sync_all(Plan &p)
{
    PZL_HIPCHK(hipStreamSynchronize(p.st4));
    PZL_HIPCHK(hipStreamSynchronize(p.st));
    PZL_HIPCHK(hipStreamSynchronize(p.st3));
    PZL_HIPCHK(hipStreamSynchronize(p.st2));
    PZL_HIPCHK(hipStreamSynchronize(p.st5));
    return 0;
}

// Candidate insertion runs on a thread of its own so that the launching
// thread keeps the GPU queue full.  Jobs are the ordered hits of whole
// batches, applied in batch order (the scan order of the reference's
// insertion); each sits in a pinned buffer of the plan's pool together with
// the event of the copy that fills it, which the insertion thread waits for.
struct SearchState {
    Plan *p = nullptr;
    CandList cl;
    double ms_host = 0;  // insertion thread busy time
    double ms_wait_pool = 0;   // launching thread waiting for a free pinned record buffer
    double ms_wait_count = 0;  // launching thread waiting for a batch's hit count (its harmonic sum)
    long long nhits = 0;
    int err = 0;         // first failure seen by the insertion thread
    std::string err_msg;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::pair<int, long long>> jobs;  // (pool buffer, records)
    bool stop = false;
    std::thread th;

    explicit SearchState(Plan *plan) : p(plan) { th = std::thread([this] { run(); }); }
    ~SearchState() { finish(); }
    void finish()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        if (th.joinable()) th.join();
    }
    // Records travel to the insertion thread in chunks of p->hit_chunk records.
    // Every pool buffer holds one chunk; it is made once (from the
    // process-wide cache, else pinned) and reused, so that a plan stops
    // pinning or freeing host memory once its pool has grown: pinning waits
    // for the device and takes milliseconds, and a launching thread doing it
    // in the middle of a search leaves the GPU without work.  A free buffer,
    // else a new one while the pool stays within its limit, else wait for the
    // insertion thread to release one.
    int acquire()
    {
        const long long chunk = p->hit_chunk;
        const long long chunk_bytes = (long long)sizeof(PzlHostHit) * chunk;
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            if (err) return -1;
            int nbusy = 0;
            for (int k = 0; k < (int)p->hpool.size(); k++) {
                auto &hb = p->hpool[k];
                if (hb.busy) {
                    nbusy++;
                } else if (hb.ptr && hb.cap >= chunk) {
                    hb.busy = true;
                    return k;
                }
            }
            if (p->hpool_bytes + chunk_bytes > p->hpool_limit) {
                if (nbusy > 0) {
                    const auto w0 = std::chrono::steady_clock::now();
                    cv.wait(lk);
                    ms_wait_pool +=
                        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
                    continue;
                }
                // nothing in flight: the idle buffers are too small for the
                // chunk (it was made larger); back to the process-wide cache
                for (auto &hb : p->hpool)
                    if (hb.ptr) {
                        pin_put(hb.ptr, hb.cap);
                        p->hpool_bytes -= (long long)sizeof(PzlHostHit) * hb.cap;
                        hb.ptr = nullptr;
                        hb.cap = 0;
                    }
            }
            int k = -1;
            for (int q = 0; q < (int)p->hpool.size(); q++)
                if (!p->hpool[q].busy && !p->hpool[q].ptr) k = q;
            if (k < 0) {
                p->hpool.emplace_back();
                k = (int)p->hpool.size() - 1;
                if (hipEventCreateWithFlags(&p->hpool[k].ev, hipEventDisableTiming) != hipSuccess) {
                    p->hpool.pop_back();
                    return -1;
                }
            }
            auto &hb = p->hpool[k];
            long long cap = chunk;
            hb.ptr = pin_take(chunk, &cap);
            if (!hb.ptr && pinned_alloc((void **)&hb.ptr, (size_t)chunk_bytes) != hipSuccess) {
                hb.ptr = nullptr;
                return -1;
            }
            hb.cap = cap;
            p->hpool_bytes += (long long)sizeof(PzlHostHit) * cap;
            hb.busy = true;
            return k;
        }
    }
    void submit(int buf, long long n)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            jobs.emplace_back(buf, n);
        }
        cv.notify_all();
    }
    void run()
    {
        for (;;) {
            std::pair<int, long long> j;
            Plan::HostBuf hb;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !jobs.empty(); });
                if (jobs.empty()) return;
                j = jobs.front();
                jobs.pop_front();
                hb = p->hpool[j.first];  // the pool only grows while jobs are pending
            }
            const hipError_t e = hipEventSynchronize(hb.ev);
            if (e == hipSuccess) {
                insert_hits(hb.ptr, j.second);
            } else {
                std::lock_guard<std::mutex> lk(mu);
                if (!err) {
                    err = PZL_ERR_HIP;
                    err_msg = hipGetErrorString(e);
                }
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                p->hpool[j.first].busy = false;
            }
            cv.notify_all();
        }
    }
    void insert_hits(const PzlHostHit *recs, long long n)
    {
        const auto t0 = std::chrono::steady_clock::now();
        const double fzlo = p->fzlo, fwlo = p->fwlo;
        for (long long i = 0; i < n; i++) {
            const PzlHostHit &h = recs[i];
            cl.insert(h.rr, h.sigma, [&]() {
                const int numharm = 1 << h.stage;
                PzlCand c;
                c.power = h.power;
                c.sigma = h.sigma;
                c.numharm = numharm;
                c.r = h.rr;
                c.z = (fzlo + h.zi * (double)PZL_DZ) / (double)numharm;
                c.w = (fwlo + h.wi * (double)PZL_DW) / (double)numharm;
                return c;
            });
        }
        if (p->last_hits_complete) {
            if ((long long)p->last_hits.size() + n <= p->keep_hits_limit) {
                p->last_hits.insert(p->last_hits.end(), recs, recs + n);
            } else {
                p->last_hits.clear();
                p->last_hits_complete = false;
            }
        }
        ms_host += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

// Order the hits of one slot on the GPU, release the slot, bring the hits over
// and queue them for the insertion thread (the caller has waited for the
// slot's count).  Nothing here waits for the GPU: the copy's event travels
// with the job.
int reset_slot(Plan &p, int slot, hipStream_t st);

int apply_slot(Plan &p, int slot, long long n, SearchState &ss)
{
    if (n <= 0) return reset_slot(p, slot, p.st2);
    PZL_HIPCHK(pzl_sort_hits(p.d_sort_temp, &p.sort_temp_bytes, p.d_hkeys[slot], p.d_hpows[slot], p.d_skeys,
                             p.d_spows, n, p.key_bits, p.st2));
    // the sort has read the slot: a later batch's harmonic sum may refill it
    // while the records are prepared and copied
    int rc = reset_slot(p, slot, p.st2);
    if (rc) return rc;
    // The records go to host memory on a stream of their own (st5), so that
    // the next slot's sort, which releases that slot, does not queue behind
    // the copies (hundreds of MB for a batch with millions of hits); only the
    // next preparation waits for them, as it rewrites the record buffer.
    PZL_HIPCHK(hipStreamWaitEvent(p.st2, p.ev_copied, 0));
    PZL_HIPCHK(pzl_launch_hitprep(p.d_skeys, p.d_spows, n, p.d_startr, p.g.corr_uselen, p.fnumz, p.fnumw,
                                  p.g.numharmstages, p.numindep_d, p.d_recs, p.st2));
    PZL_HIPCHK(hipEventRecord(p.ev_prep, p.st2));
    PZL_HIPCHK(hipStreamWaitEvent(p.st5, p.ev_prep, 0));
    // the records in chunks of pool buffers, in order (the insertion thread
    // applies the jobs first in, first out)
    for (long long off = 0; off < n; off += p.hit_chunk) {
        const long long m = std::min(p.hit_chunk, n - off);
        const int buf = ss.acquire();
        if (buf < 0)
            return fail(ss.err ? ss.err : PZL_ERR_HIP, "no pinned buffer for %lld hit records%s%s", m,
                        ss.err_msg.empty() ? "" : ": ", ss.err_msg.c_str());
        PZL_PINNED_CHECK(p.hpool[buf].ptr, sizeof(PzlHostHit) * m);
        PZL_HIPCHK(hipMemcpyAsync(p.hpool[buf].ptr, p.d_recs + off, sizeof(PzlHostHit) * m, hipMemcpyDeviceToHost,
                                  p.st5));
        PZL_HIPCHK(hipEventRecord(p.hpool[buf].ev, p.st5));
        ss.submit(buf, m);
    }
    PZL_HIPCHK(hipEventRecord(p.ev_copied, p.st5));
    ss.nhits += n;
    return 0;
}

int reset_slot(Plan &p, int slot, hipStream_t st)
{
    PZL_HIPCHK(hipMemsetAsync(p.d_counts + 2 * slot, 0, 2 * sizeof(unsigned long long), st));
    PZL_HIPCHK(hipEventRecord(p.ev_free[slot], st));
    return 0;
}

// Wait for the batch in 'slot', order and apply its hits.  A batch that
// overflowed the slot is searched again one rstep at a time (one rstep never
// exceeds a slot: hitcap >= cells per rstep x stages).
int finalize(Plan &p, int slot, long long r0, int nb, SearchState &ss, Timers *tm)
{
    const auto w0 = std::chrono::steady_clock::now();
    PZL_HIPCHK(hipEventSynchronize(p.ev_count[slot]));
    ss.ms_wait_count += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    const unsigned long long n = p.h_counts[2 * slot];
    if ((long long)n <= p.hitcap) return apply_slot(p, slot, (long long)n, ss);  // resets the slot
    // overflow: let every batch in flight finish, then search this batch
    // again one rstep at a time
    int rc = sync_all(p);
    if (rc) return rc;
    for (int b = 0; b < nb; b++) {
        // the previous rstep's sort (stream st2) reads and rewrites the slot's
        // buffers: the harmonic sum refills them only after it
        PZL_HIPCHK(hipStreamWaitEvent(p.sum_stream(), p.ev_free[slot], 0));
        rc = reset_slot(p, slot, p.sum_stream());
        if (rc) return rc;
        rc = launch_batch(p, r0 + b, 1, 0, slot, nullptr, tm);
        if (rc) return rc;
        PZL_PINNED_CHECK(p.h_counts + 2 * slot, 2 * sizeof(unsigned long long));
        PZL_HIPCHK(hipMemcpyAsync(p.h_counts + 2 * slot, p.d_counts + 2 * slot, 2 * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, p.sum_stream()));
        rc = sync_all(p);
        if (rc) return rc;
        const unsigned long long n1 = p.h_counts[2 * slot];
        if ((long long)n1 > p.hitcap)
            return fail(PZL_ERR_OVERFLOW, "hit buffer overflow: %llu hits in one rstep", n1);
        rc = apply_slot(p, slot, (long long)n1, ss);
        if (rc) return rc;
    }
    return 0;
}

double plane_alg_bytes(const Plan &p)
{
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

int pzl_bank_cache_set_budget(long long bytes)
{
    if (bytes < 0) return fail(PZL_ERR_ARG, "negative budget");
    std::lock_guard<std::mutex> lk(g_bank_mu);
    g_bank_budget = bytes;
    evict_banks_locked(0);
    return 0;
}

int pzl_bank_cache_clear(void)
{
    std::lock_guard<std::mutex> lk(g_bank_mu);
    const long long keep = g_bank_budget;
    g_bank_budget = 0;
    evict_banks_locked(0);
    g_bank_budget = keep;
    return 0;
}

int pzl_plan_bank_built(void *plan) { return plan ? (int)((Plan *)plan)->bank_built : 0; }

int pzl_device_count(int *count)
{
    PZL_HIPCHK(hipGetDeviceCount(count));
    return 0;
}

int pzl_device_memory(int device, long long *free_bytes, long long *total_bytes)
{
    if (!free_bytes || !total_bytes) return fail(PZL_ERR_ARG, "null argument");
    size_t f = 0, t = 0;
    PZL_HIPCHK(hipSetDevice(device));
    PZL_HIPCHK(hipMemGetInfo(&f, &t));
    *free_bytes = (long long)f;
    *total_bytes = (long long)t;
    return 0;
}

int pzl_geometry_from_config(const PzlConfig *cfg, PzlGeometry *g)
{
    if (!cfg || !g) return fail(PZL_ERR_ARG, "null argument");
    std::vector<PzlHarm> harms;
    int rc = compute_geometry(*cfg, *g, &harms, nullptr);
    if (rc) return rc;
    g->device_bytes = plan_device_bytes(*g, plan_sizes(*cfg, *g, harms), harms, -1);
    return 0;
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
        if (!pzl_corr_supported(H.fftlen)) {
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
    for (int s = 0; s < p->g.numharmstages; s++) p->numindep_d[s] = (double)p->g.numindep[s];
    const PlanSizes sizes = plan_sizes(*cfg, p->g, p->harms);
    p->B = sizes.B;
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
    // The harmonic sum dereferences these maps unchecked: prove they stay in
    // their planes for every rstep (r maps are monotone in r, so the ends
    // bound them) and that the integer r map equals the double formula.
    for (double s : p->startr)
        if (s != floor(s)) p->int_rlo = 0;
    for (int h = 1; h < nh; h++) {
        const PzlHarm &H = p->harms[h];
        for (int jj = 0; jj < F.numz; jj++)
            if (p->zinds[(size_t)h * F.numz + jj] >= H.numz) {
                delete p;
                return fail(PZL_ERR_UNSUPPORTED, "z map leaves harmonic %d/%d", H.harmnum, H.numharm);
            }
        for (int ii = 0; ii < F.numw; ii++)
            if (p->winds[(size_t)h * F.numw + ii] < 0 || p->winds[(size_t)h * F.numw + ii] >= H.numw) {
                delete p;
                return fail(PZL_ERR_UNSUPPORTED, "w map leaves harmonic %d/%d", H.harmnum, H.numharm);
            }
        const int rstep_bins = (int)(p->g.corr_uselen * PZL_DR);
        for (double s : p->startr) {
            const PzlHarmR hr = pzl_harm_r(H.hf, 0, s, s + rstep_bins - PZL_DR, p->g.corr_uselen);
            const int lo = pzl_rind(H.hf, s, hr.rlo, 0), hi = pzl_rind(H.hf, s, hr.rlo, p->g.corr_uselen - 1);
            bool ok = lo >= 0 && hi < H.rs_stride && hi < hr.numrs;
            if (ok && p->int_rlo) {
                // integer round-half-even of (2 s + ii) * harmnum / 2^stage
                for (int ii : {0, 1, p->g.corr_uselen / 2, p->g.corr_uselen - 1}) {
                    const long long prod = ((long long)(2.0 * s) + ii) * H.harmnum;
                    long long q = prod >> H.stage;
                    const long long rem = prod & ((1LL << H.stage) - 1), half = 1LL << (H.stage - 1);
                    if (rem > half || (rem == half && (q & 1))) q++;
                    ok = ok && (q - 2 * hr.rlo == pzl_rind(H.hf, s, hr.rlo, ii));
                }
            }
            if (!ok) {
                delete p;
                return fail(PZL_ERR_UNSUPPORTED, "r map leaves harmonic %d/%d at startr %.1f", H.harmnum,
                            H.numharm, s);
            }
        }
    }
    std::map<int, std::vector<int>> cls;
    for (int h = 0; h < nh; h++) cls[p->harms[h].fftlen].push_back(h);
    for (auto it = cls.rbegin(); it != cls.rend(); ++it) {
        SizeClass c;
        c.fftlen = it->first;
        c.harms = it->second;
        for (int h : c.harms) c.nrows.push_back(p->harms[h].nrows);
        c.variant = tuned_variant(*p, c);
        p->classes.push_back(c);
    }
    {
        const int t = tuned_group(*p), ng = pzl_corr_num_groups();
        p->side_group = t < 0 ? -1 : t % ng;
        p->group_all = t >= ng;
    }
    p->hitcap = sizes.hitcap;
    const long long maxkey = std::max<long long>(p->g.nrsteps, 1) * p->g.numharmstages * F.numw * F.numz *
                             (long long)p->g.corr_uselen;
    p->key_bits = 1;
    while (p->key_bits < 63 && (1LL << p->key_bits) < maxkey) p->key_bits++;

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
    PCHK(hipStreamCreateWithFlags(&p->st2, hipStreamNonBlocking));
    PCHK(hipStreamCreateWithFlags(&p->st3, hipStreamNonBlocking));
    PCHK(hipStreamCreateWithFlags(&p->st4, hipStreamNonBlocking));
    PCHK(hipStreamCreateWithFlags(&p->st5, hipStreamNonBlocking));
    PCHK(hipEventCreateWithFlags(&p->ev_prep, hipEventDisableTiming));
    PCHK(hipEventCreateWithFlags(&p->ev_copied, hipEventDisableTiming));
    p->st_all[0] = p->st;
    p->st_all[1] = p->st3;
    p->st_all[2] = p->st4;
    p->nslots = sizes.nslots;
    for (int s = 0; s < p->nslots; s++) {
        PCHK(hipEventCreateWithFlags(&p->ev_count[s], hipEventDisableTiming));
        PCHK(hipEventCreateWithFlags(&p->ev_free[s], hipEventDisableTiming));
    }
    p->seg_stride = sizes.seg_stride;
    p->plane_stride = p->g.plane_bytes_per_rstep / 4;
    if (p->g.plane_bytes_per_rstep >= (1LL << 31))
        return cleanup(fail(PZL_ERR_UNSUPPORTED, "the planes of one rstep need %lld bytes (at most 2 GiB)",
                            p->g.plane_bytes_per_rstep));
    p->sort_temp_bytes = 0;
    PCHK(pzl_sort_hits(nullptr, &p->sort_temp_bytes, nullptr, nullptr, nullptr, nullptr, p->hitcap, p->key_bits,
                       p->st2));
    // every buffer in one allocation, in the order plan_arena_items lists them
    const std::vector<long long> items = plan_arena_items(p->g, sizes, p->harms, (long long)p->sort_temp_bytes);
    PCHK(plan_malloc(p, &p->d_arena, (size_t)arena_bytes(items)));
    size_t item = 0;
    char *cur = p->d_arena;
    auto carve = [&]() {
        char *r = cur;
        cur += (items[item] + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
        item++;
        return r;
    };
    p->d_harms = (PzlHarm *)carve();
    int *cls_idx = (int *)carve();
    p->d_startr = (double *)carve();
    p->d_zinds = (uint16_t *)carve();
    p->d_winds = (int *)carve();
    p->d_zoff = (int *)carve();
    for (int k = 0; k < Plan::kSets; k++) {
        p->d_segs[k] = (float2 *)carve();
        p->d_rinds[k] = (uint16_t *)carve();
        p->d_hr[k] = (PzlHrTab *)carve();
        p->d_planes[k] = (float *)carve();
    }
    for (int s = 0; s < p->nslots; s++) {
        p->d_hkeys[s] = (long long *)carve();
        p->d_hpows[s] = (float *)carve();
    }
    p->d_counts = (unsigned long long *)carve();
    p->d_skeys = (long long *)carve();
    p->d_spows = (float *)carve();
    p->d_recs = (PzlHostHit *)carve();
    p->d_sort_temp = carve();
    if (item != items.size()) return cleanup(fail(PZL_ERR_ARG, "plan arena layout mismatch"));
    PCHK(hipMemcpy(p->d_harms, p->harms.data(), sizeof(PzlHarm) * nh, hipMemcpyHostToDevice));
    for (auto &c : p->classes) {
        c.d_harms = cls_idx;
        PCHK(hipMemcpy(c.d_harms, c.harms.data(), sizeof(int) * c.harms.size(), hipMemcpyHostToDevice));
        cls_idx += c.harms.size();
    }
    if (!p->startr.empty())
        PCHK(hipMemcpy(p->d_startr, p->startr.data(), sizeof(double) * p->startr.size(), hipMemcpyHostToDevice));
    PCHK(hipMemcpy(p->d_zinds, p->zinds.data(), sizeof(uint16_t) * p->zinds.size(), hipMemcpyHostToDevice));
    PCHK(hipMemcpy(p->d_winds, p->winds.data(), sizeof(int) * p->winds.size(), hipMemcpyHostToDevice));
    {
        // byte offset of every subharmonic row the harmonic sum reads for a
        // fundamental (w, z): [w][z][harmonic - 1]
        const int nq = std::max(nh - 1, 1);
        std::vector<int> zoff((size_t)F.numw * F.numz * nq, 0);
        for (int ii = 0; ii < F.numw; ii++)
            for (int jj = 0; jj < F.numz; jj++)
                for (int h = 1; h < nh; h++) {
                    const PzlHarm &H = p->harms[h];
                    const long long off = 4 * (H.plane_off + ((long long)p->winds[(size_t)h * F.numw + ii] * H.numz +
                                                              p->zinds[(size_t)h * F.numz + jj]) *
                                                                 H.rs_stride);
                    zoff[((size_t)ii * F.numz + jj) * nq + (h - 1)] = (int)off;
                }
        PCHK(hipMemcpy(p->d_zoff, zoff.data(), sizeof(int) * zoff.size(), hipMemcpyHostToDevice));
    }
    // The buffer-set events order the plan's streams on this device and
    // nothing on the host waits for them: their records release to device
    // scope, without the system-scope fence of a host-visible event.
    constexpr unsigned kDevEv = hipEventDisableTiming | hipEventDisableSystemFence;
    for (int k = 0; k < Plan::kSets; k++) {
        PCHK(hipEventCreateWithFlags(&p->ev_corr[k], kDevEv));
        PCHK(hipEventCreateWithFlags(&p->ev_sum[k], kDevEv));
        PCHK(hipEventCreateWithFlags(&p->ev_seg[k], kDevEv));
        PCHK(hipEventRecord(p->ev_seg[k], p->st));
        PCHK(hipEventCreateWithFlags(&p->ev_side[k], kDevEv));
        PCHK(hipEventRecord(p->ev_side[k], p->st));
        // recorded once so that the first waits on them are satisfied
        PCHK(hipEventRecord(p->ev_corr[k], p->st));
        PCHK(hipEventRecord(p->ev_sum[k], p->st));
    }
    PCHK(hipMemset(p->d_counts, 0, 2 * Plan::kSlots * sizeof(unsigned long long)));
    PCHK(pinned_alloc((void **)&p->h_counts, 2 * Plan::kSlots * sizeof(unsigned long long)));
    // The kernel bank: built once per shape and kept resident.
    {
        bool built = false;
        int brc = acquire_bank(p, &built);
        if (brc) return cleanup(brc);
        p->bank_built = built;
    }
    p->g.device_bytes = plan_device_bytes(p->g, sizes, p->harms, (long long)p->sort_temp_bytes);
#undef PCHK
    *plan_out = p;
    return 0;
}

void pzl_plan_destroy(void *plan) { free_plan((Plan *)plan); }

int pzl_plan_device_bytes(void *plan, long long *held)
{
    Plan *p = (Plan *)plan;
    if (!p || !held) return fail(PZL_ERR_ARG, "null argument");
    *held = p->dev_bytes + (p->d_bank ? p->g.bank_bytes : 0);
    return 0;
}

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
    PZL_HIPCHK(hipSetDevice(p->device));
    if (p->d_fft_own && p->fft_bins != numbins) {
        PZL_HIPCHK(hipFree(p->d_fft_own));
        p->d_fft_own = nullptr;
        p->dev_bytes -= (long long)sizeof(float2) * p->fft_bins;
    }
    if (!p->d_fft_own) PZL_HIPCHK(plan_malloc(p, &p->d_fft_own, sizeof(float2) * numbins));
    PZL_HIPCHK(hipMemcpy(p->d_fft_own, host_fft, sizeof(float2) * numbins, hipMemcpyHostToDevice));
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
    PZL_HIPCHK(hipSetDevice(p->device));
    const long long nr = p->g.nrsteps;
    if (rstep_lo < 0) rstep_lo = 0;
    if (rstep_hi < 0 || rstep_hi > nr) rstep_hi = nr;
    Timers tm;
    tm.on = (stats != nullptr);
    p->last_hits.clear();
    p->last_hits_complete = true;
    SearchState ss(p);
    hipEvent_t ev_start = nullptr, ev_end = nullptr;
    if (tm.on) {
        ev_start = tm.get();
        PZL_HIPCHK(hipEventRecord(ev_start, p->st));
    }
    PZL_HIPCHK(hipMemsetAsync(p->d_counts, 0, 2 * Plan::kSlots * sizeof(unsigned long long), p->st));
    // Batches in flight, oldest first; the GPU runs up to nslots batches
    // ahead of the host's candidate insertion.
    struct Pending {
        int slot, nb;
        long long r0;
    };
    std::vector<Pending> pend;
    size_t head = 0;
    int i = 0;
    // per-kernel events on every kSampleEvery-th batch only: the averages stay
    // representative and the event packets do not perturb the timed search
    constexpr int kSampleEvery = 8;
    long long sampled = 0;
    bool timed_out = false;
    for (long long r0 = rstep_lo; r0 < rstep_hi; r0 += p->B, i++) {
        if (p->has_deadline && std::chrono::steady_clock::now() >= p->deadline) {
            timed_out = true;  // launch nothing more; the batches in flight drain below
            break;
        }
        const int nb = (int)std::min<long long>(p->B, rstep_hi - r0);
        const int slot = i % p->nslots;
        tm.active = (i % kSampleEvery) == 0;
        sampled += tm.active;
        if (pend.size() - head == (size_t)p->nslots) {
            const Pending q = pend[head++];
            int rc = finalize(*p, q.slot, q.r0, q.nb, ss, &tm);
            if (rc) return rc;
        }
        // the harmonic sum writes into the slot once the host has drained it
        if (i >= p->nslots) PZL_HIPCHK(hipStreamWaitEvent(p->sum_stream(), p->ev_free[slot], 0));
        // (segments ahead on their own stream read only what is constant during
        // the search: the spectrum and the startr table)
        int rc = launch_batch(*p, r0, nb, i % Plan::kSets, slot, nullptr, &tm, true, p->use_graph,
                              p->overlap_sum && p->side_classes);
        if (rc) return rc;
        PZL_PINNED_CHECK(p->h_counts + 2 * slot, 2 * sizeof(unsigned long long));
        PZL_HIPCHK(hipMemcpyAsync(p->h_counts + 2 * slot, p->d_counts + 2 * slot, 2 * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, p->sum_stream()));
        PZL_HIPCHK(hipEventRecord(p->ev_count[slot], p->sum_stream()));
        pend.push_back({slot, nb, r0});
        // apply whatever the GPU has already finished
        while (head + 1 < pend.size() && hipEventQuery(p->ev_count[pend[head].slot]) == hipSuccess) {
            const Pending q = pend[head++];
            rc = finalize(*p, q.slot, q.r0, q.nb, ss, &tm);
            if (rc) return rc;
        }
    }
    tm.active = true;
    if (tm.on) {
        ev_end = tm.get();
        PZL_HIPCHK(hipEventRecord(ev_end, p->sum_stream()));  // the last harmonic sum follows the last correlation
    }
    while (head < pend.size()) {
        const Pending q = pend[head++];
        int rc = finalize(*p, q.slot, q.r0, q.nb, ss, &tm);
        if (rc) return rc;
    }
    {
        int rc = sync_all(*p);
        if (rc) return rc;
    }
    ss.finish();
    if (ss.err) return fail(ss.err, "hit records did not arrive: %s", ss.err_msg.c_str());
    if (timed_out) {
        p->last_cands.clear();
        p->last_hits_complete = false;
        return fail(PZL_ERR_TIMEOUT, "search deadline passed after %d of %lld batches", i,
                    (rstep_hi - rstep_lo + p->B - 1) / p->B);
    }
    p->last_cands = ss.cl.items();
    const auto &items = p->last_cands;
    if (ncands) *ncands = (long long)items.size();
    if (cands)
        for (long long k = 0; k < (long long)items.size() && k < maxcands; k++) cands[k] = items[k];
    const auto t2 = std::chrono::steady_clock::now();
    if (stats) {
        memset(stats, 0, sizeof *stats);
        float ms = 0;
        std::map<int, double> per_class;
        std::map<int, long long> per_class_n;
        for (auto &r : tm.recs) {
            (void)hipEventElapsedTime(&ms, r.a, r.b);
            if (r.kind == 0) {
                stats->ms_segments += ms;
            } else if (r.kind == 1) {
                stats->ms_corr += ms;
                if (r.fftlen == 0) {
                    stats->ms_corr_group += ms;  // grouped launch of the shorter lengths
                } else {
                    per_class[r.fftlen] += ms;
                    per_class_n[r.fftlen]++;
                }
                stats->corr_launches++;
            } else if (r.kind == 2) {
                stats->ms_sum += ms;
            } else {
                stats->ms_graph += ms;
                stats->graph_launches++;
            }
        }
        (void)hipEventElapsedTime(&ms, ev_start, ev_end);
        stats->ms_gpu = ms;
        stats->ms_wall = std::chrono::duration<double, std::milli>(t2 - t0).count();
        stats->ms_host = ss.ms_host;
        stats->ms_wait_pool = ss.ms_wait_pool;
        stats->ms_wait_count = ss.ms_wait_count;
        stats->nhits = ss.nhits;
        stats->nrsteps = rstep_hi - rstep_lo;
        stats->cells = (double)stats->nrsteps * p->g.corr_uselen * p->g.numz * std::max(p->g.numw, 1);
        double seg_bytes = 0;
        for (auto &H : p->harms) seg_bytes += H.fftlen / 2 * 8.0;
        const double planes = plane_alg_bytes(*p);
        stats->bytes_alg = stats->nrsteps * (planes + seg_bytes) + (double)p->g.bank_bytes;
        stats->bytes_alg_corr = stats->nrsteps * (planes / 2 + seg_bytes) + (double)p->g.bank_bytes;
        // sampled batches -> estimates for the whole search (launch counts exact)
        const double scale = sampled > 0 ? (double)i / (double)sampled : 0.0;
        stats->ms_segments *= scale;
        stats->ms_corr *= scale;
        stats->ms_sum *= scale;
        stats->ms_graph *= scale;
        stats->ms_corr_group *= scale;
        for (auto &kv : per_class) kv.second *= scale;
        for (auto &kv : per_class_n) kv.second = i;
        const bool grouped = p->side_group >= 0 && p->classes.size() > 1;
        stats->group_launches = grouped ? i : 0;
        stats->corr_launches = (long long)i * (grouped ? (p->group_all ? 1LL : 2LL) : (long long)p->classes.size());
        stats->graph_launches = (long long)std::llround(stats->graph_launches * scale);
        for (auto &kv : per_class) stats->ms_corr_max_class = std::max(stats->ms_corr_max_class, kv.second);
        stats->nclass = 0;
        for (auto &c : p->classes) {
            if (stats->nclass >= 16) break;
            const int k = stats->nclass++;
            stats->class_fftlen[k] = c.fftlen;
            stats->class_ms[k] = per_class.count(c.fftlen) ? per_class[c.fftlen] : 0.0;
            const long long launches = per_class_n.count(c.fftlen) ? per_class_n[c.fftlen] : 0;
            stats->class_launches[k] = launches;
            // per rstep: its plane rows written + segments read; the kernel rows
            // once per search (they are re-read by every launch: class_bank_bytes)
            double per_rstep = 0, bank = 0, flops = 0;
            for (int h : c.harms) {
                const PzlHarm &H = p->harms[h];
                const double n = H.fftlen;
                per_rstep += (double)H.nrows * H.max_numrs * 4.0 + H.fftlen / 2 * 8.0;
                bank += (double)H.nrows * H.fftlen * 8.0;
                flops += (double)H.nrows * (5.0 * n * std::log2(n) + 6.0 * n + 3.0 * H.max_numrs);
            }
            const double nrs = (double)(rstep_hi - rstep_lo);
            stats->class_bytes[k] = per_rstep * nrs + bank;
            stats->class_bank_bytes[k] = bank;
            stats->class_flops[k] = flops * nrs;
        }
    }
    return 0;
}

int pzl_autotune(void *plan, int reps, char *report, int report_len)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    if (!p->d_fft) return fail(PZL_ERR_NOSPECTRUM, "no spectrum set");
    PZL_HIPCHK(hipSetDevice(p->device));
    if (reps < 1) reps = 1;
    {
        // the variants may change below (drops the graphs), and every kernel is
        // first timed on the whole chip
        int grc = p->cu_split ? pzl_plan_set_cu_split(plan, 0, 0) : drop_graphs(*p);
        if (grc) return grc;
    }
    const long long nr = p->g.nrsteps;
    const long long r0 = std::max<long long>(0, std::min<long long>(nr - p->B, nr / 2));
    const int nb = (int)std::min<long long>(p->B, nr - r0);
    if (nb < 1) return 0;
    // the segments of one batch from the middle of the band
    int rc = launch_batch(*p, r0, nb, 0, 0, nullptr, nullptr, false);
    if (rc) return rc;
    {
        int src = sync_all(*p);
        if (src) return src;
    }
    hipEvent_t e0, e1;
    PZL_HIPCHK(hipEventCreate(&e0));
    PZL_HIPCHK(hipEventCreate(&e1));
    std::string rep;
    char line[256];
    std::vector<std::pair<float, int>> lead_times;  // (ms, variant) of the longest FFT length
    for (auto &c : p->classes) {
        const PzlCorrArgs ca = corr_args(*p, c, nb, 0);
        const int nv = pzl_corr_num_variants(c.fftlen);
        int best = c.variant;
        float best_ms = 1e30f;
        for (int v = 0; v < nv; v++) {
            hipError_t e = pzl_launch_corr(c.fftlen, v, &ca, c.nrows.data(), p->st);
            if (e == hipSuccess) e = hipEventRecord(e0, p->st);
            for (int k = 0; k < reps && e == hipSuccess; k++)
                e = pzl_launch_corr(c.fftlen, v, &ca, c.nrows.data(), p->st);
            if (e == hipSuccess) e = hipEventRecord(e1, p->st);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                snprintf(line, sizeof line, "fftlen %d variant %d (%s): failed: %s\n", c.fftlen, v,
                         pzl_corr_variant_name(c.fftlen, v), hipGetErrorString(e));
                rep += line;
                continue;
            }
            ms /= reps;
            snprintf(line, sizeof line, "fftlen %d variant %d (%s): %.4f ms\n", c.fftlen, v,
                     pzl_corr_variant_name(c.fftlen, v), ms);
            rep += line;
            if (&c == &p->classes[0]) lead_times.emplace_back(ms, v);
            if (ms < best_ms) {
                best_ms = ms;
                best = v;
            }
        }
        c.variant = best;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[tune_key(*p, c)] = best;
    }
    // harmonic-sum workgroup order, on the planes the correlations just wrote
    std::vector<std::pair<float, int>> sum_times;  // (ms, form = order | rpt log2 << 1)
    {
        PZL_HIPCHK(hipMemsetAsync(p->d_counts, 0, 2 * sizeof(unsigned long long), p->st));
        float best_ms = 1e30f;
        int best = p->sum_order, best_rpt = p->sum_rpt;
        const bool ov = p->overlap_sum;
        // on p->st behind the correlations, or, with a CU split, on the summing
        // stream's own compute units after them
        p->overlap_sum = p->cu_split > 0;
        hipStream_t ts = p->sum_stream();
        if (ts != p->st) {
            PZL_HIPCHK(hipEventRecord(e0, p->st));
            PZL_HIPCHK(hipStreamWaitEvent(ts, e0, 0));
        }
        for (int form = 0; form < 6; form++) {
            const int order = form & 1, rpt = 1 << (form >> 1);
            p->sum_order = order;
            p->sum_rpt = rpt;
            int rc2 = launch_sum_part(*p, r0, nb, 0, 0, nullptr, nullptr);
            hipError_t e = rc2 ? hipErrorUnknown : hipEventRecord(e0, ts);
            for (int k = 0; k < reps && e == hipSuccess; k++)
                e = launch_sum_part(*p, r0, nb, 0, 0, nullptr, nullptr) ? hipErrorUnknown : hipSuccess;
            if (e == hipSuccess) e = hipEventRecord(e1, ts);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) continue;
            ms /= reps;
            snprintf(line, sizeof line, "harmonic sum order %d, %d r per thread: %.4f ms\n", order, rpt, ms);
            rep += line;
            sum_times.emplace_back(ms, form);
            if (ms < best_ms) {
                best_ms = ms;
                best = order;
                best_rpt = rpt;
            }
        }
        p->sum_order = best;
        p->sum_rpt = best_rpt;
        p->overlap_sum = ov;
        PZL_HIPCHK(hipStreamSynchronize(ts));
        PZL_HIPCHK(hipMemsetAsync(p->d_counts, 0, 2 * sizeof(unsigned long long), p->st));
        PZL_HIPCHK(hipStreamSynchronize(p->st));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // In the search the longest length's correlation shares the CUs with the
    // harmonic sum and the shorter lengths: the variants of it within 8% of
    // the fastest alone are timed on a band of the real pipeline (best of two
    // searches of 8 batches), and the fastest there is kept.
    std::sort(lead_times.begin(), lead_times.end());
    const long long band = 8LL * p->B;
    if (lead_times.size() >= 2 && nr >= band) {
        SizeClass &c = p->classes[0];
        const long long lo = (nr - band) / 2;
        float best_ms = 1e30f;
        int best = c.variant;
        for (size_t i = 0; i < lead_times.size() && i < 5; i++) {
            if (lead_times[i].first > 1.08f * lead_times[0].first) break;
            c.variant = lead_times[i].second;
            rc = drop_graphs(*p);  // captured graphs hold the variants of their capture
            if (rc) return rc;
            float ms = 1e30f;
            for (int k = 0; k < 2; k++) {
                long long n = 0;
                const auto t0 = std::chrono::steady_clock::now();
                rc = pzl_search(plan, lo, lo + band, nullptr, 0, &n, nullptr);
                if (rc) return rc;
                ms = std::min(ms, (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                                                  t0).count());
            }
            snprintf(line, sizeof line, "fftlen %d variant %d in a search band of %lld rsteps: %.3f ms\n", c.fftlen,
                     c.variant, band, ms);
            rep += line;
            if (ms < best_ms) {
                best_ms = ms;
                best = c.variant;
            }
        }
        c.variant = best;
        rc = drop_graphs(*p);
        if (rc) return rc;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[tune_key(*p, c)] = best;
    }
    // likewise the harmonic-sum forms within 8% of the fastest alone
    std::sort(sum_times.begin(), sum_times.end());
    if (sum_times.size() >= 2 && nr >= band && sum_times[1].first <= 1.08f * sum_times[0].first) {
        const long long lo = (nr - band) / 2;
        float best_ms = 1e30f;
        int best_order = p->sum_order, best_rpt = p->sum_rpt;
        for (size_t i = 0; i < sum_times.size() && i < 3; i++) {
            if (sum_times[i].first > 1.08f * sum_times[0].first) break;
            p->sum_order = sum_times[i].second & 1;
            p->sum_rpt = 1 << (sum_times[i].second >> 1);
            float ms = 1e30f;
            for (int k = 0; k < 2; k++) {
                long long n = 0;
                const auto t0 = std::chrono::steady_clock::now();
                rc = pzl_search(plan, lo, lo + band, nullptr, 0, &n, nullptr);
                if (rc) return rc;
                ms = std::min(ms, (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                                                  t0).count());
            }
            snprintf(line, sizeof line, "harmonic sum order %d, %d r per thread in a search band of %lld rsteps: %.3f ms\n",
                     p->sum_order, p->sum_rpt, band, ms);
            rep += line;
            if (ms < best_ms) {
                best_ms = ms;
                best_order = p->sum_order;
                best_rpt = p->sum_rpt;
            }
        }
        p->sum_order = best_order;
        p->sum_rpt = best_rpt;
    }
    // every length in one grouped launch, the shorter lengths in one, or one
    // launch per length: each form a group allows is timed on the same band
    // (best of three searches), in that order (the last group first: the
    // leanest), and a form replaces the one kept so far only when it is faster
    // by more than 2% (the forms differ by about that much, and the choice
    // should not follow the timing noise)
    if (p->classes.size() >= 2 && nr >= band) {
        const long long lo = (nr - band) / 2;
        const int ng = pzl_corr_num_groups();
        std::vector<int> forms;  // as g_tuned_group
        for (int all = 1; all >= 0; all--)
            for (int g = ng - 1; g >= 0; g--)
                if (group_covers(*p, g, all != 0)) forms.push_back(g + all * ng);
        forms.push_back(-1);
        if (forms.size() > 1) {
            float best_ms = 1e30f;
            int best = forms[0];
            for (int f : forms) {
                p->side_group = f < 0 ? -1 : f % ng;
                p->group_all = f >= ng;
                rc = drop_graphs(*p);
                if (rc) return rc;
                float ms = 1e30f;
                for (int k = 0; k < 3; k++) {
                    long long n = 0;
                    const auto t0 = std::chrono::steady_clock::now();
                    rc = pzl_search(plan, lo, lo + band, nullptr, 0, &n, nullptr);
                    if (rc) {
                        p->side_group = -1;
                        p->group_all = false;
                        return rc;
                    }
                    ms = std::min(ms, (float)std::chrono::duration<double, std::milli>(
                                          std::chrono::steady_clock::now() - t0)
                                          .count());
                }
                snprintf(line, sizeof line, "%s%s in a search band of %lld rsteps: %.3f ms\n",
                         f < 0 ? "one launch per FFT length" : (f >= ng ? "every length in " : "shorter lengths in "),
                         f < 0 ? "" : pzl_corr_group_name(f % ng), band, ms);
                rep += line;
                if (ms < 0.98f * best_ms) {
                    best_ms = ms;
                    best = f;
                }
            }
            p->side_group = best < 0 ? -1 : best % ng;
            p->group_all = best >= ng;
            rc = drop_graphs(*p);
            if (rc) return rc;
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tuned_group[group_key(*p)] = best;
        }
    }
    // CU split (pzl_plan_set_cu_split): the harmonic sums on 3 or 2 CUs per
    // shader engine, the segments and correlations on every CU (keeping CUs
    // for the sums alone measured slower), each with the harmonic sum's 4 and
    // 1 r per thread (on few CUs the sum is bound by address issue, which 4 r
    // per thread cuts); a later split replaces an earlier one when it is 1%
    // faster,
    // timed on a band of 384 rsteps and at least 16 batches (best of two; in a
    // short band the last batch's sum, slower on fewer CUs, would weigh as much
    // as the steady state gains); the fastest split is kept when it is more
    // than 2% faster than the whole chip for every kernel
    const long long sband = std::min<long long>(nr, std::max<long long>(16LL * p->B, 384));
    if (nr >= 16LL * p->B && p->overlap_sum) {
        const long long lo = (nr - sband) / 2;
        int ncus = 0, nxcc = 1;
        PZL_HIPCHK(hipDeviceGetAttribute(&ncus, hipDeviceAttributeMultiprocessorCount, p->device));
        if (hipDeviceGetAttribute(&nxcc, hipDeviceAttributeNumberOfXccs, p->device) != hipSuccess) {
            (void)hipGetLastError();
            nxcc = 1;
        }
        const int u = 4 * std::max(nxcc, 1);  // one CU per shader engine
        struct Split {
            int sum, only, rpt;
        };
        std::vector<Split> splits = {{0, 0, p->sum_rpt}};
        for (int k = 3; k >= 2; k--)
            for (int rpt : {4, 1})
                if ((k + 1) * u <= ncus) splits.push_back({k * u, 0, rpt});
        const int rpt0 = p->sum_rpt;
        float base_ms = 1e30f, best_ms = 1e30f;  // the whole chip; the fastest split
        Split best = splits[0];
        for (const Split &sp : splits) {
            if (pzl_plan_set_cu_split(plan, sp.sum, sp.only)) {
                // CU masks refused (not a failure of the search): keep the whole chip
                snprintf(line, sizeof line, "harmonic sums on %d CUs: not available (%s)\n", sp.sum,
                         pzl::g_err.c_str());
                rep += line;
                (void)hipGetLastError();
                pzl::g_err.clear();
                break;
            }
            p->sum_rpt = sp.rpt;
            float ms = 1e30f;
            for (int k = 0; k < 2; k++) {
                long long n = 0;
                const auto t0 = std::chrono::steady_clock::now();
                rc = pzl_search(plan, lo, lo + sband, nullptr, 0, &n, nullptr);
                if (rc) {
                    p->sum_rpt = rpt0;
                    (void)pzl_plan_set_cu_split(plan, 0, 0);
                    return rc;
                }
                ms = std::min(ms, (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                                                  t0).count());
            }
            snprintf(line, sizeof line,
                     "harmonic sums on %d CUs (%d of them only for the sums), %d r per thread, in a search band of "
                     "%lld rsteps: %.3f ms\n", sp.sum, sp.only, sp.rpt, sband, ms);
            rep += line;
            if (&sp == &splits[0]) {
                base_ms = ms;
            } else if (best_ms > 1e29f || ms < 0.99f * best_ms) {
                // the first split, or one clearly faster: more summing CUs
                // (tried first) leave the sums more room when hits come in bursts
                best_ms = ms;
                best = sp;
            }
        }
        if (!(best_ms < 0.98f * base_ms)) best = splits[0];
        rc = pzl_plan_set_cu_split(plan, best.sum, best.only);
        if (rc) return rc;
        p->sum_rpt = best.rpt;
    }
    if (report && report_len > 0) {
        snprintf(report, (size_t)report_len, "%s", rep.c_str());
    }
    return 0;
}

int pzl_plan_set_deadline(void *plan, double ms)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    p->has_deadline = ms > 0;
    if (p->has_deadline)
        p->deadline = std::chrono::steady_clock::now() +
                      std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                          std::chrono::duration<double, std::milli>(ms));
    return 0;
}

int pzl_plan_set_overlap(void *plan, int on)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    int rc = sync_all(*p);
    if (rc) return rc;
    p->overlap_sum = on != 0;
    p->side_classes = on >= 2;
    return 0;
}

// The correlation's waves take every register of a SIMD, so a harmonic sum
// launched beside it only gets workgroup slots the correlation gives up, and
// the two kernels' times add.  A CU split confines the summing stream to
// sum_cus compute units, the first sum_only_cus of which the segments and
// correlations do not use: the memory-bound sum of one batch then runs beside
// the VALU-bound correlation of the next.  Mask bit i is the i-th CU in the
// driver's order, which deals consecutive bits to the dies and, within a die,
// to its shader engines in turn, so the first n bits spread n CUs evenly;
// workgroups are spread evenly over the shader engines, so a split is worth
// having only in whole CUs per shader engine.
int pzl_plan_set_cu_split(void *plan, int sum_cus, int sum_only_cus)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    PZL_HIPCHK(hipSetDevice(p->device));
    int ncus = 0, nxcc = 1;
    PZL_HIPCHK(hipDeviceGetAttribute(&ncus, hipDeviceAttributeMultiprocessorCount, p->device));
    if (hipDeviceGetAttribute(&nxcc, hipDeviceAttributeNumberOfXccs, p->device) != hipSuccess) {
        (void)hipGetLastError();
        nxcc = 1;  // a runtime that does not report dies: one
    }
    // every shader engine of every die keeps compute units in each mask (a
    // queue whose mask leaves a die without any never finishes that die's
    // share of a launch): the same count on each die, at least one per shader
    // engine (4 per die)
    nxcc = std::max(nxcc, 1);
    auto ok = [&](int n) { return n == 0 || (n % nxcc == 0 && n >= 4 * nxcc && ncus - n >= 4 * nxcc); };
    if (sum_cus < 0 || sum_only_cus < 0 || sum_only_cus > sum_cus || !ok(sum_cus) || !ok(sum_only_cus))
        return fail(PZL_ERR_ARG, "cu split %d/%d: counts of 0, or multiples of %d from %d to %d, the second at "
                                 "most the first", sum_cus, sum_only_cus, nxcc, 4 * nxcc, ncus - 4 * nxcc);
    int rc = drop_graphs(*p);
    if (rc) return rc;
    const int words = (ncus + 31) / 32;
    std::vector<uint32_t> sum_mask(words, 0), rest_mask(words, 0);
    for (int i = 0; i < ncus; i++) {
        if (i < sum_cus) sum_mask[i / 32] |= 1u << (i % 32);
        if (i >= sum_only_cus) rest_mask[i / 32] |= 1u << (i % 32);
    }
    // st, st3, st4: the sums on their CUs, the segments and correlations on
    // the rest (every CU when sum_only_cus is 0).  The three are made together,
    // once per split, and kept: a summing stream made alone beside the plan's
    // first streams has had its launches wait for the correlation's in some
    // processes (as if both shared one hardware dispatcher), never one of
    // three made at once
    hipStream_t s[3] = {p->st_all[0], p->st_all[1], p->st_all[2]};
    if (sum_cus) {
        bool have = false;
        for (const auto &m : p->masked)
            if (m.first == sum_only_cus && m.count == sum_cus) {
                s[m.role] = m.s;
                have = true;
            }
        if (!have) {
            hipError_t e = hipSuccess;
            for (int k = 0; k < 3 && e == hipSuccess; k++) {
                e = hipExtStreamCreateWithCUMask(&s[k], (uint32_t)words, (k == 1 ? sum_mask : rest_mask).data());
                if (e == hipSuccess) p->masked.push_back({k, sum_only_cus, sum_cus, s[k]});
            }
            if (e != hipSuccess)
                return fail(PZL_ERR_HIP, "cu split %d/%d: %s", sum_cus, sum_only_cus, hipGetErrorString(e));
        }
    }
    p->st = s[0];
    p->st3 = s[1];
    p->st4 = s[2];
    p->cu_split = sum_cus;
    p->cu_split_only = sum_only_cus;
    return 0;
}

int pzl_plan_cu_split(void *plan, int *sum_cus, int *sum_only_cus)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    if (sum_cus) *sum_cus = p->cu_split;
    if (sum_only_cus) *sum_only_cus = p->cu_split_only;
    return 0;
}

int pzl_plan_set_graph(void *plan, int on)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    p->use_graph = on != 0;
    return 0;
}

int pzl_tune_batchsize(const PzlConfig *cfg, int device, const float *host_fft, long long numbins,
                       const int *candidates, int ncand, long long band_rsteps, int *best, double *ms_per_rstep)
{
    if (!cfg || !host_fft || !candidates || ncand < 1 || !best) return fail(PZL_ERR_ARG, "bad arguments");
    double best_ms = 1e300;
    *best = candidates[0];
    for (int k = 0; k < ncand; k++) {
        PzlConfig c = *cfg;
        c.batchsize = candidates[k];
        void *plan = nullptr;
        int rc = pzl_plan_create(&c, device, &plan);
        if (rc) return rc;
        Plan *p = (Plan *)plan;
        rc = pzl_set_spectrum(plan, host_fft, numbins);
        const long long nr = p->g.nrsteps;
        const long long band = std::max<long long>(1, std::min<long long>(band_rsteps, nr));
        const long long lo = std::max<long long>(0, std::min<long long>(nr / 2, nr - band));
        if (rc == 0) rc = pzl_autotune(plan, 2, nullptr, 0);
        if (rc == 0) rc = pzl_search(plan, lo, lo + band, nullptr, 0, nullptr, nullptr);  // warm
        double ms = 0;
        if (rc == 0) {
            const auto t0 = std::chrono::steady_clock::now();
            rc = pzl_search(plan, lo, lo + band, nullptr, 0, nullptr, nullptr);
            ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / band;
        }
        pzl_plan_destroy(plan);
        if (rc) return rc;
        if (ms_per_rstep) ms_per_rstep[k] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            *best = candidates[k];
        }
    }
    return 0;
}

int pzl_plan_corr_group(void *plan, int *group, int *all)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    if (group) *group = p->side_group;
    if (all) *all = p->side_group >= 0 && p->group_all;
    return 0;
}

int pzl_plan_set_corr_group(void *plan, int group, int all)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    if (group >= 0 && !group_covers(*p, group, all != 0))
        return fail(PZL_ERR_ARG, "correlation group %d does not hold every %sFFT length of this plan", group,
                    all ? "" : "shorter ");
    if (group < 0) {
        group = -1;
        all = 0;
    }
    if (group != p->side_group || (all != 0) != p->group_all) {
        int rc = drop_graphs(*p);
        if (rc) return rc;
    }
    p->side_group = group;
    p->group_all = all != 0;
    return 0;
}

int pzl_plan_sum_form(void *plan, int *order, int *rpt)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    if (order) *order = p->sum_order;
    if (rpt) *rpt = p->sum_rpt;
    return 0;
}

int pzl_plan_set_sum_form(void *plan, int order, int rpt)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    if ((order != 0 && order != 1) || (rpt != 1 && rpt != 2 && rpt != 4))
        return fail(PZL_ERR_ARG, "harmonic-sum form: order 0 or 1, 1, 2 or 4 r per thread");
    int rc = sync_all(*p);
    if (rc) return rc;
    p->sum_order = order;
    p->sum_rpt = rpt;
    return 0;
}

int pzl_plan_variant(void *plan, int fftlen, int *variant)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    for (auto &c : p->classes)
        if (c.fftlen == fftlen) {
            if (variant) *variant = c.variant;
            return 0;
        }
    return fail(PZL_ERR_ARG, "no FFT length %d in this plan", fftlen);
}

int pzl_plan_set_variant(void *plan, int fftlen, int variant)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    if (variant < 0 || variant >= pzl_corr_num_variants(fftlen))
        return fail(PZL_ERR_ARG, "FFT length %d has no variant %d", fftlen, variant);
    for (auto &c : p->classes)
        if (c.fftlen == fftlen) {
            if (c.variant != variant) {
                int rc = drop_graphs(*p);
                if (rc) return rc;
            }
            c.variant = variant;
            return 0;
        }
    return fail(PZL_ERR_ARG, "no FFT length %d in this plan", fftlen);
}

int pzl_get_cands(void *plan, PzlCand *cands, long long maxcands, long long *ncands)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    if (ncands) *ncands = (long long)p->last_cands.size();
    for (long long k = 0; k < (long long)p->last_cands.size() && k < maxcands; k++) cands[k] = p->last_cands[k];
    return 0;
}

int pzl_plan_set_hit_chunk(void *plan, long long records)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    if (records < 1 || records * (long long)sizeof(PzlHostHit) > p->hpool_limit)
        return fail(PZL_ERR_ARG, "hit chunk of %lld records: 1 .. %lld", records,
                    p->hpool_limit / (long long)sizeof(PzlHostHit));
    p->hit_chunk = records;
    return 0;
}

int pzl_plan_set_keep_hits(void *plan, long long limit)
{
    Plan *p = (Plan *)plan;
    if (!p || limit < 0) return fail(PZL_ERR_ARG, "bad arguments");
    p->keep_hits_limit = limit;
    return 0;
}

int pzl_get_hits(void *plan, long long *keys, float *powers, long long maxhits, long long *nhits)
{
    Plan *p = (Plan *)plan;
    if (!p) return fail(PZL_ERR_ARG, "null plan");
    if (!p->last_hits_complete)
        return fail(PZL_ERR_UNSUPPORTED, "the last search had more than %lld hits; they were not kept",
                    p->keep_hits_limit);
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
    PZL_HIPCHK(hipSetDevice(p->device));
    PZL_HIPCHK(hipMemcpy(out, p->d_bank + H.bank_off + (long long)row * H.fftlen, sizeof(float2) * H.fftlen,
                         hipMemcpyDeviceToHost));
    // power-of-two rows are held as K / fftlen (an exact scale): undo it
    if ((H.fftlen & (H.fftlen - 1)) == 0)
        for (int k = 0; k < 2 * H.fftlen; k++) out[k] *= (float)H.fftlen;
    return 0;
}

int pzl_compute_plane(void *plan, long long rstep, int idx, float *out, long long maxfloats, int *numrs,
                      long long *rlo)
{
    Plan *p = (Plan *)plan;
    if (!p || idx < 0 || idx >= (int)p->harms.size() || rstep < 0 || rstep >= p->g.nrsteps)
        return fail(PZL_ERR_ARG, "bad plane request");
    if (!p->d_fft) return fail(PZL_ERR_NOSPECTRUM, "no spectrum set");
    PZL_HIPCHK(hipSetDevice(p->device));
    PZL_HIPCHK(hipMemsetAsync(p->d_counts, 0, 4 * sizeof(unsigned long long), p->st));
    int rc = launch_batch(*p, rstep, 1, 0, 0, nullptr, nullptr);
    if (rc) return rc;
    {
        int src = sync_all(*p);
        if (src) return src;
    }
    PzlHrTab hr;
    PZL_HIPCHK(hipMemcpy(&hr, p->d_hr[0] + idx, sizeof hr, hipMemcpyDeviceToHost));
    const PzlHarm &H = p->harms[idx];
    if (numrs) *numrs = hr.numrs;
    if (rlo) *rlo = hr.rlo;
    if (out) {
        const long long need = (long long)H.nrows * hr.numrs;
        if (need > maxfloats) return fail(PZL_ERR_ARG, "output too small (%lld floats needed)", need);
        PZL_HIPCHK(hipMemcpy2D(out, sizeof(float) * hr.numrs, p->d_planes[0] + H.plane_off,
                               sizeof(float) * H.rs_stride, sizeof(float) * hr.numrs, H.nrows,
                               hipMemcpyDeviceToHost));
    }
    PZL_HIPCHK(hipMemset(p->d_counts, 0, 4 * sizeof(unsigned long long)));
    return 0;
}

int pzl_compute_plane_at(void *plan, double fullrlo, int idx, float *out, long long maxfloats, int *numrs,
                         long long *rlo)
{
    Plan *p = (Plan *)plan;
    if (!p || idx < 0 || idx >= (int)p->harms.size()) return fail(PZL_ERR_ARG, "bad plane request");
    if (!p->d_fft) return fail(PZL_ERR_NOSPECTRUM, "no spectrum set");
    PZL_HIPCHK(hipSetDevice(p->device));
    const long long spare = p->g.nrsteps;
    PZL_HIPCHK(hipMemcpyAsync(p->d_startr + spare, &fullrlo, sizeof(double), hipMemcpyHostToDevice, p->st));
    int rc = launch_batch(*p, spare, 1, 0, 0, nullptr, nullptr, false);
    if (rc) return rc;
    {
        int src = sync_all(*p);
        if (src) return src;
    }
    PzlHrTab hr;
    PZL_HIPCHK(hipMemcpy(&hr, p->d_hr[0] + idx, sizeof hr, hipMemcpyDeviceToHost));
    const PzlHarm &H = p->harms[idx];
    if (numrs) *numrs = hr.numrs;
    if (rlo) *rlo = hr.rlo;
    if (out) {
        const long long need = (long long)H.nrows * hr.numrs;
        if (need > maxfloats) return fail(PZL_ERR_ARG, "output too small (%lld floats needed)", need);
        PZL_HIPCHK(hipMemcpy2D(out, sizeof(float) * hr.numrs, p->d_planes[0] + H.plane_off,
                               sizeof(float) * H.rs_stride, sizeof(float) * hr.numrs, H.nrows,
                               hipMemcpyDeviceToHost));
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
    PZL_HIPCHK(hipSetDevice(p->device));
    if (!p->d_sums) PZL_HIPCHK(plan_malloc(p, &p->d_sums, sizeof(float) * need));
    PZL_HIPCHK(hipMemsetAsync(p->d_counts, 0, 4 * sizeof(unsigned long long), p->st));
    int rc = launch_batch(*p, rstep, 1, 0, 0, p->d_sums, nullptr);
    if (rc) return rc;
    {
        int src = sync_all(*p);
        if (src) return src;
    }
    PZL_HIPCHK(hipMemcpy(out, p->d_sums, sizeof(float) * need, hipMemcpyDeviceToHost));
    PZL_HIPCHK(hipMemset(p->d_counts, 0, 4 * sizeof(unsigned long long)));
    return 0;
}

}  // extern "C"

// Host-side candidate insertion on its own (no GPU): applies n hits in order
// with insert_new_accelcand semantics.  Exposed for testing the list logic.
extern "C" int pzl_insert_cands(const double *rr, const float *sigma, const float *power, const int *numharm,
                                const double *z, const double *w, long long n, PzlCand *out, long long maxout,
                                long long *nout)
{
    if (n < 0 || (n > 0 && (!rr || !sigma || !power || !numharm || !z || !w)))
        return fail(PZL_ERR_ARG, "bad hit arrays");
    CandList cl;
    for (long long i = 0; i < n; i++) {
        cl.insert(rr[i], sigma[i], [&]() {
            PzlCand c;
            c.power = power[i];
            c.sigma = sigma[i];
            c.numharm = numharm[i];
            c.r = rr[i];
            c.z = z[i];
            c.w = w[i];
            return c;
        });
    }
    const auto &v = cl.items();
    if (nout) *nout = (long long)v.size();
    for (long long i = 0; i < (long long)v.size() && i < maxout; i++) out[i] = v[i];
    return 0;
}
