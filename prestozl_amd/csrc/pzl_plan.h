// pzl_plan.h -- host-side state of one search plan (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <chrono>
#include <vector>

#include "pzl_accel.h"
#include "pzl_kernels.h"

namespace pzl {

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
extern thread_local std::string g_err;

#define PZL_HIPCHK(call)                                                                                \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return ::pzl::fail(PZL_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                               __LINE__);                                                               \
    } while (0)

struct SizeClass {
    int fftlen;
    std::vector<int> harms;
    std::vector<int> nrows;  // kernel rows of each harmonic (host copy for the launcher)
    int *d_harms = nullptr;
    int variant = 0;         // correlation kernel variant (autotuned)
};

// create_accelobs (accel_utils.c:1636-2037) + the harmonic shapes of
// create_subharminfos (:236-282) + the rstep list of accelsearch.c:183-188.
int compute_geometry(const PzlConfig &c, PzlGeometry &g, std::vector<PzlHarm> *harms_out,
                     std::vector<double> *startr_out);

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
    int int_rlo = 1;  // every startr is an integer (integer r maps on the GPU)
    int B = 8;
    int device = 0;
    hipStream_t st = nullptr;   // segments and correlations
    hipStream_t st2 = nullptr;  // hit ordering
    hipStream_t st3 = nullptr;  // harmonic sums (when overlapped)
    hipStream_t st4 = nullptr;  // segments (when overlapped), ahead of the correlation
    hipStream_t st5 = nullptr;  // hit records to host memory
    hipEvent_t ev_prep = nullptr;    // the last slot's records are prepared (st2)
    hipEvent_t ev_copied = nullptr;  // the last slot's records are copied out (st5)
    bool overlap_sum = true;    // harmonic sums on st3, concurrent with the next batch's correlation
    bool side_classes = true;   // with overlap_sum: the next batch's segments ahead on st4, and the shorter
                                // FFT lengths correlated there beside the longest (unless grouped)
    int cu_split = 0;           // > 0: st3 on the first cu_split compute units ...
    int cu_split_only = 0;      // ... and st, st4 on all but the first cu_split_only
    hipStream_t st_all[3] = {};  // st, st3, st4 as created with the plan (every compute unit)
    // CU-masked streams of each split used (role 0: st, 1: st3, 2: st4; the
    // split's sum_only_cus and sum_cus), made together and kept until the plan
    // is destroyed
    struct MaskedStream {
        int role, first, count;
        hipStream_t s;
    };
    std::vector<MaskedStream> masked;
    // resident
    PzlHarm *d_harms = nullptr;
    float2 *d_bank = nullptr;  // shared, owned by the bank cache
    bool bank_built = false;   // this plan's creation built the bank
    double *d_startr = nullptr;
    double numindep_d[6] = {0, 0, 0, 0, 0, 0};
    uint16_t *d_zinds = nullptr;
    int *d_winds = nullptr;
    int *d_zoff = nullptr;  // [numw][numz][nh - 1] subharmonic row byte offsets (harmonic sum)
    const float2 *d_fft = nullptr;
    float2 *d_fft_own = nullptr;
    long long dev_bytes = 0;  // device memory this plan has allocated and holds (the bank aside)
    char *d_arena = nullptr;  // one allocation holding every buffer below (spectrum and d_sums aside)
    long long fft_bins = 0;
    // per batch, three sets: the correlation of one batch fills one set while
    // the harmonic sum of the previous batch reads another, and the segments
    // of the next batch are written into the third beside both, so that no
    // batch's correlation waits for its segments
    static constexpr int kSets = 3;
    float2 *d_segs[kSets] = {};
    long long seg_stride = 0;
    uint16_t *d_rinds[kSets] = {};
    PzlHrTab *d_hr[kSets] = {};
    float *d_planes[kSets] = {};
    long long plane_stride = 0;
    hipEvent_t ev_corr[kSets] = {};  // the set's planes are written
    hipEvent_t ev_sum[kSets] = {};   // the set's planes have been summed
    hipEvent_t ev_seg[kSets] = {};   // the set's segments are ready
    hipEvent_t ev_side[kSets] = {};  // the set's shorter FFT lengths are correlated (st4)
    // hits: a ring of unordered slots (one per batch in flight), one ordered copy
    static constexpr int kSlots = 4;
    int nslots = 2;
    long long hitcap = 0;
    int key_bits = 64;
    long long *d_hkeys[kSlots] = {};
    float *d_hpows[kSlots] = {};
    unsigned long long *d_counts = nullptr;  // [slot][count, spare]
    unsigned long long *h_counts = nullptr;  // pinned mirror
    long long *d_skeys = nullptr;
    float *d_spows = nullptr;
    PzlHostHit *d_recs = nullptr;
    void *d_sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
    // pinned hit-record buffers handed to the insertion thread, one chunk of a
    // batch's records each, with the event of the copy that fills it; the
    // pool grows up to hpool_limit bytes and keeps its buffers, so that bursts
    // of hits never make the launching thread pin memory
    struct HostBuf {
        PzlHostHit *ptr = nullptr;
        long long cap = 0;
        hipEvent_t ev = nullptr;
        bool busy = false;
    };
    std::vector<HostBuf> hpool;
    long long hpool_bytes = 0;
    long long hpool_limit = 2LL << 30;
    long long hit_chunk = 1LL << 20;  // records per pool buffer (pzl_plan_set_hit_chunk)
    hipEvent_t ev_count[kSlots] = {};
    hipEvent_t ev_free[kSlots] = {};
    float *d_sums = nullptr;
    // graph replay of full batches (one executable graph per buffer set)
    bool use_graph = true;
    static constexpr int kMaxGraphFailures = 8;
    int graph_failures = 0;  // batch graphs the runtime refused (those batches ran eagerly)
    hipGraphExec_t gexec[kSets] = {};  // correlation part, per buffer set
    // pzl_plan_set_deadline: searches stop launching batches after this time
    bool has_deadline = false;
    std::chrono::steady_clock::time_point deadline{};
    hipStream_t sum_stream() const { return overlap_sum ? st3 : st; }
    int side_group = -1;     // >= 0: the shorter FFT lengths in one grouped launch (autotuned) ...
    bool group_all = false;  // ... and the longest length in it too: one correlation launch per batch
    int sum_order = 0;  // harmonic-sum workgroup order (autotuned)
    int sum_rpt = 1;    // harmonic-sum fundamental r per thread (autotuned)
    // results of the last search
    std::vector<PzlHostHit> last_hits;
    std::vector<PzlCand> last_cands;
    bool last_hits_complete = true;
    long long keep_hits_limit = 1LL << 22;
};

}  // namespace pzl
