// Host-only HIP runtime stand-in and fake search kernels (see hip/hip_runtime.h).
//
// The fake kernels keep the contract the host code relies on, not the
// arithmetic: the harmonic sum appends a deterministic set of hits per rstep
// (a function of the rstep only) to the batch's slot with an atomic count, the
// sort orders (key, power) pairs, and the record preparation decodes keys into
// host records.  Every kernel touches the first and last element of the
// buffers it would write, so a wrong size or a freed buffer shows up under
// AddressSanitizer.  Test infrastructure only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <memory>
#include <thread>
#include <vector>

#include "pzl_kernels.h"
#include "pzl_launch.h"

using Clock = std::chrono::steady_clock;

struct StubStream {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::function<void()>> q;
    bool busy = false, stop = false;
    std::thread th;
    std::vector<std::function<void()>> *capture = nullptr;  // being captured (by the capturing thread)
    StubStream()
    {
        th = std::thread([this] {
            for (;;) {
                std::function<void()> f;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return stop || !q.empty(); });
                    if (q.empty()) return;
                    f = std::move(q.front());
                    q.pop_front();
                    busy = true;
                }
                f();
                {
                    std::lock_guard<std::mutex> lk(mu);
                    busy = false;
                }
                cv.notify_all();
            }
        });
    }
    void push(std::function<void()> f)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            q.push_back(std::move(f));
        }
        cv.notify_all();
    }
    void drain()
    {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return q.empty() && !busy; });
    }
    ~StubStream()
    {
        drain();
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        th.join();
    }
};

struct StubEvent {
    std::mutex mu;
    std::condition_variable cv;
    unsigned long long recorded = 0, done = 0;
    Clock::time_point t{};
};

namespace {
std::mutex g_streams_mu;
std::set<StubStream *> g_streams;
std::mutex g_mem_mu;
std::map<void *, size_t> g_dev;
long long g_dev_bytes = 0;

// Run f on the stream, or now when the stream is the null stream; a stream
// being captured records f into its graph instead.
void enqueue(hipStream_t st, std::function<void()> f)
{
    if (!st) {
        f();
        return;
    }
    if (st->capture) {
        st->capture->push_back(std::move(f));
        return;
    }
    st->push(std::move(f));
}

void drain_all()
{
    std::vector<StubStream *> v;
    {
        std::lock_guard<std::mutex> lk(g_streams_mu);
        v.assign(g_streams.begin(), g_streams.end());
    }
    for (auto *s : v) s->drain();
}

// A kernel's write to the first and last byte of a buffer it fills ...
template <typename T> void touch(T *p, long long n)
{
    if (!p || n <= 0) return;
    volatile char *a = (volatile char *)p;
    volatile char *b = (volatile char *)(p + n) - 1;
    *a = *a;
    *b = *b;
}

// ... and its read of a buffer it only reads (inputs shared by searches).
template <typename T> void touch_r(const T *p, long long n)
{
    if (!p || n <= 0) return;
    const volatile char *a = (const volatile char *)p;
    const volatile char *b = (const volatile char *)(p + n) - 1;
    (void)*a;
    (void)*b;
}
}  // namespace

const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "stub error"; }
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipGetDevice(int *d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int *n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipDeviceSynchronize()
{
    drain_all();
    return hipSuccess;
}
hipError_t hipMemGetInfo(size_t *free_b, size_t *total_b)
{
    std::lock_guard<std::mutex> lk(g_mem_mu);
    *total_b = 288ull << 30;
    *free_b = *total_b - (size_t)g_dev_bytes;
    return hipSuccess;
}
long long hipstub_device_bytes()
{
    std::lock_guard<std::mutex> lk(g_mem_mu);
    return g_dev_bytes;
}
hipError_t hipMalloc(void **p, size_t n)
{
    *p = aligned_alloc(256, (n + 255) / 256 * 256);
    if (!*p) return hipErrorOutOfMemory;
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_dev[*p] = n;
    g_dev_bytes += (long long)n;
    return hipSuccess;
}
hipError_t hipFree(void *p)
{
    if (!p) return hipSuccess;
    drain_all();
    {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        auto it = g_dev.find(p);
        if (it == g_dev.end()) {
            fprintf(stderr, "hipstub: hipFree of %p, not a device allocation\n", p);
            abort();
        }
        g_dev_bytes -= (long long)it->second;
        g_dev.erase(it);
    }
    free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned)
{
    *p = malloc(n);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void *p)
{
    drain_all();
    free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind)
{
    drain_all();
    memcpy(dst, src, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind, hipStream_t st)
{
    enqueue(st, [=] { memcpy(dst, src, n); });
    return hipSuccess;
}
hipError_t hipMemcpy2D(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                       hipMemcpyKind)
{
    drain_all();
    for (size_t r = 0; r < height; r++) memcpy((char *)dst + r * dpitch, (const char *)src + r * spitch, width);
    return hipSuccess;
}
hipError_t hipMemset(void *p, int v, size_t n)
{
    drain_all();
    memset(p, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t st)
{
    enqueue(st, [=] { memset(p, v, n); });
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *st, unsigned)
{
    *st = new StubStream();
    std::lock_guard<std::mutex> lk(g_streams_mu);
    g_streams.insert(*st);
    return hipSuccess;
}
hipError_t hipExtStreamCreateWithCUMask(hipStream_t *st, uint32_t, const uint32_t *)
{
    return hipStreamCreateWithFlags(st, 0);
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int)
{
    *v = a == hipDeviceAttributeMultiprocessorCount ? 256 : a == hipDeviceAttributeNumberOfXccs ? 8 : 0;
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t st)
{
    {
        std::lock_guard<std::mutex> lk(g_streams_mu);
        g_streams.erase(st);
    }
    delete st;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t st)
{
    if (st) st->drain();
    else drain_all();
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t st, hipEvent_t ev, unsigned)
{
    unsigned long long g;
    {
        std::lock_guard<std::mutex> lk(ev->mu);
        g = ev->recorded;
    }
    enqueue(st, [=] {
        std::unique_lock<std::mutex> lk(ev->mu);
        ev->cv.wait(lk, [&] { return ev->done >= g; });
    });
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *ev)
{
    *ev = new StubEvent();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *ev, unsigned) { return hipEventCreate(ev); }
hipError_t hipEventDestroy(hipEvent_t ev)
{
    drain_all();  // a pending record or wait may still name it
    delete ev;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t ev, hipStream_t st)
{
    unsigned long long g;
    {
        std::lock_guard<std::mutex> lk(ev->mu);
        g = ++ev->recorded;
    }
    enqueue(st, [=] {
        {
            std::lock_guard<std::mutex> lk(ev->mu);
            ev->done = std::max(ev->done, g);
            ev->t = Clock::now();
        }
        ev->cv.notify_all();
    });
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t ev)
{
    std::lock_guard<std::mutex> lk(ev->mu);
    return ev->done >= ev->recorded ? hipSuccess : hipErrorNotReady;
}
hipError_t hipEventSynchronize(hipEvent_t ev)
{
    std::unique_lock<std::mutex> lk(ev->mu);
    const unsigned long long g = ev->recorded;
    ev->cv.wait(lk, [&] { return ev->done >= g; });
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
    std::scoped_lock lk(a->mu, b->mu);
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}
// Graphs: operations captured from a stream, or added as kernel nodes (each
// a recorded launch of the stand-in's kernels), replayed in order on the
// stream a launch names.
struct StubGraph {
    std::vector<std::function<void()>> ops;
};
struct StubGraphExec {
    std::vector<std::function<void()>> ops;
};

hipError_t hipGraphCreate(hipGraph_t *g, unsigned)
{
    if (!g) return hipErrorInvalidValue;
    *g = new StubGraph{};
    return hipSuccess;
}

hipError_t hipGraphAddKernelNode(hipGraphNode_t *n, hipGraph_t g, const hipGraphNode_t *, size_t,
                                 const hipKernelNodeParams *p)
{
    if (!g || !p || !p->kernelParams) return hipErrorInvalidValue;
    const std::function<void()> *op = *(std::function<void()> *const *)p->kernelParams[0];
    g->ops.push_back(*op);
    if (n) *n = (hipGraphNode_t)(uintptr_t)g->ops.size();
    return hipSuccess;
}

hipError_t hipLaunchKernel(const void *, dim3, dim3, void **, size_t, hipStream_t) { return hipErrorInvalidValue; }

hipError_t hipStreamBeginCapture(hipStream_t st, hipStreamCaptureMode)
{
    if (!st || st->capture) return hipErrorInvalidValue;
    st->capture = new std::vector<std::function<void()>>();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t st, hipGraph_t *g)
{
    if (!st || !st->capture || !g) return hipErrorInvalidValue;
    *g = new StubGraph{std::move(*st->capture)};
    delete st->capture;
    st->capture = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, char *, size_t)
{
    if (!e || !g) return hipErrorInvalidValue;
    *e = new StubGraphExec{g->ops};
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t st)
{
    if (!e) return hipErrorInvalidValue;
    for (auto &f : e->ops) enqueue(st, f);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e)
{
    delete e;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete g;
    return hipSuccess;
}

// ------------------------------------------------------------- fake kernels
namespace {
unsigned mix(unsigned long long x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    return (unsigned)x;
}
}  // namespace

extern "C" {

hipError_t pzl_launch_bank_build(const PzlHarm *, int, int total_rows, int max_fftlen, float2 *d_bank,
                                 hipStream_t st)
{
    enqueue(st, [=] {
        touch(d_bank, 1);
        (void)total_rows;
        (void)max_fftlen;
    });
    return hipSuccess;
}

extern "C++" {
thread_local std::vector<PzlKernelRec> *pzl_recorder = nullptr;
}

namespace {
// The stand-in's kernels: enqueued, or -- while the engine records a batch
// -- recorded, the argument block holding the address of the operation (kept
// until the process ends) for hipGraphAddKernelNode.
std::mutex g_ops_mu;
std::vector<std::unique_ptr<std::function<void()>>> g_ops;

hipError_t launch_op(hipStream_t st, std::function<void()> f)
{
    if (!pzl_recorder) {
        enqueue(st, std::move(f));
        return hipSuccess;
    }
    std::function<void()> *op;
    {
        std::lock_guard<std::mutex> lk(g_ops_mu);
        g_ops.push_back(std::make_unique<std::function<void()>>(std::move(f)));
        op = g_ops.back().get();
    }
    PzlKernelRec r{nullptr, dim3(1), dim3(1), 0, std::vector<unsigned char>(sizeof op)};
    std::memcpy(r.arg.data(), &op, sizeof op);
    pzl_recorder->push_back(std::move(r));
    return hipSuccess;
}
}  // namespace

hipError_t pzl_launch_segments(const PzlBatchArgs *a, int, hipStream_t st)
{
    const PzlBatchArgs b = *a;
    return launch_op(st, [=] {
        touch(b.segs, b.seg_stride * b.nb);
        touch(b.hr_tab, (long long)b.nb * b.nh);
        touch_r(b.fft, b.numbins);
        if (b.rstep0_dev) touch_r(b.rstep0_dev, 1);
    });
}

hipError_t pzl_launch_corr(int, int, const PzlCorrArgs *a, const int *, hipStream_t st)
{
    const PzlCorrArgs c = *a;
    return launch_op(st, [=] {
        // the class's own plane rows only: classes run concurrently on two streams
        for (int b = 0; b < c.nb; b++)
            for (int k = 0; k < c.nclass; k++) {
                const PzlHarm &H = c.harms[c.class_harms[k]];
                touch(c.planes + b * c.plane_stride + H.plane_off, (long long)H.nrows * H.rs_stride);
                touch_r(c.segs + b * c.seg_stride + H.seg_off, H.fftlen / 2);
            }
    });
}

int pzl_corr_num_variants(int) { return 1; }
const char *pzl_corr_variant_name(int, int) { return "stub"; }

// One grouped launch: each class's plane rows, as pzl_launch_corr.
hipError_t pzl_launch_corr_group(int group, int ncls, const int *, const PzlCorrArgs *const *args, const int *const *,
                                 hipStream_t st)
{
    if (group != 0) return hipErrorInvalidValue;
    for (int c = 0; c < ncls; c++) {
        hipError_t e = pzl_launch_corr(0, 0, args[c], nullptr, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
int pzl_corr_num_groups(void) { return 1; }
const char *pzl_corr_group_name(int group) { return group == 0 ? "stub group" : nullptr; }
int pzl_corr_group_has(int group, int n) { return group == 0 && n >= 8 && n <= 4096; }
int pzl_corr_supported(int n) { return n >= 8 && n <= 4096; }

// Hits of rstep r: a count and cells that depend on r alone.
hipError_t pzl_launch_harmsum(const PzlSumArgs *a, hipStream_t st)
{
    const PzlSumArgs s = *a;
    enqueue(st, [=] {
        touch_r(s.planes, s.plane_stride * s.nb);
        const long long r0 = s.rstep0_dev ? *s.rstep0_dev : s.rstep0;
        for (int b = 0; b < s.nb; b++) {
            const long long rstep = r0 + b;
            const unsigned h = mix((unsigned long long)rstep + 12345);
            const int n = (h % 7 == 0) ? 0 : (int)(h % 37);
            for (int k = 0; k < n; k++) {
                const unsigned g = mix(((unsigned long long)rstep << 20) + (unsigned)k);
                const int stage = (int)(g % (unsigned)s.nstages);
                const int w = (int)((g >> 4) % (unsigned)s.numw), z = (int)((g >> 10) % (unsigned)s.numz);
                const int r = (k * s.corr_uselen) / (n > 0 ? n : 1);  // distinct r per hit
                const long long key =
                    ((((rstep * s.nstages + stage) * s.numw + w) * (long long)s.numz + z) * s.corr_uselen) + r;
                const unsigned long long idx = __atomic_fetch_add(s.nhits, 1ull, __ATOMIC_RELAXED);
                if (idx < s.maxhits) {
                    s.hit_keys[idx] = key;
                    s.hit_pows[idx] = s.powcut[stage] + (float)(g % 1000) * 0.01f;
                }
            }
        }
    });
    return hipSuccess;
}

hipError_t pzl_sort_hits(void *temp, size_t *temp_bytes, long long *keys_in, float *pows_in, long long *keys_out,
                         float *pows_out, long long n, int, hipStream_t st)
{
    if (!temp) {
        *temp_bytes = 4096;
        return hipSuccess;
    }
    enqueue(st, [=] {
        touch((char *)temp, 1);
        std::vector<std::pair<long long, float>> v((size_t)n);
        for (long long i = 0; i < n; i++) v[i] = {keys_in[i], pows_in[i]};
        std::sort(v.begin(), v.end());
        for (long long i = 0; i < n; i++) {
            keys_out[i] = v[i].first;
            pows_out[i] = v[i].second;
        }
    });
    return hipSuccess;
}

hipError_t pzl_launch_hitprep(const long long *keys, const float *pows, long long n, const double *startr,
                              int corr_uselen, int numz, int numw, int nstages, const double *, PzlHostHit *out,
                              hipStream_t st)
{
    enqueue(st, [=] {
        for (long long i = 0; i < n; i++) {
            long long k = keys[i];
            const int r = (int)(k % corr_uselen);
            k /= corr_uselen;
            const int z = (int)(k % numz);
            k /= numz;
            const int w = (int)(k % numw);
            k /= numw;
            const int stage = (int)(k % nstages);
            const long long rstep = k / nstages;
            PzlHostHit h;
            h.rr = (startr[rstep] + 0.5 * r) / (double)(1 << stage);
            h.key = keys[i];
            h.power = pows[i];
            h.sigma = 5.0f + pows[i] * 0.01f;
            h.zi = (short)z;
            h.wi = (short)w;
            h.stage = stage;
            out[i] = h;
        }
    });
    return hipSuccess;
}

hipError_t pzl_launch_add_plane(float *, int, int, int, const float *, int, int, const int *, const uint16_t *,
                                const uint16_t *, hipStream_t)
{
    return hipErrorUnknown;
}

hipError_t pzl_launch_threshold(const float *, long long, float, long long *, unsigned long long *, long long,
                                hipStream_t)
{
    return hipErrorUnknown;
}

}  // extern "C"
