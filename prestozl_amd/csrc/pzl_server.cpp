// pzl_server.cpp -- the request loop: searches of different lengths and
// (zmax, wmax, numharm) served on one GPU with continuous batching.
//
// Replaces the reference's multi-process xargs/pool pipeline
// (bin/accelsearch_pipeline_cu.py: one CPU1/GPU/CPU2 thread triple per
// process, pool_size processes on a GPU).  Here one process keeps everything
// resident in HBM -- spectra, kernel banks (shared per shape through the bank
// cache), power planes and hit buffers (plans are pooled per shape and reused)
// -- and admits queued requests in arrival order as soon as their computed
// HBM footprint fits the budget and an execution slot is free.  Admitted
// searches run concurrently on their own streams, so one search's kernels
// fill the gaps of another's (host insertion, small FFT lengths, tails).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "pzl_accel.h"
#include "pzl_plan.h"

using namespace pzl;

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0, Clock::time_point t1)
{
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

// Every field of the request that shapes the plan (the spectrum aside).
std::vector<double> shape_key(const PzlConfig &c)
{
    return {c.N,
            c.dt,
            (double)c.numbins,
            (double)c.lobin,
            (double)c.zmax,
            (double)c.wmax,
            (double)c.wmaxP,
            (double)c.numharm,
            (double)c.sigma,
            (double)c.floP,
            (double)c.fhiP,
            (double)c.rloP,
            (double)c.rhiP,
            c.flo,
            c.fhi,
            c.rlo,
            c.rhi,
            c.nph,
            (double)c.batchsize,
            (double)c.maxhits};
}

struct Request {
    long long id = 0;
    PzlConfig cfg{};
    const float *fft = nullptr;  // caller-owned until the request completes
    bool on_device = false;      // fft is device memory of the server's device
    long long numbins = 0;
    long long footprint = 0;
    double cells = 0;
    Clock::time_point t_submit, t_admit, t_end;
    bool has_deadline = false;
    Clock::time_point deadline{};
    bool done = false;
    int status = 0;
    std::string error;
    bool plan_reused = false;
    long long nhits = 0;
    std::vector<PzlCand> cands;
};

struct PooledPlan {
    std::vector<double> key;
    void *plan = nullptr;
    long long bytes = 0;
    bool busy = false;
};

struct Server {
    int device = 0;
    long long budget = 0;
    int max_inflight = 1;
    std::mutex mu;
    std::condition_variable cv;       // queue / admission changes
    std::condition_variable done_cv;  // a request completed
    std::deque<std::shared_ptr<Request>> queue;
    std::map<long long, std::shared_ptr<Request>> reqs;
    std::vector<PooledPlan> pool;
    long long admitted_bytes = 0;  // footprints of the searches in flight
    long long resident_bytes = 0;  // all pooled plans (busy or idle) and plans being created
    long long peak_admitted = 0, peak_resident = 0, timed_out = 0;
    int inflight = 0, peak_inflight = 0;
    long long next_id = 1, completed = 0;
    bool stop = false;
    std::vector<std::thread> workers;

    // Free idle pooled plans, least recently pooled first, until 'need' more
    // bytes fit beside what is resident, then drop the kernel banks that no
    // plan references any more (each plan's footprint counted its bank; a
    // bank left in the process-wide cache would be HBM the budget no longer
    // sees).  Caller holds mu.
    void evict_idle_locked(long long need)
    {
        bool evicted = false;
        for (size_t i = 0; i < pool.size() && resident_bytes + need > budget;) {
            if (!pool[i].busy) {
                pzl_plan_destroy(pool[i].plan);
                resident_bytes -= pool[i].bytes;
                pool.erase(pool.begin() + i);
                evicted = true;
            } else {
                i++;
            }
        }
        if (evicted) (void)pzl_bank_cache_clear();
    }

    static bool expired(const Request &r, Clock::time_point now) { return r.has_deadline && now >= r.deadline; }

    // The head of the queue may start: a slot is free and its footprint fits
    // beside the searches in flight (idle pooled plans are evictable).
    bool head_admissible_locked() const
    {
        if (queue.empty() || inflight >= max_inflight) return false;
        return admitted_bytes + queue.front()->footprint <= budget;
    }

    // Requests whose deadline passed while queued fail without running.
    // Caller holds mu.
    bool drop_expired_locked()
    {
        const auto now = Clock::now();
        bool any = false;
        for (auto it = queue.begin(); it != queue.end();) {
            if (expired(**it, now)) {
                Request &r = **it;
                r.status = PZL_ERR_TIMEOUT;
                r.error = "deadline passed before the request was admitted";
                r.t_admit = r.t_end = now;
                r.done = true;
                completed++;
                timed_out++;
                it = queue.erase(it);
                any = true;
            } else {
                ++it;
            }
        }
        return any;
    }

    void worker()
    {
        for (;;) {
            std::shared_ptr<Request> r;
            void *plan = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if (drop_expired_locked()) done_cv.notify_all();
                    if (head_admissible_locked() || (stop && queue.empty())) break;
                    // wake up now and then to fail queued requests past their deadline
                    cv.wait_for(lk, std::chrono::milliseconds(20));
                }
                if (!head_admissible_locked()) return;  // stopping, nothing left to admit
                r = queue.front();
                queue.pop_front();
                inflight++;
                peak_inflight = std::max(peak_inflight, inflight);
                admitted_bytes += r->footprint;
                peak_admitted = std::max(peak_admitted, admitted_bytes);
                r->t_admit = Clock::now();
                const auto key = shape_key(r->cfg);
                for (auto &p : pool)
                    if (!p.busy && p.key == key) {
                        p.busy = true;
                        plan = p.plan;
                        r->plan_reused = true;
                        break;
                    }
                if (!plan) {
                    // reserve the new plan's bytes in the same critical section as
                    // the eviction, so that two workers creating plans at once
                    // both see each other's share of the budget
                    evict_idle_locked(r->footprint);
                    resident_bytes += r->footprint;
                    peak_resident = std::max(peak_resident, resident_bytes);
                }
            }
            int rc = 0;
            if (!plan) {
                rc = pzl_plan_create(&r->cfg, device, &plan);
                std::lock_guard<std::mutex> lk(mu);
                if (rc == 0) {
                    PooledPlan p;
                    p.key = shape_key(r->cfg);
                    p.plan = plan;
                    p.bytes = r->footprint;  // counted at the reservation
                    p.busy = true;
                    pool.push_back(p);
                } else {
                    resident_bytes -= r->footprint;
                    plan = nullptr;
                }
            }
            if (rc == 0 && r->has_deadline) {
                const double left = std::chrono::duration<double, std::milli>(r->deadline - Clock::now()).count();
                rc = left > 0 ? pzl_plan_set_deadline(plan, left)
                              : fail(PZL_ERR_TIMEOUT, "deadline passed before the search started");
            }
            if (rc == 0)
                rc = r->on_device ? pzl_set_spectrum_device(plan, r->fft, r->numbins)
                                  : pzl_set_spectrum(plan, r->fft, r->numbins);
            PzlStats st{};
            long long n = 0;
            if (rc == 0) rc = pzl_search(plan, 0, -1, nullptr, 0, &n, &st);
            if (rc == 0) {
                r->cands.resize((size_t)n);
                rc = pzl_get_cands(plan, r->cands.data(), n, &n);
            }
            if (plan) (void)pzl_plan_set_deadline(plan, 0);
            {
                std::lock_guard<std::mutex> lk(mu);
                r->status = rc;
                if (rc) r->error = pzl_last_error();
                if (rc == PZL_ERR_TIMEOUT) timed_out++;
                r->nhits = st.nhits;
                r->cells = st.cells;
                r->t_end = Clock::now();
                r->done = true;
                completed++;
                for (auto &p : pool)
                    if (p.plan == plan) p.busy = false;
                if (rc && rc != PZL_ERR_TIMEOUT && plan) {
                    // a failed plan is not reused (a timed-out search has drained: its plan is)
                    for (size_t i = 0; i < pool.size(); i++)
                        if (pool[i].plan == plan) {
                            pzl_plan_destroy(plan);
                            resident_bytes -= pool[i].bytes;
                            pool.erase(pool.begin() + i);
                            break;
                        }
                }
                admitted_bytes -= r->footprint;
                inflight--;
            }
            cv.notify_all();
            done_cv.notify_all();
        }
    }
};

}  // namespace

extern "C" {

int pzl_server_create(int device, long long hbm_budget, int max_inflight, void **srv)
{
    if (!srv || max_inflight < 1) return fail(PZL_ERR_ARG, "bad server arguments");
    int ndev = 0;
    PZL_HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(PZL_ERR_ARG, "no device %d", device);
    auto *s = new Server();
    s->device = device;
    if (hbm_budget <= 0) {
        size_t free_b = 0, total_b = 0;
        hipError_t e = hipSetDevice(device);
        if (e == hipSuccess) e = hipMemGetInfo(&free_b, &total_b);
        if (e != hipSuccess) {
            delete s;
            return fail(PZL_ERR_HIP, "hipMemGetInfo failed: %s", hipGetErrorString(e));
        }
        hbm_budget = (long long)(free_b * 0.8);
    }
    s->budget = hbm_budget;
    s->max_inflight = max_inflight;
    for (int i = 0; i < max_inflight; i++) s->workers.emplace_back([s] { s->worker(); });
    *srv = s;
    return 0;
}

static int submit(void *srv, const PzlConfig *cfg, const float *fft, bool on_device, long long numbins,
                  long long *req_id, double deadline_ms = 0)
{
    Server *s = (Server *)srv;
    if (!s || !cfg || !fft || numbins <= 0) return fail(PZL_ERR_ARG, "bad request");
    PzlGeometry g;
    int rc = pzl_geometry_from_config(cfg, &g);
    if (rc) return rc;
    auto r = std::make_shared<Request>();
    r->cfg = *cfg;
    r->fft = fft;
    r->on_device = on_device;
    r->numbins = numbins;
    // plan buffers, resident spectrum and kernel bank (counted per request,
    // although requests of one shape share a cached bank), and the runtime's
    // memory behind the plan's streams
    r->footprint = g.device_bytes + PZL_RUNTIME_BYTES_PER_SEARCH;
    r->t_submit = Clock::now();
    if (deadline_ms > 0) {
        r->has_deadline = true;
        r->deadline = r->t_submit + std::chrono::duration_cast<Clock::duration>(
                                        std::chrono::duration<double, std::milli>(deadline_ms));
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return fail(PZL_ERR_ARG, "server is shutting down");
        if (r->footprint > s->budget)
            return fail(PZL_ERR_ARG, "request needs %lld bytes of HBM, above the server budget %lld", r->footprint,
                        s->budget);
        r->id = s->next_id++;
        s->reqs[r->id] = r;
        s->queue.push_back(r);
    }
    s->cv.notify_all();
    if (req_id) *req_id = r->id;
    return 0;
}

int pzl_server_submit(void *srv, const PzlConfig *cfg, const float *host_fft, long long numbins, long long *req_id)
{
    return submit(srv, cfg, host_fft, false, numbins, req_id);
}

int pzl_server_submit_device(void *srv, const PzlConfig *cfg, const void *dev_fft, long long numbins,
                             long long *req_id)
{
    return submit(srv, cfg, (const float *)dev_fft, true, numbins, req_id);
}

int pzl_server_submit_deadline(void *srv, const PzlConfig *cfg, const void *fft, int on_device, long long numbins,
                               double deadline_ms, long long *req_id)
{
    return submit(srv, cfg, (const float *)fft, on_device != 0, numbins, req_id, deadline_ms);
}

int pzl_server_wait(void *srv, long long req_id, PzlCand *cands, long long maxcands, long long *ncands,
                    PzlRequestStats *st)
{
    Server *s = (Server *)srv;
    if (!s) return fail(PZL_ERR_ARG, "null server");
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(req_id);
        if (it == s->reqs.end()) return fail(PZL_ERR_ARG, "unknown request %lld", req_id);
        r = it->second;
        s->done_cv.wait(lk, [&] { return r->done; });
        s->reqs.erase(req_id);
    }
    if (ncands) *ncands = (long long)r->cands.size();
    for (long long k = 0; cands && k < (long long)r->cands.size() && k < maxcands; k++) cands[k] = r->cands[k];
    if (st) {
        memset(st, 0, sizeof *st);
        st->ms_queue = ms_since(r->t_submit, r->t_admit);
        st->ms_run = ms_since(r->t_admit, r->t_end);
        st->ms_total = ms_since(r->t_submit, r->t_end);
        st->t_submit_ms = std::chrono::duration<double, std::milli>(r->t_submit.time_since_epoch()).count();
        st->t_admit_ms = std::chrono::duration<double, std::milli>(r->t_admit.time_since_epoch()).count();
        st->t_end_ms = std::chrono::duration<double, std::milli>(r->t_end.time_since_epoch()).count();
        st->footprint = r->footprint;
        st->nhits = r->nhits;
        st->cells = r->cells;
        st->plan_reused = r->plan_reused;
    }
    if (r->status) return fail(r->status, "request %lld failed: %s", req_id, r->error.c_str());
    return 0;
}

int pzl_server_poll(void *srv, long long *done_ids, int max_ids, int *n)
{
    Server *s = (Server *)srv;
    if (!s) return fail(PZL_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    int k = 0;
    for (auto &kv : s->reqs)
        if (kv.second->done) {
            if (done_ids && k < max_ids) done_ids[k] = kv.first;
            k++;
        }
    if (n) *n = k;
    return 0;
}

int pzl_server_stats(void *srv, PzlServerStats *st)
{
    Server *s = (Server *)srv;
    if (!s || !st) return fail(PZL_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    memset(st, 0, sizeof *st);
    st->budget = s->budget;
    st->admitted_bytes = s->admitted_bytes;
    st->peak_admitted_bytes = s->peak_admitted;
    st->resident_bytes = s->resident_bytes;
    st->queued = (long long)s->queue.size();
    st->inflight = s->inflight;
    st->peak_inflight = s->peak_inflight;
    st->completed = s->completed;
    st->pooled_plans = (long long)s->pool.size();
    st->peak_resident_bytes = s->peak_resident;
    st->timed_out = s->timed_out;
    return 0;
}

void pzl_server_destroy(void *srv)
{
    Server *s = (Server *)srv;
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    for (auto &t : s->workers) t.join();
    for (auto &p : s->pool) pzl_plan_destroy(p.plan);
    delete s;
}

}  // extern "C"
