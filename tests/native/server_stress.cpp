// Host-logic stress of the request loop and the search driver, built with g++
// against the HIP runtime stand-in (hipstub) under AddressSanitizer or
// ThreadSanitizer.  Every request's candidate list must equal the one its
// plan gives searched alone; requests run three at a time, from host and
// "device" spectra, with plan eviction (a budget below the mix), deadlines
// that pass in the queue and while running, and repeated rounds so that
// pinned hit buffers move through the process-wide cache between plans.
// Before that, one plan per shape searches again under CU splits, hit-record
// chunk sizes and graph settings changed between its searches.
// Exit status 0 when every check holds.  Test infrastructure only.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pzl_accel.h"

static PzlConfig config(double N, double dt, int zmax, int wmax, int numharm)
{
    PzlConfig c;
    memset(&c, 0, sizeof c);
    c.N = N;
    c.dt = dt;
    c.numbins = (long long)N / 2;
    c.zmax = zmax;
    c.wmax = wmax;
    c.wmaxP = wmax > 0;
    c.numharm = numharm;
    c.sigma = 3.0f;
    c.floP = 1;
    c.flo = 1.0;
    c.fhiP = 1;
    c.fhi = 1e4;
    c.batchsize = 4;
    return c;
}

static bool same(const std::vector<PzlCand> &a, const std::vector<PzlCand> &b)
{
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); i++)
        if (a[i].power != b[i].power || a[i].sigma != b[i].sigma || a[i].numharm != b[i].numharm ||
            a[i].r != b[i].r || a[i].z != b[i].z || a[i].w != b[i].w)
            return false;
    return true;
}

static std::vector<PzlCand> alone(const PzlConfig &c, const float *fft)
{
    void *plan = nullptr;
    if (pzl_plan_create(&c, 0, &plan)) {
        fprintf(stderr, "plan: %s\n", pzl_last_error());
        exit(2);
    }
    long long n = 0;
    if (pzl_set_spectrum(plan, fft, c.numbins) || pzl_search(plan, 0, -1, nullptr, 0, &n, nullptr)) {
        fprintf(stderr, "search: %s\n", pzl_last_error());
        exit(2);
    }
    std::vector<PzlCand> v((size_t)n);
    pzl_get_cands(plan, v.data(), n, &n);
    pzl_plan_destroy(plan);
    return v;
}

int main(int argc, char **argv)
{
    setvbuf(stdout, nullptr, _IOLBF, 0);
    const int rounds = argc > 1 ? atoi(argv[1]) : 3;
    std::vector<PzlConfig> shapes = {config(1 << 17, 1e-3, 8, 0, 2), config(1 << 17, 1e-3, 10, 20, 4),
                                     config(1 << 16, 2e-3, 4, 0, 8), config(1 << 18, 5e-4, 6, 20, 2)};
    std::vector<std::vector<float>> spectra;
    std::vector<std::vector<PzlCand>> ref;
    std::vector<void *> dev;
    long long biggest = 0;
    for (auto &c : shapes) {
        std::vector<float> f(2 * c.numbins);
        for (size_t i = 0; i < f.size(); i++) f[i] = (float)((i * 2654435761u) % 1000) * 1e-3f;
        spectra.push_back(f);
        ref.push_back(alone(c, spectra.back().data()));
        void *d = nullptr;
        hipMalloc(&d, sizeof(float) * f.size());
        hipMemcpy(d, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice);
        dev.push_back(d);
        PzlGeometry g;
        pzl_geometry_from_config(&c, &g);
        biggest = std::max(biggest, g.device_bytes + PZL_RUNTIME_BYTES_PER_SEARCH);
        printf("shape %zu: %zu candidates alone\n", ref.size() - 1, ref.back().size());
    }
    int bad = 0;
    // plan options between searches of one plan: CU splits (masked streams
    // made once per split and kept), hit-record chunks from 16 records up
    // (the pool's undersized buffers go back to the cache), graphs on and off
    struct Opt {
        int sum, only;
        long long chunk;
        int graph;
    };
    const Opt opts[] = {{96, 0, 16, 1}, {64, 64, 1 << 20, 1}, {96, 0, 64, 0}, {0, 0, 16, 1}, {96, 0, 1 << 20, 1}};
    for (size_t s = 0; s < shapes.size(); s++) {
        void *plan = nullptr;
        if (pzl_plan_create(&shapes[s], 0, &plan) || pzl_set_spectrum(plan, spectra[s].data(), shapes[s].numbins)) {
            fprintf(stderr, "plan: %s\n", pzl_last_error());
            return 2;
        }
        for (const Opt &o : opts) {
            long long n = 0;
            if (pzl_plan_set_cu_split(plan, o.sum, o.only) || pzl_plan_set_hit_chunk(plan, o.chunk) ||
                pzl_plan_set_graph(plan, o.graph) || pzl_search(plan, 0, -1, nullptr, 0, &n, nullptr)) {
                fprintf(stderr, "options: %s\n", pzl_last_error());
                return 2;
            }
            std::vector<PzlCand> v((size_t)n);
            pzl_get_cands(plan, v.data(), n, &n);
            if (!same(v, ref[s])) {
                fprintf(stderr, "shape %zu split %d/%d chunk %lld graph %d: %lld candidates vs %zu alone\n", s, o.sum,
                        o.only, o.chunk, o.graph, n, ref[s].size());
                bad++;
            }
        }
        pzl_plan_destroy(plan);
    }
    printf("plan options: %s\n", bad ? "mismatch" : "identical");
    for (int round = 0; round < rounds; round++) {
        // odd rounds: room for the largest request and a little (plans are
        // evicted and rebuilt); even rounds: room for three at once
        void *srv = nullptr;
        if (pzl_server_create(0, (round % 2) ? biggest * 3 / 2 : biggest * 4, 3, &srv)) {
            fprintf(stderr, "server: %s\n", pzl_last_error());
            return 2;
        }
        struct Sub {
            long long id;
            int shape;
            bool may_time_out;
        };
        std::vector<Sub> subs;
        for (int k = 0; k < 24; k++) {
            const int s = (k * 7 + round) % (int)shapes.size();
            long long id = 0;
            int rc;
            bool dl = false;
            if (k % 5 == 3) {
                rc = pzl_server_submit_deadline(srv, &shapes[s], (k % 2) ? dev[s] : spectra[s].data(), k % 2,
                                                shapes[s].numbins, (k % 3 == 0) ? 0.05 : 2.0 * (k % 3), &id);
                dl = true;
            } else if (k % 2) {
                rc = pzl_server_submit_device(srv, &shapes[s], dev[s], shapes[s].numbins, &id);
            } else {
                rc = pzl_server_submit(srv, &shapes[s], spectra[s].data(), shapes[s].numbins, &id);
            }
            if (rc) {
                fprintf(stderr, "submit: %s\n", pzl_last_error());
                return 2;
            }
            subs.push_back({id, s, dl});
        }
        int timeouts = 0;
        for (auto &sb : subs) {
            std::vector<PzlCand> v(1 << 16);
            long long n = 0;
            PzlRequestStats rs;
            const int rc = pzl_server_wait(srv, sb.id, v.data(), (long long)v.size(), &n, &rs);
            if (rc == PZL_ERR_TIMEOUT && sb.may_time_out) {
                timeouts++;
                continue;
            }
            v.resize((size_t)n);
            if (rc || !same(v, ref[sb.shape])) {
                fprintf(stderr, "round %d request %lld (shape %d): rc %d, %lld candidates vs %zu alone\n", round,
                        sb.id, sb.shape, rc, n, ref[sb.shape].size());
                bad++;
            }
        }
        PzlServerStats st;
        pzl_server_stats(srv, &st);
        printf("round %d: %lld completed, %lld timed out (%d seen), peak inflight %d, peak resident %lld of %lld\n",
               round, st.completed, st.timed_out, timeouts, st.peak_inflight, st.peak_resident_bytes, st.budget);
        if (st.peak_resident_bytes > st.budget || st.inflight != 0 || st.admitted_bytes != 0) bad++;
        pzl_server_destroy(srv);
    }
    for (void *d : dev) hipFree(d);
    pzl_bank_cache_clear();
    printf("%s\n", bad ? "FAILED" : "OK");
    return bad ? 1 : 0;
}
