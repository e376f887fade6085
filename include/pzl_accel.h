/*
 * pzl_accel.h -- C ABI of the MI355X jerk-search engine (libaccelsearch_lib.so).
 *
 * This is the thin layer between host C and the HIP kernels.  It replaces the
 * reference's extern "C" device launchers (accel_kernels.cu:59-71:
 * do_fft_batch, fuse_add_search_batch, sort_search_results,
 * init_constant_device, clear_cache, gen_kernmatrix_cu) and the host helpers
 * that fed them (accel_utils_cu.c:378-612, 857-929).  The drop-in accel.h
 * API (accelsearch_CPU1/GPU/CPU2, subharm_fderivs_vol, ...) is declared in
 * include/accel.h and is implemented on top of these entry points.
 *
 * Plain C types only: pointers, sizes, and the structs below.  Every entry
 * point returns 0 on success, PZL_ERR_OVERFLOW (1) when the hit buffer
 * overflows (the reference's too_large path, accelsearch_lib.c:748-782), or
 * a negative code; pzl_last_error() describes the last failure of the
 * calling thread.  A plan is not safe for concurrent use by two threads;
 * different plans are independent.
 */
#ifndef PZL_ACCEL_H
#define PZL_ACCEL_H

#ifdef __cplusplus
extern "C" {
#endif

#define PZL_ERR_OVERFLOW 1
#define PZL_ERR_ARG (-1)
#define PZL_ERR_HIP (-2)
#define PZL_ERR_UNSUPPORTED (-3)
#define PZL_ERR_NOSPECTRUM (-4)
#define PZL_ERR_TIMEOUT (-5)

/* Search request: the fields of Cmdline (accelsearch_cmd.h:11-87) and
 * infodata (makeinf.h:23-56) that create_accelobs (accel_utils.c:1636-2037)
 * reads.  flo/fhi/rlo/rhi follow the *P flags exactly as the reference. */
typedef struct PzlConfig {
    double N;            /* samples in the time series (.inf N) */
    double dt;           /* sample time in s (.inf dt) */
    long long numbins;   /* complex bins in the .fft file */
    long long lobin;     /* -lobin */
    int zmax, wmax;      /* -zmax, -wmax as given */
    int wmaxP;           /* -wmax given (0: acceleration-only search) */
    int numharm;         /* 1, 2, 4, 8, 16, 32 */
    float sigma;         /* -sigma */
    int floP, fhiP, rloP, rhiP;
    double flo, fhi, rlo, rhi;
    double nph;          /* > 0: photon normalisation (-photon) */
    int batchsize;       /* rsteps per launch group (-batchsize), >= 1 */
    long long maxhits;   /* hit-buffer capacity, 0 = default */
} PzlConfig;

/* Search geometry (the accelobs fields, accel.h:37-82). */
typedef struct PzlGeometry {
    long long N, numbins, lobin, highestbin, nrsteps;
    int maxkernlen, corr_uselen, fftlen, numharmstages, numz, numw, numharm_kernels;
    double dt, T, rlo, rhi, zlo, zhi, wlo, whi;
    float sigma;
    float powcut[6];
    long long numindep[6];
    long long bank_bytes, plane_bytes_per_rstep, device_bytes;
} PzlGeometry;

/* One harmonic's shape (subharminfo, accel.h:111-122). */
typedef struct PzlHarmInfo {
    int numharm, harmnum, stage, zmax, wmax, numz, numw, fftlen, halfwidth, max_numrs;
} PzlHarmInfo;

/* Candidate (the accelcand fields the search fills, accel.h:84-97). */
typedef struct PzlCand {
    float power, sigma;
    int numharm;
    double r, z, w;
} PzlCand;

/* Timing of the last search (milliseconds; HIP events on the search stream). */
typedef struct PzlStats {
    double ms_wall;        /* host wall clock of pzl_search */
    double ms_gpu;         /* first launch to last kernel end */
    double ms_segments, ms_corr, ms_sum;  /* summed kernel durations */
    double ms_corr_max_class;             /* the FFT length with the most time */
    double ms_host;        /* sort, sigma, candidate insertion */
    long long corr_launches, nhits, nrsteps;
    double cells;          /* nrsteps * corr_uselen * numz * max(numw,1) */
    double bytes_alg;      /* algorithmic HBM bytes of the search */
    double bytes_alg_corr; /* ... of the correlation kernels alone */
    int nclass;            /* correlation launches are grouped by FFT length */
    int class_fftlen[16];
    double class_ms[16];          /* summed duration of that class's launches */
    long long class_launches[16];
    double class_bytes[16];       /* algorithmic bytes of that class's launches: plane rows
                                     written and segments read per rstep, the class's kernel
                                     rows once per search */
    double ms_graph;              /* summed duration of graph-replayed batches */
    long long graph_launches;
    double class_bank_bytes[16];  /* the class's kernel rows (read once per launch) */
    double class_flops[16];       /* nominal FP32 operations of the class's launches:
                                     5 N log2 N per inverse FFT, 6 N per product with
                                     conj(kernel), 3 per stored power */
    double ms_corr_group;         /* summed duration of the grouped correlation launches (the
                                     class_ms of the lengths they hold are then 0) */
    long long group_launches;
    double ms_wait_count;  /* launching thread waiting for the oldest batch in flight (its hit count) */
    double ms_wait_pool;   /* launching thread waiting for a pinned hit-record buffer (the insertion
                              thread behind) */
} PzlStats;

const char *pzl_last_error(void);
int pzl_device_count(int *count);
/* Free and total device memory of 'device' (the cudaMemGetInfo measurement
 * that accelsearch_maxsize.c:564-566 leaves commented out). */
int pzl_device_memory(int device, long long *free_bytes, long long *total_bytes);

/* Kernel banks are cached per (device, harmonic shapes) and shared by plans;
 * unreferenced banks are evicted oldest-first above the byte budget
 * (default 64 GiB).  _clear drops every unreferenced bank. */
int pzl_bank_cache_set_budget(long long bytes);
int pzl_bank_cache_clear(void);
/* 1 if creating this plan built its bank, 0 if it came from the cache. */
int pzl_plan_bank_built(void *plan);

/* Geometry and the kernel bank on 'device'; the bank stays in HBM until
 * pzl_plan_destroy.  accel_utils.c:1636-2037 + create_subharminfos :236-282 */
int pzl_plan_create(const PzlConfig *cfg, int device, void **plan);
void pzl_plan_destroy(void *plan);
int pzl_plan_geometry(void *plan, PzlGeometry *g);
int pzl_plan_harm(void *plan, int idx, PzlHarmInfo *h);
/* Search shape only, without touching the GPU.  g->device_bytes is the HBM
 * footprint model that admission uses (replaces accelsearch_maxsize,
 * accelsearch_maxsize.c:551-619): a plan's exact allocations item by item,
 * with the sort scratch bounded before the plan exists; pzl_plan_geometry
 * reports it with the plan's actual sort scratch. */
int pzl_geometry_from_config(const PzlConfig *cfg, PzlGeometry *g);
/* Device memory the plan holds: what it allocated (its spectrum once set
 * from the host included) plus its kernel bank.  accelsearch_maxsize.c:551-619 */
int pzl_plan_device_bytes(void *plan, long long *held);

/* The spectrum: numbins complex64 values, realfft packing (read_fft.c:4-47).
 * _device takes a caller-owned device pointer that must stay valid. */
int pzl_set_spectrum(void *plan, const float *host_fft, long long numbins);
int pzl_set_spectrum_device(void *plan, const void *dev_fft, long long numbins);

/* Search rsteps [rstep_lo, rstep_hi) (rstep_hi < 0: to the end) and insert
 * the hits into a candidate list with insert_new_accelcand semantics
 * (accel_utils.c:382-480).  Writes min(ncands, maxcands) candidates. */
int pzl_search(void *plan, long long rstep_lo, long long rstep_hi, PzlCand *cands, long long maxcands,
               long long *ncands, PzlStats *stats);
/* A deadline for the plan's searches: once 'ms' milliseconds have passed
 * since this call, pzl_search launches no further batch, waits for the ones
 * in flight and returns PZL_ERR_TIMEOUT (its candidate list is then partial
 * and not returned).  ms <= 0 clears the deadline. */
int pzl_plan_set_deadline(void *plan, double ms);

/* Autotuner: times every correlation-kernel variant (FFT radices, threads and
 * kernel rows per workgroup) of each FFT length of the plan on one batch of
 * its own data and keeps the fastest, for this plan and for later plans of the
 * same shape on the device.  report (optional) receives one line per variant.
 * Replaces the reference's per-shape cuFFT plan choice (accel_kernels.cu:96-158). */
int pzl_autotune(void *plan, int reps, char *report, int report_len);
/* Graph replay: the correlations of every full batch of rsteps are launched
 * as one hipGraph (one per buffer set, built from the launches of the set's
 * first full batch); the batch's segments are launched before it (ahead on
 * their own stream, beside the previous batch's correlation, when
 * overlapped), and the harmonic sum stays a launch of its own, beside the
 * next batch.  On by
 * default; on = 0 launches every kernel one by one.  Per-kernel timing is then
 * reported per batch (ms_graph) instead of per kernel. */
int pzl_plan_set_graph(void *plan, int on);
/* CU split: the harmonic sums run on the first sum_cus compute units (spread
 * evenly over the dies and shader engines), and the segments and correlations
 * on every compute unit but the first sum_only_cus (0 <= sum_only_cus <=
 * sum_cus), so that the memory-bound sum of a batch runs beside the next
 * batch's correlation rather than in the workgroup slots it leaves.  Counts
 * are multiples of the die count, at least 4 per die (one per shader engine).
 * sum_cus = 0 (the default): every kernel may use every compute unit.  Waits
 * for the plan's work and drops its batch graphs; the autotuner chooses a
 * split for the plan.  No counterpart in the reference (its kernels run on
 * one stream). */
int pzl_plan_set_cu_split(void *plan, int sum_cus, int sum_only_cus);
int pzl_plan_cu_split(void *plan, int *sum_cus, int *sum_only_cus);
/* Overlap: on = 1: the harmonic sum of a batch runs on its own stream,
 * concurrently with the segments and correlations of the next batch (three
 * buffer sets), which stay in one stream (the arrangement of a replayed
 * graph); on = 2 (the default): also the segments of the next batch ahead on a
 * third stream, and the correlations of the shorter FFT lengths there beside
 * the longest length's (when they are not grouped); on = 0: every kernel of
 * the search in one stream. */
int pzl_plan_set_overlap(void *plan, int on);
/* Rsteps per launch: for each candidate batch size, plan, autotune and time a
 * band of band_rsteps rsteps from the middle of the search on this spectrum;
 * *best gets the fastest, ms_per_rstep (optional, ncand entries) the times.
 * Results do not depend on the batch size. */
int pzl_tune_batchsize(const PzlConfig *cfg, int device, const float *host_fft, long long numbins,
                       const int *candidates, int ncand, long long band_rsteps, int *best, double *ms_per_rstep);
/* The harmonic-sum form in use (chosen by pzl_autotune): workgroup order
 * (0: r chunk, z block, w fastest; 1: z block, w, r chunk fastest) and
 * consecutive fundamental r per thread (1, 2 or 4).  Results do not depend on
 * it. */
int pzl_plan_sum_form(void *plan, int *order, int *rpt);
int pzl_plan_set_sum_form(void *plan, int order, int rpt);
/* Correlation launches per batch (chosen by pzl_autotune): one per FFT length
 * (group -1), or one grouped launch of correlation group 'group' for the
 * shorter lengths (all = 0; the longest keeps its own launch) or for every
 * length (all = 1: one correlation launch per batch).  A group holds one
 * kernel variant per length; it must hold every length it is given.  Results
 * do not depend on it beyond the variants' rounding. */
int pzl_plan_corr_group(void *plan, int *group, int *all);
int pzl_plan_set_corr_group(void *plan, int group, int all);
int pzl_corr_num_groups(void);
const char *pzl_corr_group_name(int group);
/* The variant in use for an FFT length of the plan; set one explicitly. */
int pzl_plan_variant(void *plan, int fftlen, int *variant);
int pzl_plan_set_variant(void *plan, int fftlen, int variant);

/* The candidate list of the last pzl_search (the same as it returned). */
int pzl_get_cands(void *plan, PzlCand *cands, long long maxcands, long long *ncands);

/* Hits of the last pzl_search in PRESTO C scan order: key encodes
 * (((rstep*stages + stage)*numw + w)*numz + z)*corr_uselen + r. */
int pzl_get_hits(void *plan, long long *keys, float *powers, long long maxhits, long long *nhits);
/* How many hits a search keeps for pzl_get_hits (default 2^22; a search with
 * more keeps none and pzl_get_hits fails). */
int pzl_plan_set_keep_hits(void *plan, long long limit);
/* Hit records travel to the host in chunks of this many records (default
 * 2^20), each in a pinned buffer that the plan keeps and reuses (2 GiB in
 * all); results do not depend on it.  No counterpart in the reference (it
 * copies each batch's hits in one piece). */
int pzl_plan_set_hit_chunk(void *plan, long long records);

/* Parity hooks: one FFT'd kernel (fftlen complex64) of harmonic idx, and the
 * power plane ([numw][numz][numrs]) of (rstep, harmonic) as
 * subharm_fderivs_vol (accel_utils.c:1094-1293) returns it. */
int pzl_get_kernel(void *plan, int idx, int row, float *out);
int pzl_compute_plane(void *plan, long long rstep, int idx, float *out, long long maxfloats, int *numrs,
                      long long *rlo);
/* The same at an arbitrary fundamental startr (subharm_fderivs_vol's
 * fullrlo; fullrhi = fullrlo + corr_uselen/2 - 0.5). */
int pzl_compute_plane_at(void *plan, double fullrlo, int idx, float *out, long long maxfloats, int *numrs,
                         long long *rlo);
/* Fundamental plane of 'rstep' after all harmonic stages were added. */
int pzl_compute_sums(void *plan, long long rstep, float *out, long long maxfloats);

/* The candidate-list logic alone (host only): apply n hits in order with
 * insert_new_accelcand semantics (accel_utils.c:382-480). */
int pzl_insert_cands(const double *rr, const float *sigma, const float *power, const int *numharm,
                     const double *z, const double *w, long long n, PzlCand *out, long long maxout,
                     long long *nout);

/* ---------------- candidate optimisation (accelsearch_CPU2) --------------
 * optimize_accelcand with harmonic polishing (accel_utils.c:573-612):
 * max_rzw_arr_harmonics (maximize_rzw.c:85-191) -- the maximum of the
 * 8x-oversampled F/Fdot/Fdotdot volume around the candidate, summed over its
 * harmonics with each harmonic's local power -- or, for w = 0 searches,
 * max_rz_arr_harmonics' downhill simplex -- then, at the maximum, the local
 * power and the power/phase derivatives of every harmonic
 * (get_localpower3d / get_derivs3d, characteristics.c:75-191).  All
 * candidates are done together on the GPU (the simplex steps of all
 * candidates evaluated in one batch per round). */
#define PZL_MAX_NUMHARM 32
typedef struct PzlRDerivs { /* rderivs, presto.h:161-169 */
    double pow, phs, dpow, dphs, d2pow, d2phs, locpow;
} PzlRDerivs;
typedef struct PzlOptCand {
    /* in: the search-grid candidate; r counts bins of the spectrum given
     * (file bin minus the spectrum's first bin) */
    double r, z, w;
    int numharm;
    /* 1 for a w = 0 search: the downhill simplex in (r, z) of
     * max_rz_arr_harmonics (maximize_rz.c:141-220, amoeba.c) instead of the
     * rzw volume; w is ignored */
    int fdot_only;
    /* out: the maximum (fundamental), and per harmonic ii (number
     * ii + 1) its derivatives at (r, z, w) * (ii + 1) and its power */
    double ropt, zopt, wopt;
    PzlRDerivs derivs[PZL_MAX_NUMHARM];
    double pows[PZL_MAX_NUMHARM];
} PzlOptCand;
/* With the plan's resident spectrum (pzl_set_spectrum). */
int pzl_optimize(void *plan, PzlOptCand *cands, long long n);
/* With a host spectrum of numbins complex bins (interleaved re, im). */
int pzl_optimize_spectrum(int device, const float *host_fft, long long numbins, PzlOptCand *cands,
                          long long n);
/* One harmonic per entry (numharm must be 1), as optimize_accelcand's
 * -noharmpolish branch calls per harmonic (accel_utils.c:610-652): for w > 0
 * searches max_rzw_arr (maximize_rzw.c:3-50; fdot_only 0): the first
 * raw-power maximum of the 9 x 9 x 24 grid over z +- ACCEL_DZ, w +- ACCEL_DW
 * and three bins from int(r) - 1, derivatives there with the local power of
 * (r, z, w); for w = 0 searches max_rz_arr (maximize_rz.c:25-85; fdot_only 1):
 * the (r, z) simplex on the raw power.  Out: (ropt, zopt, wopt), derivs[0],
 * pows[0] = the maximum's raw power (the functions' return value). */
int pzl_maximize(void *plan, PzlOptCand *cands, long long n);
int pzl_maximize_spectrum(int device, const float *host_fft, long long numbins, PzlOptCand *cands,
                          long long n);

/* ---------------- request loop (continuous batching on one GPU) ----------
 * Replaces the reference's multi-process pipeline
 * (bin/accelsearch_pipeline_cu.py: pool_size processes of CPU1/GPU/CPU2
 * threads per GPU).  Requests are admitted in arrival order when their HBM
 * footprint (plan buffers + kernel bank) fits the budget beside the searches
 * in flight and one of max_inflight execution slots is free; admitted
 * searches run concurrently.  Plans are pooled per shape and reused; kernel
 * banks are shared through the bank cache. */
typedef struct PzlRequestStats {
    double ms_queue, ms_run, ms_total;       /* submit->admit, admit->done, submit->done */
    double t_submit_ms, t_admit_ms, t_end_ms; /* steady-clock time stamps */
    long long footprint, nhits;
    double cells;
    int plan_reused;
} PzlRequestStats;

typedef struct PzlServerStats {
    long long budget, admitted_bytes, peak_admitted_bytes, resident_bytes;
    long long queued, completed, pooled_plans;
    int inflight, peak_inflight;
    long long peak_resident_bytes;  /* pooled plans (busy or idle) plus plans being created */
    long long timed_out;            /* requests failed by their deadline */
} PzlServerStats;

/* A request is admitted by its footprint: the geometry's device_bytes plus
 * PZL_RUNTIME_BYTES_PER_SEARCH for the HIP runtime's own memory behind a
 * plan's streams (measured at most 27 MiB on MI355X).
 * hbm_budget <= 0: 80% of the device's free memory. */
#define PZL_RUNTIME_BYTES_PER_SEARCH (32LL << 20)
int pzl_server_create(int device, long long hbm_budget, int max_inflight, void **server);
/* Queue a search; host_fft (numbins complex64) must stay valid until the
 * request completes.  Fails if the request alone exceeds the budget. */
int pzl_server_submit(void *server, const PzlConfig *cfg, const float *host_fft, long long numbins,
                      long long *req_id);
/* The same with the spectrum in device memory of the server's device (a DM
 * trial from pzl_dedisperse + pzl_realfft_device_batch, say); it stays the
 * caller's and must stay valid until the request is done. */
int pzl_server_submit_device(void *server, const PzlConfig *cfg, const void *dev_fft, long long numbins,
                             long long *req_id);
/* The same with a deadline: 'deadline_ms' after submission the request
 * fails with PZL_ERR_TIMEOUT -- still queued, it is dropped; running, its
 * search stops launching batches, drains the ones in flight and returns --
 * and its slot and admitted bytes are freed for the next request.  The other
 * requests are not affected.  deadline_ms <= 0: none.  The reference
 * pipeline's per-thread timeout and task re-queue
 * (bin/accelsearch_pipeline_cu.py:440-490) leave the retry to the caller here. */
int pzl_server_submit_deadline(void *server, const PzlConfig *cfg, const void *fft, int on_device,
                               long long numbins, double deadline_ms, long long *req_id);
/* Block until the request is done; copy its candidates and timing and forget it. */
int pzl_server_wait(void *server, long long req_id, PzlCand *cands, long long maxcands, long long *ncands,
                    PzlRequestStats *stats);
/* Ids of completed requests not yet waited for (n: how many there are). */
int pzl_server_poll(void *server, long long *done_ids, int max_ids, int *n);
int pzl_server_stats(void *server, PzlServerStats *stats);
/* Finish every queued and running request, then free the pooled plans. */
void pzl_server_destroy(void *server);

/* ---------------- realfft: the .dat -> .fft producer ----------------------
 * realfft(idata, n, isign) (fastffts.c:163-235) on the GPU, in place on n
 * floats: isign -1 is the forward transform, n/2 bins in the Numerical
 * Recipes packing (bin 0 = (DC, Nyquist)), unnormalised; +1 the inverse, every
 * value times 2/n.  n even, n/2 a product of 2, 3, 5, 7 and 11 up to 2^26
 * (the lengths choose_good_N picks, misc_utils.c:327-366; the reference,
 * through FFTW, takes any even n).  Calls on one device run one at
 * a time and return when the transform is done. */
int pzl_realfft(float *idata, long long n, int isign);      /* host memory */
int pzl_realfft_device(int device, void *d_data, long long n, int isign);
/* nseries (<= 65535) series of n floats each, back to back, in one set of
 * launches: the DM trials of a dedispersion run (realfft per .dat file,
 * realfft.c:66-251). */
int pzl_realfft_device_batch(int device, void *d_data, long long n, int nseries, int isign);
/* The same transforms without allocating anything: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * workspace is a fixed device array of 16.5 MiB that each GPU holds until the
 * process ends, so n * nseries is at most 2^22 (one series of 4M samples, 16
 * DM trials of 2^18); larger requests fail with PZL_ERR_ARG.  The launches are
 * enqueued on 'stream' (a hipStream_t; NULL: the null stream) and the call
 * returns without waiting for them.  Every call on a GPU shares that GPU's
 * workspace: calls whose transforms overlap on one GPU (other streams, other
 * threads) race and must be ordered by the caller. */
int pzl_realfft_device_noalloc(int device, void *d_data, long long n, int nseries, int isign, void *stream);

/* ---------------- dedispersion: the DM-trial producer ---------------------
 * prepsubband_cu's GPU kernels (prepsubband_kernels.cu:23-24) on device
 * memory, synchronous, with their argument meaning and layouts:
 * DownsampDataArray (:26-46, :79-83): currentdsdata[t][s] = mean of bs
 * consecutive samples of currentdata[t bs + j][s], t < worklen, s < nsub.
 * getDedisp (:48-77, :85-90): outdata[dm][x] = sum over subbands s (ascending)
 * of the sample x + offsets[dm][s] of the block pair (lastdsdata, then
 * currentdsdata), each [worklen][nsub]; offsets [numdms][nsub] on the device,
 * each in [0, worklen] (an offset outside is an error; the reference would
 * read outside its blocks).  Results are bit-identical to the reference's. */
int pzl_downsamp_data_array(float *currentdsdata_gpu, const float *currentdata_gpu, int worklen, int bs, int nsub);
int pzl_get_dedisp(const float *currentdsdata_gpu, const float *lastdsdata_gpu, float *outdata_gpu, int worklen,
                   int nsub, int numdms, const int *offsets_gpu);
/* The same with output rows out_stride floats apart (>= worklen), so that
 * consecutive blocks land in one [numdms][series] array. */
int pzl_get_dedisp_strided(const float *currentdsdata_gpu, const float *lastdsdata_gpu, float *outdata_gpu,
                           int worklen, int nsub, int numdms, const int *offsets_gpu, long long out_stride);
/* prepsubband_cu.c's block loop (:1117-1164) over a whole subband series in
 * device memory: subbands [nsamp][nsub] (nsamp whole blocks of worklen) ->
 * out [numdms][nsamp]; output block k reads blocks k and k + 1 (zeros after
 * the last, as the reference pads a short read).  The DM trials stay in HBM
 * for pzl_realfft_device_batch and the search. */
int pzl_dedisperse(const float *subbands_gpu, long long nsamp, int nsub, int worklen, int numdms,
                   const int *offsets_gpu, float *out_gpu);

#ifdef __cplusplus
}
#endif
#endif
