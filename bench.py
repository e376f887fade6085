#!/usr/bin/env python3
"""Benchmark of the MI355X jerk search (one GPU).

Workload (default): N = 2^24 float32 samples at dt = 1e-4 s, Gaussian noise
plus three jerked sinusoids, real-FFT'd to 2^23 complex64 bins; search
zmax = 200, wmax = 300, numharm = 16, sigma = 5 (11,680 rsteps).  One step is
one complete search of that spectrum; the kernel bank is built once per plan
and the spectrum is resident in HBM before timing starts.

Prints one JSON line: cells/s (nrsteps * corr_uselen * numz * numw per
second), the roofline of the dominant kernel (the 2048-point correlation),
p50/p99 per-search latency, and the oracle timed on a bounded sample of the
same workload on the host cores.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3        # MI355X published FP32 (vector, packed) peak
FP32_NONPACKED_TFLOPS = 78.6    # 256 CUs x 64 lanes x 2 (FMA) x 2.4 GHz without packed FP32


def hip_sync():
    try:
        ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()
    except OSError:
        pass


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(fft, kw, rsteps, nthreads, cells_per_rstep, start, rsteps_1core=16):
    """The oracle (PRESTO C's per-rstep loop restated, FP32) on a band of
    rsteps from the middle of the search: on 'nthreads' host threads (OpenMP
    over kernel rows, as -ncpus does) and on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    obs = O.create_accelobs(O.make_params(**kw))

    def run(nt, n):
        ctx = O.Ctx(obs, O.FP32, nthreads=nt)
        t0 = time.perf_counter()
        ctx.search(fft, start, start + n, maxhits=1 << 16)
        dt = time.perf_counter() - t0
        ctx.close()
        return dt

    dt = run(nthreads, rsteps)
    dt1 = run(1, rsteps_1core) if rsteps_1core else None
    out = {"value": cells_per_rstep * rsteps / dt, "unit": "cells/s", "cores": nthreads, "kind": "port",
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
           "sample": "oracle FP32 (PRESTO C restated, OpenMP over kernel rows) on rsteps [%d, %d) of the same "
                     "spectrum (%.1f s), kernel-bank build excluded as on the GPU" % (start, start + rsteps, dt)}
    if dt1:
        out["single_core"] = {"value": cells_per_rstep * rsteps_1core / dt1, "unit": "cells/s", "cores": 1,
                              "sample": "rsteps [%d, %d) on one thread (%.1f s)" % (start, start + rsteps_1core,
                                                                                     dt1)}
    return out


def cpu_optimize_sample(fft, cands, ncand=8):
    """The oracle's optimize_accelcand (harmonic polishing, one core) on the
    first 'ncand' candidates: seconds per candidate and mean harmonics."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    sample = cands[:ncand]
    t0 = time.perf_counter()
    for r, z, w, nh in sample:
        O.max_rzw_arr_harmonics(fft, nh, r, z, w)
    dt = time.perf_counter() - t0
    return {"ms_per_candidate": dt * 1e3 / max(len(sample), 1), "cores": 1, "kind": "port",
            "sample": "oracle max_rzw_arr_harmonics + derivatives on the first %d candidates "
                      "(mean %.1f harmonics)" % (len(sample), np.mean([c[3] for c in sample]) if sample else 0)}


def cpu_maximize_sample(fft, points, npts=64):
    """The oracle's max_rzw_arr (one harmonic, one core) on the first 'npts'
    harmonic points of the -noharmpolish optimisation."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    sample = points[:npts]
    t0 = time.perf_counter()
    for r, z, w in sample:
        O.max_rzw_arr(fft, r, z, w)
    dt = time.perf_counter() - t0
    return {"ms_per_point": dt * 1e3 / max(len(sample), 1), "cores": 1, "kind": "port",
            "sample": "oracle max_rzw_arr + derivatives on the first %d harmonic points" % len(sample)}


def copy_calibrate(mib):
    """A plain copy kernel of known size: under rocprofv3 --pmc its FETCH_SIZE
    and WRITE_SIZE per dispatch give the counters' scale (bytes moved / KiB read)."""
    import torch
    n = mib * 2 ** 20 // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "dispatches": 5}))


def serve_bench(n, rate, max_inflight, seed=4):
    """A served stream of the uneven-search workload: n requests arriving as
    Poisson at 'rate' per second, lengths uniform in 2^21..2^24 samples at
    dt = 1e-4 s, (zmax, wmax, numharm) uniform in (50, 0, 8), (200, 150, 16),
    (200, 300, 16), sigma 5.  One synthetic spectrum per length (noise, plus a
    jerked sinusoid for half of the lengths) is shared by the requests of that
    length."""
    from prestozl_amd import binding as B
    from prestozl_amd import synth
    rng = np.random.Generator(np.random.PCG64(seed))
    spectra = {}
    for lg in (21, 22, 23, 24):
        sig = synth.random_signals(rng, 1) if rng.random() < 0.5 else []
        spectra[lg] = synth.realfft_packed(synth.timeseries(1 << lg, 1e-4, sig, seed * 100 + lg))
    shapes = [(50, 0, 8), (200, 150, 16), (200, 300, 16)]
    reqs = [(int(rng.integers(21, 25)), shapes[int(rng.integers(0, 3))]) for _ in range(n)]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
    srv = B.Server(budget=0, max_inflight=max_inflight)
    ids = []
    t0 = time.perf_counter()
    for k, (lg, (z, w, nh)) in enumerate(reqs):
        dt_wait = t0 + arrivals[k] - time.perf_counter()
        if dt_wait > 0:
            time.sleep(dt_wait)
        fft = spectra[lg]
        cfg = B.make_config(N=1 << lg, dt=1e-4, numbins=len(fft), zmax=z, wmax=w, numharm=nh, sigma=5.0,
                            wmaxP=w > 0, batchsize=8)
        ids.append(srv.submit(cfg, fft))
        if (k + 1) % 32 == 0:
            print("served stream: %d of %d submitted at %.1f s" % (k + 1, n, time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
    res = [srv.wait(i) for i in ids]
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = [r[1]["ms_total"] for r in res]
    cells = sum(r[1]["cells"] for r in res)
    return {"workload": "served stream: %d requests, Poisson %.2f/s, N in 2^21..2^24, (zmax,wmax,numharm) in "
                        "(50,0,8),(200,150,16),(200,300,16), sigma 5" % (n, rate),
            "throughput": cells / wall, "unit": "cells/s", "wall_s": wall,
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "max": float(max(lat))},
            "queue_ms_mean": float(np.mean([r[1]["ms_queue"] for r in res])),
            "max_inflight": max_inflight, "peak_inflight": st["peak_inflight"],
            "peak_admitted_GiB": st["peak_admitted_bytes"] / 2 ** 30,
            "plans_reused": int(sum(r[1]["plan_reused"] for r in res))}


def dm_trials_bench(n, max_inflight):
    """The DM-trial batch after dedispersion: n files of 262,144 samples at
    dt = 6.4e-4 s with a common jerked sinusoid (37.2 Hz, z = 40, w = -80)
    whose amplitude peaks at trial 282; zmax 200, wmax 300, numharm 8, sigma 5,
    all served through the request loop (one kernel bank for every trial)."""
    from prestozl_amd import binding as B
    from prestozl_amd import synth
    files = [synth.dm_trial(i)[0] for i in range(n)]
    srv = B.Server(budget=0, max_inflight=max_inflight)
    t0 = time.perf_counter()
    ids = []
    for fft in files:
        cfg = B.make_config(N=262144, dt=6.4e-4, numbins=len(fft), zmax=200, wmax=300, numharm=8, sigma=5.0,
                            wmaxP=True, batchsize=8)
        ids.append(srv.submit(cfg, fft))
    res = [srv.wait(i) for i in ids]
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = [r[1]["ms_total"] for r in res]
    best = max(range(n), key=lambda i: max([c[1] for c in res[i][0]] or [0.0]))
    return {"workload": "DM-trial batch: %d files of N=2^18, dt=6.4e-4 s, zmax 200, wmax 300, numharm 8, sigma 5, "
                        "served" % n,
            "throughput": sum(r[1]["cells"] for r in res) / wall, "unit": "cells/s", "wall_s": wall,
            "files_per_s": n / wall, "latency_ms": {"p50": float(np.percentile(lat, 50)),
                                                     "p99": float(np.percentile(lat, 99))},
            "peak_inflight": st["peak_inflight"], "plans_created": int(n - sum(r[1]["plan_reused"] for r in res)),
            "strongest_trial": best}


def realfft_bench(reps=20, cpu=True):
    """realfft (fastffts.c:163-235), the .dat -> .fft producer, on the GPU:
    the forward transform of config 2's series (N = 2^24 samples, dt = 1e-4 s,
    noise + 3 jerked sinusoids) and of config 5's DM-trial length (2^18),
    device-resident (one synchronous call per transform), and through the
    host-memory entry point (PCIe both ways).  Algorithmic bytes per transform:
    n floats read + n floats written (8 n)."""
    from prestozl_amd import binding as B
    from prestozl_amd import synth
    hip = ctypes.CDLL("libamdhip64.so")
    res = {}
    for label, n in (("config2_N2^24", 1 << 24), ("config5_N2^18", 1 << 18)):
        sigs = [(0.05, 123.4, 40.0, -80.0), (0.03, 311.0, -120.0, 200.0), (0.08, 57.7, 10.0, 0.0)]
        x = synth.timeseries(n, 1e-4, sigs, seed=21)
        d = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(d), ctypes.c_size_t(4 * n)) == 0
        assert hip.hipMemcpy(d, x.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(4 * n), 1) == 0
        for _ in range(2):
            B.realfft_device(d.value, n, -1)
        t0 = time.perf_counter()
        for _ in range(reps):
            B.realfft_device(d.value, n, -1)
        ms = (time.perf_counter() - t0) * 1e3 / reps
        hip.hipFree(d)
        t0 = time.perf_counter()
        got = B.realfft(x, -1)
        ms_host = (time.perf_counter() - t0) * 1e3
        gbs = 8.0 * n / (ms * 1e-3) / 1e9
        r = {"n": n, "ms_per_transform": ms, "reps": reps, "alg_bytes": 8 * n,
             "achieved_GBs": gbs, "frac_hbm": gbs / HBM_PEAK_GBS, "ms_host_api": ms_host}
        if cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle as O
            t0 = time.perf_counter()
            ref = O.realfft(x, -1, O.FP32)
            r["cpu_oracle_fp32_ms"] = (time.perf_counter() - t0) * 1e3
            r["cpu_cores"] = 1
            d64 = got.astype(np.float64) - ref
            r["rms_rel_vs_oracle_fp32"] = float(np.sqrt(np.mean(d64 * d64) / np.mean(ref.astype(np.float64) ** 2)))
        res[label] = r
    # config 5: the DM trials of one dedispersion run, transformed together
    nfiles, n = 564, 1 << 18
    x = np.stack([synth.timeseries(n, 6.4e-4, [(0.1, 37.2, 40.0, -80.0)], seed=300 + i) for i in range(8)])
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), ctypes.c_size_t(4 * n * nfiles)) == 0
    for i in range(0, nfiles, 8):
        k = min(8, nfiles - i)
        assert hip.hipMemcpy(ctypes.c_void_p(d.value + 4 * n * i), x.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_size_t(4 * n * k), 1) == 0
    B.realfft_device(d.value, n, -1, nseries=nfiles)
    t0 = time.perf_counter()
    for _ in range(reps):
        B.realfft_device(d.value, n, -1, nseries=nfiles)
    ms = (time.perf_counter() - t0) * 1e3 / reps
    hip.hipFree(d)
    gbs = 8.0 * n * nfiles / (ms * 1e-3) / 1e9
    res["config5_batch_564xN2^18"] = {"n": n, "series": nfiles, "ms_per_batch": ms, "files_per_s": nfiles / ms * 1e3,
                                      "achieved_GBs": gbs, "frac_hbm": gbs / HBM_PEAK_GBS}
    res["note"] = ("device-resident: one synchronous pzl_realfft_device call per transform (column FFTs, row FFTs "
                   "written transposed, NR packing: three passes over the series); ms_host_api includes both "
                   "PCIe copies; cpu_oracle_fp32_ms is the oracle's float realfft on one core")
    return res


def dedisp_bench(cpu=True, nsearch=48, serve=True):
    """The DM-trial producer on the GPU, device-resident end to end: 64
    subbands of N = 2^18 samples (dt = 6.4e-4 s, blocks of 16384 as
    prepsubband_cu reads them) with a sinusoid dispersed at trial 282 of 564,
    dedispersed into 564 trials (getDedisp per block, prepsubband_cu.c:
    1117-1164), one batched realfft, then searches (zmax 200, wmax 300, numharm
    8, sigma 5) straight from device memory: 'nsearch' trials around 282 with
    one plan, and all 564 through the request loop.
    Algorithmic bytes of the dedispersion: the subbands read once and the
    trials written once."""
    from prestozl_amd import binding as B
    rng = np.random.default_rng(282)
    nsub, numdms, worklen, nsamp, dt = 64, 564, 16384, 1 << 18, 6.4e-4
    f = np.linspace(0.6, 1.0, nsub)
    d = 1.0 / f ** 2 - 1.0 / f[-1] ** 2
    offs = np.rint(np.linspace(0.0, 1.0, numdms)[:, None] * d[None, :] / d.max() * (worklen // 2)).astype(np.int32)
    t = np.arange(nsamp)
    sub = rng.standard_normal((nsamp, nsub)).astype(np.float32)
    for s in range(nsub):
        sub[:, s] += (0.15 * np.cos(2 * np.pi * 37.2 * (t - offs[282, s]) * dt)).astype(np.float32)
    ds, do, dser = B.DeviceArray(sub.nbytes), B.DeviceArray(offs.nbytes), B.DeviceArray(numdms * nsamp * 4)
    ds.upload(sub)
    do.upload(offs)
    B.dedisperse(ds.ptr, nsamp, nsub, worklen, numdms, do.ptr, dser.ptr)  # warm
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        B.dedisperse(ds.ptr, nsamp, nsub, worklen, numdms, do.ptr, dser.ptr)
    ms_dd = (time.perf_counter() - t0) * 1e3 / reps
    check = dser.download((2, nsamp), np.float32, offset=4 * nsamp * 282)
    t0 = time.perf_counter()
    B.realfft_device(dser.ptr, nsamp, -1, nseries=numdms)
    ms_fft = (time.perf_counter() - t0) * 1e3
    cfg = B.make_config(N=nsamp, dt=dt, numbins=nsamp // 2, zmax=200, wmax=300, numharm=8, sigma=5.0, wmaxP=True)
    plan = B.Plan(cfg)
    lo = 282 - nsearch // 2
    plan.set_spectrum_device(dser.ptr + lo * nsamp * 4, nsamp // 2)
    plan.search()  # warm
    best = (0.0, -1, 0.0)
    t0 = time.perf_counter()
    for dm in range(lo, lo + nsearch):
        plan.set_spectrum_device(dser.ptr + dm * nsamp * 4, nsamp // 2)
        cands, _ = plan.search(stats=False)
        for c in cands:
            if c[1] > best[0]:
                best = (c[1], dm, c[3] / (nsamp * dt))
    ms_search = (time.perf_counter() - t0) * 1e3 / nsearch
    plan.close()
    # every trial through the request loop, straight from device memory
    served = None
    if serve:
        served = serve_trials(B, cfg, dser, numdms, nsamp)
    alg = 4.0 * nsamp * nsub + 4.0 * nsamp * numdms
    r = {"workload": "64 subbands x N=2^18 (dt=6.4e-4 s) -> 564 DM trials (blocks of 16384), realfft of every trial, "
                     "search of %d trials from device memory" % nsearch,
         "ms_dedisperse": ms_dd, "dm_samples_per_s": numdms * nsamp / ms_dd * 1e3,
         "alg_bytes": alg, "achieved_GBs": alg / (ms_dd * 1e-3) / 1e9, "frac_hbm": alg / (ms_dd * 1e-3) / 1e9 / HBM_PEAK_GBS,
         "ms_realfft_batch": ms_fft, "ms_search_per_trial": ms_search,
         "strongest": {"sigma": best[0], "trial": best[1], "freq_hz": best[2]}}
    if served:
        served["chain_s"] = (ms_dd + ms_fft) / 1e3 + served["wall_s"]
        r["served_all_trials"] = served
    return _dedisp_cpu(r, cpu, sub, offs, worklen, check, ds, do, dser)


def serve_trials(B, cfg, dser, numdms, nsamp):
    """Every DM trial through the request loop from device memory."""
    srv = B.Server(budget=0, max_inflight=3)
    t0 = time.perf_counter()
    ids = [srv.submit_device(cfg, dser.ptr + dm * nsamp * 4, nsamp // 2) for dm in range(numdms)]
    res = [srv.wait(i) for i in ids]
    wall = time.perf_counter() - t0
    srv.close()
    lat = [x[1]["ms_total"] for x in res]
    served_best = max(range(numdms), key=lambda i: max([c[1] for c in res[i][0]] or [0.0]))
    return {"trials": numdms, "wall_s": wall, "trials_per_s": numdms / wall,
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
            "strongest_trial": served_best,
            "note": "dedispersion + batched realfft + all 564 searches (3 in flight), every spectrum read from "
                    "device memory"}


def _dedisp_cpu(r, cpu, sub, offs, worklen, check, ds, do, dser):
    """The oracle's dedispersion of one block (the CPU baseline) and a
    bit-identity check of block 7 for two trials; frees the device arrays."""
    numdms, nsamp = offs.shape[0], sub.shape[0]
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        t0 = time.perf_counter()
        got = O.dedisp(sub[worklen * 8:worklen * 9], sub[worklen * 7:worklen * 8], offs[:64])
        cms = (time.perf_counter() - t0) * 1e3
        r["cpu_baseline"] = {"ms_per_block_of_64_trials": cms, "cores": 1, "kind": "port",
                             "sample": "oracle getDedisp on one block (16384 samples, 64 subbands) for 64 trials",
                             "ms_dedisperse_scaled": cms * (numdms / 64.0) * (nsamp / worklen)}
        r["bit_identical_block_7"] = bool(np.array_equal(
            O.dedisp(sub[worklen * 8:worklen * 9], sub[worklen * 7:worklen * 8], offs[282:284]),
            check[:, worklen * 7:worklen * 8]))
    ds.free()
    do.free()
    dser.free()
    return r


def config3_bench(steps):
    """The batched form at zmax 200, wmax 500, numharm 16 (4096-point
    fundamental): N = 2^23 samples at dt = 1e-4 s with five injected jerked
    sinusoids, sigma 5.  Rsteps per launch (8 to 64) from the batch-size tuner,
    kernel variants from the autotuner, full batches replayed from hipGraphs;
    the same search at 8 rsteps per launch beside it."""
    from prestozl_amd import binding as B
    from prestozl_amd import synth
    fft, N, dt, opts, sigs = synth.config3()
    kw = dict(N=N, dt=dt, numbins=len(fft), zmax=200, wmax=500, numharm=16, sigma=5.0, wmaxP=True)
    bt, ms = B.tune_batchsize(B.make_config(batchsize=8, **kw), fft, candidates=(8, 16, 32, 64), band=128)
    res = {}
    variants = None
    for bs, graph in ((bt, True), (8, False)):
        plan = B.Plan(B.make_config(batchsize=bs, **kw))
        plan.set_spectrum(fft)
        if variants is None:
            plan.autotune(reps=2)
            variants = {f: plan.variant(f) for f in plan.fftlens()}
            group = plan.corr_group()
            split, sum_form = plan.cu_split(), plan.sum_form()
        else:  # the tuned plan's kernel variants: the same arithmetic, so the same candidates
            for f, v in variants.items():
                plan.set_variant(f, v)
            plan.set_corr_group(*group)
            plan.set_cu_split(*split)
            plan.set_sum_form(*sum_form)
        plan.set_graph(graph)
        plan.search(stats=False)
        t0 = time.perf_counter()
        for _ in range(steps):
            cands, st = plan.search(stats=False)
        res[bs, graph] = ((time.perf_counter() - t0) / steps, cands)
        g = plan.geom
        cells = g.nrsteps * g.corr_uselen * g.numz * max(g.numw, 1)
        plan.close()
    t8, c8 = res[8, False]
    tb, cb = res[bt, True]
    return {"workload": "N=2^23, dt=1e-4 s, zmax=200, wmax=500, numharm=16, sigma=5",
            "value": cells / tb, "unit": "cells/s", "ms_per_search": tb * 1e3, "batchsize": bt, "graph": True,
            "batchsize_ms_per_rstep": ms, "at_batchsize_8": {"value": cells / t8, "ms_per_search": t8 * 1e3},
            "same_candidates": cb == c8, "nrsteps": g.nrsteps, "fftlen": g.fftlen, "candidates": len(cb),
            "cu_split": dict(zip(("sum_cus", "sum_only_cus"), split))}


def measured_traffic(kernel):
    """HBM bytes per launch of 'kernel' at this workload from the committed
    rocprofv3 --pmc measurement (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)
        return t.get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def corr_group_desc(plan):
    """How a plan launches its correlations per batch."""
    from prestozl_amd import binding as B
    g, every = plan.corr_group()
    if g < 0:
        return "one launch per FFT length"
    return ("every FFT length in one launch: " if every else "the shorter FFT lengths in one launch: ") + \
        B.corr_groups()[g]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batchsize", type=int, default=8)
    ap.add_argument("--log2n", type=int, default=24)
    ap.add_argument("--cpu-rsteps", type=int, default=120)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-rstep-per-launch (plain GPU form) search")
    ap.add_argument("--serve", type=int, default=256,
                    help="requests in the served-stream leg (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=1.2, help="Poisson arrival rate of the served stream (1/s)")
    ap.add_argument("--serve-inflight", type=int, default=3)
    ap.add_argument("--dm-trials", type=int, default=564, help="files in the DM-trial leg (0: skip)")
    ap.add_argument("--no-tune", action="store_true", help="use the default correlation variants")
    ap.add_argument("--no-legs", action="store_true",
                    help="only the timed searches (no graph-replay / one-stream / plain / serving / DM-trial legs)")
    ap.add_argument("--rsteps", type=int, default=0,
                    help="profiling runs only: search this many rsteps from the middle of the band")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the served-stream leg (--serve requests at --serve-rate) and print its line")
    ap.add_argument("--realfft-only", action="store_true",
                    help="run only the realfft (.dat -> .fft) leg and print its line")
    ap.add_argument("--dedisp-only", action="store_true",
                    help="run only the dedispersion -> realfft -> search leg and print its line")
    ap.add_argument("--dedisp-no-serve", action="store_true",
                    help="with --dedisp-only: leave out the request-loop search of every trial")
    ap.add_argument("--sigma", type=float, default=5.0,
                    help="profiling runs: the search threshold (the headline search uses 5; a high one leaves almost no hits)")
    ap.add_argument("--graph", type=int, default=-1,
                    help="1: replay each full batch's segments and correlations from a captured graph, 0: launch "
                         "them one by one, -1: the plan's default")
    ap.add_argument("--overlap", type=int, default=-1,
                    help="profiling runs: 0 = one stream, 1 = harmonic sums on a side stream, 2 = also the next "
                         "batch's segments ahead and the shorter FFT lengths beside (the default)")
    ap.add_argument("--cu-split", type=int, nargs=2, default=None, metavar=("SUM", "SUM_ONLY"),
                    help="harmonic sums on the first SUM compute units, segments and correlations on all but the "
                         "first SUM_ONLY (0 0: every kernel on every CU; default: the autotuner's choice)")
    ap.add_argument("--sum-form", type=int, nargs=2, default=None, metavar=("ORDER", "RPT"),
                    help="harmonic-sum workgroup order and r per thread (default: the autotuner's choice)")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="run a plain device copy of MIB MiB five times and exit (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibrate:
        copy_calibrate(args.copy_calibrate)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each rstep and each file is "
              "independent: one process per GPU, each searching its own files or rstep range, would scale "
              "the search).", file=sys.stderr)
        sys.exit(2)

    from prestozl_amd import binding as B
    from prestozl_amd import synth

    if args.serve_only:
        print(json.dumps(serve_bench(args.serve, args.serve_rate, args.serve_inflight)), flush=True)
        return
    if args.realfft_only:
        print(json.dumps(realfft_bench(cpu=not args.no_cpu)), flush=True)
        return
    if args.dedisp_only:
        print(json.dumps(dedisp_bench(cpu=not args.no_cpu, serve=not args.dedisp_no_serve)), flush=True)
        return
    fft, N, dt, opts, sigs = synth.config2(N=1 << args.log2n)
    kw = dict(N=N, dt=dt, numbins=len(fft), zmax=opts["zmax"], wmax=opts["wmax"], numharm=opts["numharm"],
              sigma=args.sigma, wmaxP=True)
    plan = B.Plan(B.make_config(batchsize=args.batchsize, **kw))
    plan.set_spectrum(fft)
    g = plan.geom
    lat, stats_all = [], []
    lo, hi = 0, -1
    if args.rsteps:
        lo = g.nrsteps // 2
        hi = lo + args.rsteps
    tune = [] if args.no_tune else plan.autotune(reps=3)
    if args.cu_split is not None:
        plan.set_cu_split(*args.cu_split)
    if args.sum_form is not None:
        plan.set_sum_form(*args.sum_form)
    if args.overlap >= 0:
        plan.set_overlap(args.overlap)
    if args.graph >= 0:
        plan.set_graph(bool(args.graph))
    for _ in range(args.warmup):
        plan.search(lo, hi, stats=False)
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        cands, st = plan.search(lo, hi, stats=True)
        lat.append((time.perf_counter() - ts) * 1e3)
        stats_all.append(st)
    hip_sync()
    total = time.perf_counter() - t0
    cells = stats_all[0]["cells"]
    value = cells * args.steps / total
    # dominant kernel: the correlation class with the most time.  Replayed
    # graphs are timed as a whole: the per-kernel durations then come from one
    # search with the same kernels launched one by one, right after
    roof_stats, roof_note = stats_all, "kernel durations from the timed searches"
    if any(st["graph_launches"] > 0 for st in stats_all):
        # the same kernels launched one by one in the replay's arrangement:
        # segments and correlations in one stream, harmonic sums beside them
        plan.set_graph(False)
        plan.set_overlap(1)
        roof_stats = [plan.search(lo, hi, stats=True)[1]]
        plan.set_overlap(args.overlap if args.overlap >= 0 else 2)
        plan.set_graph(True)
        roof_note = "kernel durations from one search after the timed graph-replay searches, launching the same " \
                    "kernels one by one in the same arrangement (segments and correlations in one stream, " \
                    "harmonic sums beside them)"
    split = plan.cu_split()
    if split[0] > 0:
        roof_note += "; the harmonic sums confined to %d compute units that the correlation also uses (the CU " \
                     "split), so the correlation's launches share the chip with them by design" % split[0]
    cls = {}
    for st in roof_stats:
        for c in st["classes"]:
            d = cls.setdefault(c["fftlen"], dict(ms=0.0, launches=0, bytes=0.0, flops=0.0, bank=0.0))
            d["ms"] += c["ms"]
            d["launches"] += c["launches"]
            d["bytes"] += c["bytes"]
            d["flops"] += c["flops"]
            d["bank"] = c["bank_bytes"]
    # a grouped launch holds the shorter lengths (or every length): one kernel
    gid, every = plan.corr_group()
    kern = {"k_corr<%d>" % k: v for k, v in cls.items() if v["launches"] > 0}
    if gid >= 0:
        longest = max(cls)
        members = [k for k in cls if every or k != longest]
        kern["k_corr_group"] = dict(ms=sum(st["ms_corr_group"] for st in roof_stats),
                                    launches=sum(st["group_launches"] for st in roof_stats),
                                    bytes=sum(cls[k]["bytes"] for k in members),
                                    flops=sum(cls[k]["flops"] for k in members),
                                    bank=sum(cls[k]["bank"] for k in members))
    dom_name = max(kern, key=lambda k: kern[k]["ms"])
    dom = kern[dom_name]
    avg_ms = dom["ms"] / dom["launches"]
    # algorithmic bytes per launch: the plane rows written and the segments read
    # by one launch, plus the class's kernel rows amortised over the search's
    # launches (the bank is streamed once per search)
    bytes_per_launch = dom["bytes"] / dom["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    flops_per_launch = dom["flops"] / dom["launches"]
    tflops = flops_per_launch / (avg_ms * 1e-3) / 1e12
    traffic = measured_traffic(dom_name)
    # the same kernel alone on the whole chip: a band of the search with every
    # kernel in one stream (nothing runs beside the correlation), eager
    alone = None
    if dom_name == "k_corr_group" and g.nrsteps >= 2048:
        plan.set_graph(False)
        plan.set_overlap(0)
        blo = g.nrsteps // 2 - 512
        sta = plan.search(blo, blo + 1024, stats=True)[1]
        plan.set_overlap(args.overlap if args.overlap >= 0 else 2)
        plan.set_graph(args.graph != 0)
        if sta["group_launches"] > 0:
            ams = sta["ms_corr_group"] / sta["group_launches"]
            alone = {"avg_launch_ms": ams, "achieved": bytes_per_launch / (ams * 1e-3) / 1e9,
                     "frac": bytes_per_launch / (ams * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "timing": "rsteps [%d, %d) with every kernel in one stream, launched one by one "
                               "(%d launches, HIP events on every 8th)" % (blo, blo + 1024, sta["group_launches"])}
    out = {
        "metric": "f-fdot-fdotdot cells/s at zmax200/wmax300/16 harm; p50/p99 latency per search",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Gaussian noise + 3 jerked sinusoids (seeded), realfft-packed complex64",
        "config": {"workload": "plain GPU drop-in at the production shape: N=2^%d, dt=1e-4 s, zmax=200, "
                               "wmax=300, numharm=16, sigma=5, batchsize=%d" % (args.log2n, args.batchsize),
                   "nrsteps": g.nrsteps if not args.rsteps else args.rsteps, "corr_uselen": g.corr_uselen, "numz": g.numz, "numw": g.numw,
                   "cells_per_search": cells, "bank_MiB": g.bank_bytes / 2 ** 20,
                   "planes_MiB_per_rstep": g.plane_bytes_per_rstep / 2 ** 20},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "samples": len(lat)},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved, "timing": roof_note,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "avg_launch_ms": avg_ms, "alg_bytes_per_launch": bytes_per_launch,
                     "bank_bytes_reread_per_launch": dom["bank"],
                     "alone": alone,
                     "traffic_over_alg": traffic / bytes_per_launch if traffic else None,
                     "binding": "FP32 vector issue (VALU), not HBM: see compute",
                     "compute": {"flops_per_launch": flops_per_launch, "achieved": tflops, "unit": "TFLOP/s",
                                 "peak": FP32_PEAK_TFLOPS, "frac": tflops / FP32_PEAK_TFLOPS,
                                 "peak_nonpacked": FP32_NONPACKED_TFLOPS,
                                 "frac_nonpacked": tflops / FP32_NONPACKED_TFLOPS,
                                 "note": "nominal 5 N log2 N per inverse FFT + 6 N per product + 3 per power; "
                                         "packed FP32 is disabled in this build, so one FP32 FMA per lane and "
                                         "cycle (78.6 TFLOP/s) is the issue ceiling"}},
        "search_roofline": {"bytes_alg_per_search": stats_all[0]["bytes_alg"],
                            "achieved_GBs": stats_all[0]["bytes_alg"] * args.steps / total / 1e9,
                            "frac": stats_all[0]["bytes_alg"] * args.steps / total / 1e9 / HBM_PEAK_GBS},
        "kernel_ms_per_search": {"segments": np.mean([s["ms_segments"] for s in roof_stats]),
                                 "corr": np.mean([s["ms_corr"] for s in roof_stats]),
                                 "harmsum": np.mean([s["ms_sum"] for s in roof_stats]),
                                 "host": np.mean([s["ms_host"] for s in stats_all]),
                                 "host_wait_count": np.mean([s["ms_wait_count"] for s in stats_all]),
                                 "host_wait_pool": np.mean([s["ms_wait_pool"] for s in stats_all]),
                                 "by_fftlen": {str(k): v["ms"] / len(roof_stats) for k, v in sorted(cls.items())},
                                 "grouped": np.mean([s["ms_corr_group"] for s in roof_stats]),
                                 "graph_replay": np.mean([s["ms_graph"] for s in stats_all])},
        "autotune": tune,
        "variants": {str(c["fftlen"]): B.lib().pzl_corr_variant_name(c["fftlen"], plan.variant(c["fftlen"])).decode()
                     for c in stats_all[0]["classes"]},
        "corr_launches": corr_group_desc(plan),
        "cu_split": dict(zip(("sum_cus", "sum_only_cus"), plan.cu_split())),
        "sum_form": dict(zip(("order", "r_per_thread"), plan.sum_form())),
        "candidates": len(cands),
        "hits": stats_all[0]["nhits"],
    }
    # accelsearch_CPU2's candidate optimisation (harmonic polishing) of every
    # candidate of the search, one GPU batch
    opt_cands = [(c[3], c[4], c[5], c[2]) for c in cands]  # (r, z, w, numharm)
    if not args.rsteps and opt_cands:
        plan.optimize(opt_cands[:4])  # warm
        ts = time.perf_counter()
        plan.optimize(opt_cands)
        topt = time.perf_counter() - ts
        out["cpu2_optimize"] = {"candidates": len(opt_cands), "ms": topt * 1e3,
                                "candidates_per_s": len(opt_cands) / topt,
                                "mean_numharm": float(np.mean([c[3] for c in opt_cands])),
                                "note": "optimize_accelcand with harmonic polishing of every search candidate "
                                        "(a superset of CPU2's harmonic-filtered list)"}
        # -noharmpolish: every harmonic of every candidate on its own
        nhp_pts = [(r * (ii + 1), z * (ii + 1), w * (ii + 1)) for r, z, w, nh in opt_cands for ii in range(nh)]
        plan.maximize(nhp_pts[:16])  # warm
        ts = time.perf_counter()
        plan.maximize(nhp_pts)
        tmax = time.perf_counter() - ts
        out["cpu2_noharmpolish"] = {"points": len(nhp_pts), "ms": tmax * 1e3, "points_per_s": len(nhp_pts) / tmax,
                                    "note": "optimize_accelcand -noharmpolish (max_rzw_arr per harmonic) of every "
                                            "search candidate's harmonics"}
    # variants of the execution: the other launch form of full batches (one
    # by one when the timed searches replayed graphs, and the reverse), and
    # everything in one stream without overlap
    if args.no_legs:
        args.no_plain, args.serve, args.dm_trials = True, 0, 0
    if not args.no_legs:
        timed_graph = args.graph != 0  # the plans replay graphs by default
        plan.set_graph(not timed_graph)
        plan.search(lo, hi, stats=False)  # captures, when replaying
        ts = time.perf_counter()
        cg, _ = plan.search(lo, hi, stats=False)
        tg = time.perf_counter() - ts
        out["eager_launches" if timed_graph else "graph_replay"] = {
            "value": cells / tg, "unit": "cells/s", "ms_per_search": tg * 1e3, "same_candidates": cg == cands}
        plan.set_graph(False)
        plan.set_overlap(False)
        ts = time.perf_counter()
        cs, _ = plan.search(lo, hi, stats=False)
        tsr = time.perf_counter() - ts
        out["serial_one_stream"] = {"value": cells / tsr, "unit": "cells/s", "ms_per_search": tsr * 1e3,
                                    "same_candidates": cs == cands}
        plan.set_overlap(1)
        ts = time.perf_counter()
        co, _ = plan.search(lo, hi, stats=False)
        tso = time.perf_counter() - ts
        out["overlap_sum_only"] = {"value": cells / tso, "unit": "cells/s", "ms_per_search": tso * 1e3,
                                   "same_candidates": co == cands,
                                   "note": "harmonic sums on their own stream, segments and every correlation "
                                           "in the main stream"}
        plan.set_overlap(True)
    if not args.no_legs and not args.rsteps:
        # the autotuner's rsteps-per-launch choice, timed on a 64-rstep band
        bt, ms = B.tune_batchsize(B.make_config(batchsize=args.batchsize, **kw), fft, candidates=(4, 8, 16, 32))
        out["batchsize_tuning"] = {"best": bt, "ms_per_rstep": ms, "band_rsteps": 64,
                                   "note": "value above uses the configured batchsize %d" % args.batchsize}
    if not args.no_plain and not args.rsteps:
        plan1 = B.Plan(B.make_config(batchsize=1, **kw))
        plan1.set_spectrum(fft)
        for c in stats_all[0]["classes"]:  # the same kernel variants: identical arithmetic
            plan1.set_variant(c["fftlen"], plan.variant(c["fftlen"]))
        plan1.set_corr_group(*plan.corr_group())
        ts = time.perf_counter()
        c1, _ = plan1.search(stats=False)
        out["plain_gpu"] = {"value": cells / (time.perf_counter() - ts), "unit": "cells/s", "batchsize": 1,
                            "same_candidates": c1 == cands}
        plan1.close()
    plan.close()
    if args.serve and not args.rsteps:
        out["serving"] = serve_bench(args.serve, args.serve_rate, args.serve_inflight)
    if not args.no_legs and not args.rsteps:
        out["config3"] = config3_bench(1)
    if args.dm_trials and not args.rsteps:
        out["dm_trials"] = dm_trials_bench(args.dm_trials, args.serve_inflight)
    if not args.no_legs and not args.rsteps:
        out["realfft"] = realfft_bench(cpu=not args.no_cpu)
        out["dedisp"] = dedisp_bench(cpu=not args.no_cpu)
    if not args.no_cpu:
        nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(fft, kw, args.cpu_rsteps, nthreads,
                                           g.corr_uselen * g.numz * max(g.numw, 1), g.nrsteps // 2)
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
        if "cpu2_noharmpolish" in out:
            out["cpu_baseline"]["cpu2_noharmpolish"] = cpu_maximize_sample(fft, nhp_pts)
            out["cpu2_noharmpolish"]["gpu_over_cpu"] = (out["cpu_baseline"]["cpu2_noharmpolish"]["ms_per_point"] *
                                                       out["cpu2_noharmpolish"]["points_per_s"] / 1e3)
        if "cpu2_optimize" in out:
            out["cpu_baseline"]["cpu2_optimize"] = cpu_optimize_sample(fft, opt_cands)
            out["cpu2_optimize"]["gpu_over_cpu"] = (out["cpu_baseline"]["cpu2_optimize"]["ms_per_candidate"] *
                                                   out["cpu2_optimize"]["candidates_per_s"] / 1e3)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
