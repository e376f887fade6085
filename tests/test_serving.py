"""The request loop (pzl_server_*): requests of different lengths and shapes
admitted by HBM footprint and run concurrently on one GPU.

Each request's candidate list must be the same served in a mix as searched
alone (identical values: the same kernels run on the same data, only the
interleaving on the device differs); admission must respect the byte budget
and the arrival order; with room for several, searches must overlap.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(config1_data):
    from prestozl_amd import binding as B
    from prestozl_amd import synth
    fft1, N1, dt1, _ = config1_data
    rng = np.random.Generator(np.random.PCG64(77))
    x = synth.timeseries(1 << 19, 1e-4, synth.random_signals(rng, 2, zspan=30.0, wspan=60.0, amp=(0.1, 0.2)), 7)
    fft2 = synth.realfft_packed(x)
    shapes = [(20, 0, 4, True), (40, 60, 8, True), (10, 0, 2, True), (30, 40, 4, True), (50, 0, 8, False),
              (20, 0, 4, True), (24, 20, 16, True), (40, 60, 8, True)]
    reqs = []
    for i, (zmax, wmax, nh, wp) in enumerate(shapes):
        fft, N, dt = (fft1, N1, dt1) if i % 2 == 0 else (fft2, 1 << 19, 1e-4)
        cfg = B.make_config(N=N, dt=dt, numbins=len(fft), zmax=zmax, wmax=wmax, numharm=nh, sigma=3.0,
                            wmaxP=wp, batchsize=4)
        reqs.append((cfg, fft))
    return reqs


def _alone(cfg, fft):
    from prestozl_amd import binding as B
    plan = B.Plan(cfg)
    plan.set_spectrum(fft)
    cands, _ = plan.search(stats=False)
    plan.close()
    return cands


def test_same_alone_and_in_mix(config1_data):
    from prestozl_amd import binding as B
    reqs = _requests(config1_data)
    alone = [_alone(cfg, fft) for cfg, fft in reqs]
    srv = B.Server(budget=0, max_inflight=3)
    ids = [srv.submit(cfg, fft) for cfg, fft in reqs]
    got = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    for k, ((cands, rs), ref) in enumerate(zip(got, alone)):
        print("request %d: %d candidates, queue %.1f ms, run %.1f ms, reused plan %d" %
              (k, len(cands), rs["ms_queue"], rs["ms_run"], rs["plan_reused"]))
        assert len(ref) > 0
        assert cands == ref, k
    print("server:", st)
    assert st["completed"] == len(reqs)
    assert st["peak_inflight"] >= 2  # searches overlapped on the device
    assert sum(r[1]["plan_reused"] for r in got) >= 1  # shapes repeat: pooled plans were reused


def test_admission_by_footprint(config1_data):
    from prestozl_amd import binding as B
    reqs = _requests(config1_data)
    foot = []
    for cfg, fft in reqs:
        g = B.geometry(cfg)
        foot.append(g.device_bytes + B.RUNTIME_BYTES_PER_SEARCH)
    budget = int(max(foot) * 1.5)  # room for one large request and at most a small one beside it
    srv = B.Server(budget=budget, max_inflight=4)
    ids = [srv.submit(cfg, fft) for cfg, fft in reqs]
    res = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    print("budget %d, peak admitted %d, peak inflight %d" % (budget, st["peak_admitted_bytes"], st["peak_inflight"]))
    assert st["peak_admitted_bytes"] <= budget
    assert st["peak_admitted_bytes"] >= max(foot)
    admits = [r[1]["t_admit_ms"] for r in res]
    assert admits == sorted(admits)  # arrival order
    for (cands, _), (cfg, fft) in zip(res, reqs):
        assert cands == _alone(cfg, fft)


def test_request_above_budget_is_refused(config1_data):
    from prestozl_amd import binding as B
    cfg, fft = _requests(config1_data)[0]
    srv = B.Server(budget=1 << 20, max_inflight=1)
    with pytest.raises(B.PzlError):
        srv.submit(cfg, fft)
    srv.close()


def test_device_spectra_in_mix_equal_host(config1_data):
    """Spectra already in device memory (submitted with
    pzl_server_submit_device), served in a mix with host-memory requests of
    the same shapes, give the same candidates as alone."""
    from prestozl_amd import binding as B
    reqs = _requests(config1_data)[:6]
    alone = [_alone(cfg, fft) for cfg, fft in reqs]
    devs = []
    for cfg, fft in reqs:
        a = np.ascontiguousarray(fft, dtype=np.complex64)
        d = B.DeviceArray(a.nbytes)
        d.upload(a)
        devs.append(d)
    srv = B.Server(budget=0, max_inflight=3)
    ids = []
    for k, ((cfg, fft), d) in enumerate(zip(reqs, devs)):
        ids.append(srv.submit_device(cfg, d.ptr, len(fft)) if k % 2 == 0 else srv.submit(cfg, fft))
    got = [srv.wait(i) for i in ids]
    srv.close()
    for k, ((cands, _), ref) in enumerate(zip(got, alone)):
        assert cands == ref, k
    for d in devs:
        d.free()


# ---------------------------------------------------------------------------
# Served results against the oracle, the served stream's shapes, deadlines and
# the HBM actually held.

def _oracle_cands(cfg_kw, fft):
    import pyoracle as O
    okw = {k: v for k, v in cfg_kw.items() if k not in ("batchsize", "maxhits")}
    obs = O.create_accelobs(O.make_params(**okw))
    ctx = O.Ctx(obs, O.FP32, nthreads=16)
    co, _, _ = ctx.search(fft, maxhits=0)
    ctx.close()
    return co


def _config4_requests():
    """The served stream's three shapes at N = 2^21 (dt = 1e-4 s), each on an
    r band of about 8 rsteps (-rlo/-rhi) around the injected jerked sinusoid
    (211.3 Hz: mean r about 44,300), sigma 5."""
    from prestozl_amd import synth
    N, dt = 1 << 21, 1e-4
    fft = synth.realfft_packed(synth.timeseries(N, dt, [(0.05, 211.3, 30.0, -40.0)], seed=41))
    out = []
    for zmax, wmax, nh in ((50, 0, 8), (200, 150, 16), (200, 300, 16)):
        kw = dict(N=N, dt=dt, numbins=len(fft), zmax=zmax, wmax=wmax, numharm=nh, sigma=5.0, wmaxP=wmax > 0,
                  rlo=41000.0, rhi=47000.0, batchsize=8)
        out.append((kw, fft))
    return out


def test_served_requests_match_oracle(config1_data):
    """Every request of the mixed stream above and the served stream's three
    shapes: the candidate list served in the mix equals the oracle's (apart
    from candidates within 1e-4 sigma of the threshold, counted)."""
    from prestozl_amd import binding as B
    from paritylib import compare_cands
    reqs = []
    for cfg, fft in _requests(config1_data):
        kw = dict(N=cfg.N, dt=cfg.dt, numbins=cfg.numbins, zmax=cfg.zmax, wmax=cfg.wmax, numharm=cfg.numharm,
                  sigma=cfg.sigma, wmaxP=bool(cfg.wmaxP), batchsize=cfg.batchsize)
        reqs.append((kw, fft))
    reqs += _config4_requests()
    srv = B.Server(budget=0, max_inflight=3)
    ids = [srv.submit(B.make_config(**kw), fft) for kw, fft in reqs]
    got = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["peak_inflight"] >= 2
    for k, ((kw, fft), (cands, rs)) in enumerate(zip(reqs, got)):
        co = _oracle_cands(kw, fft)
        removed = compare_cands(cands, co, kw["sigma"])
        print("request %d (zmax %d, wmax %d, numharm %d): %d served / %d oracle candidates, %d near threshold "
              "removed, run %.1f ms" % (k, kw["zmax"], kw["wmax"], kw["numharm"], len(cands), len(co), removed,
                                        rs["ms_run"]))
        assert len(co) > 0


def _long_request():
    """zmax 200 / wmax 300 / 16 harmonics over N = 2^21 (1,460 rsteps)."""
    from prestozl_amd import binding as B
    from prestozl_amd import synth
    fft, N, dt, opts, sigs = synth.config2(N=1 << 21)
    return B.make_config(N=N, dt=dt, numbins=len(fft), zmax=200, wmax=300, numharm=16, sigma=5.0, wmaxP=True,
                         batchsize=8), fft


def test_plan_deadline_stops_search_and_plan_stays_usable():
    """A deadline that passes during a search stops it (PZL_ERR_TIMEOUT) after
    the batches in flight drain; the same plan then searches normally and
    gives the candidates of an undisturbed search."""
    import time
    from prestozl_amd import binding as B
    cfg, fft = _long_request()
    plan = B.Plan(cfg)
    plan.set_spectrum(fft)
    t0 = time.perf_counter()
    ref, _ = plan.search(stats=False)
    t_full = time.perf_counter() - t0
    plan.set_deadline(t_full * 1e3 / 3)
    t0 = time.perf_counter()
    with pytest.raises(B.PzlError) as e:
        plan.search(stats=False)
    t_cut = time.perf_counter() - t0
    print("full search %.1f ms, stopped after %.1f ms: %s" % (t_full * 1e3, t_cut * 1e3, e.value))
    assert e.value.rc == B.PZL_ERR_TIMEOUT
    assert t_cut < 0.8 * t_full
    plan.set_deadline(0)
    again, _ = plan.search(stats=False)
    assert again == ref
    plan.close()


def test_server_deadline_fails_only_that_request(config1_data):
    """Requests with a deadline that passes (one while queued behind a long
    search, one while running) fail with PZL_ERR_TIMEOUT; their slots and
    bytes are freed; the other requests complete with their alone results."""
    from prestozl_amd import binding as B
    cfg_long, fft_long = _long_request()
    small = _requests(config1_data)[:3]
    alone_long = _alone(cfg_long, fft_long)
    alone_small = [_alone(cfg, fft) for cfg, fft in small]
    srv = B.Server(budget=0, max_inflight=1)
    i_long = srv.submit(cfg_long, fft_long)
    i_queued = srv.submit(cfg_long, fft_long, deadline_ms=5.0)  # expires behind the first
    i_small = [srv.submit(cfg, fft) for cfg, fft in small]
    i_plain = srv.submit(cfg_long, fft_long, deadline_ms=0.0)  # no deadline
    got_long = srv.wait(i_long)
    with pytest.raises(B.PzlError) as e:
        srv.wait(i_queued)
    assert e.value.rc == B.PZL_ERR_TIMEOUT
    got_small = [srv.wait(i) for i in i_small]
    assert srv.wait(i_plain)[0] == alone_long
    # a deadline shorter than the search itself: it passes while running
    t_run = got_long[1]["ms_run"]
    i_cut = srv.submit(cfg_long, fft_long, deadline_ms=max(t_run / 4, 1.0))
    with pytest.raises(B.PzlError) as e2:
        srv.wait(i_cut)
    assert e2.value.rc == B.PZL_ERR_TIMEOUT
    after = srv.submit(*small[0])
    assert srv.wait(after)[0] == alone_small[0]
    st = srv.stats()
    srv.close()
    print("server after deadlines:", st)
    assert got_long[0] == alone_long
    assert [g[0] for g in got_small] == alone_small
    assert st["timed_out"] == 2
    assert st["inflight"] == 0 and st["admitted_bytes"] == 0 and st["queued"] == 0


def test_hbm_held_within_budget(config1_data):
    """Mixed shapes, three in flight, a budget of 1.5 x the largest request:
    the bytes the server counts stay within the budget, and the device memory
    the process actually holds (sampled every few ms while the requests run)
    never grows by more than the budget plus 64 MiB over what it held before."""
    import threading
    import time
    from prestozl_amd import binding as B
    reqs = _requests(config1_data) + [(B.make_config(**kw), fft) for kw, fft in _config4_requests()]
    foot = [B.geometry(cfg).device_bytes + B.RUNTIME_BYTES_PER_SEARCH for cfg, _ in reqs]
    budget = int(max(foot) * 1.5)
    B.bank_cache_clear()
    free0, _ = B.device_memory()
    low = [free0]
    stop = threading.Event()

    def sample():
        while not stop.is_set():
            low[0] = min(low[0], B.device_memory()[0])
            time.sleep(0.003)

    th = threading.Thread(target=sample)
    th.start()
    srv = B.Server(budget=budget, max_inflight=3)
    try:
        ids = [srv.submit(cfg, fft) for cfg, fft in reqs + reqs[::-1]]
        for i in ids:
            srv.wait(i)
        st = srv.stats()
    finally:
        stop.set()
        th.join()
        srv.close()
    held = free0 - low[0]
    print("budget %.1f MiB, peak counted resident %.1f MiB, peak admitted %.1f MiB, peak held %.1f MiB"
          % (budget / 2 ** 20, st["peak_resident_bytes"] / 2 ** 20, st["peak_admitted_bytes"] / 2 ** 20,
             held / 2 ** 20))
    assert st["peak_resident_bytes"] <= budget
    assert st["peak_admitted_bytes"] <= budget
    assert held <= budget + (64 << 20)
