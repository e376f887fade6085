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
