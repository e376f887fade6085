"""prepsubband's GPU dedispersion (DownsampDataArray / getDedisp,
prepsubband_kernels.cu:26-90) on the GPU against the oracle's restatement:
bit-identical (the same float sums in the same order) on random data, and the
device-resident chain subbands -> DM trials -> spectra -> search with no host
round trip."""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


def _offsets(numdms, nsub, maxoff, rng):
    f = np.linspace(0.6, 1.0, nsub)  # subband 0 lowest; the top subband has no delay
    d = 1.0 / f ** 2 - 1.0 / f[-1] ** 2
    dm = np.sort(rng.uniform(0.0, 1.0, numdms))
    return np.rint(dm[:, None] * d[None, :] / d.max() * maxoff).astype(np.int32)


@pytest.mark.parametrize("bs", [1, 2, 3, 8])
def test_downsamp_bit_identical(bs):
    from prestozl_amd import binding as B
    rng = np.random.default_rng(bs)
    worklen, nsub = 3000, 48
    x = rng.standard_normal((worklen * bs, nsub)).astype(np.float32)
    din, dout = B.DeviceArray(x.nbytes), B.DeviceArray(worklen * nsub * 4)
    din.upload(x)
    B.downsamp_data_array(dout.ptr, din.ptr, worklen, bs, nsub)
    got = dout.download((worklen, nsub), np.float32)
    assert np.array_equal(got, O.downsamp(x, bs))


@pytest.mark.parametrize("worklen,nsub,numdms", [(1000, 16, 7), (4096, 64, 33), (2500, 96, 5), (600, 300, 5)])
def test_dedisp_bit_identical(worklen, nsub, numdms):
    from prestozl_amd import binding as B
    rng = np.random.default_rng(worklen + nsub)
    last = rng.standard_normal((worklen, nsub)).astype(np.float32)
    cur = rng.standard_normal((worklen, nsub)).astype(np.float32)
    offs = _offsets(numdms, nsub, worklen, rng)
    offs[0, :] = 0
    offs[-1, 0] = worklen  # both ends of the allowed range
    dl, dc = B.DeviceArray(last.nbytes), B.DeviceArray(cur.nbytes)
    do, dout = B.DeviceArray(offs.nbytes), B.DeviceArray(numdms * worklen * 4)
    dl.upload(last)
    dc.upload(cur)
    do.upload(offs)
    B.get_dedisp(dc.ptr, dl.ptr, dout.ptr, worklen, nsub, numdms, do.ptr)
    got = dout.download((numdms, worklen), np.float32)
    assert np.array_equal(got, O.dedisp(cur, last, offs))


@pytest.mark.parametrize("worklen,nsub,numdms", [(3000, 24, 19), (1024, 8, 8)])
def test_dedisp_scattered_offsets_bit_identical(worklen, nsub, numdms):
    """Unsorted offsets over the whole [0, worklen]: DM groups whose offsets in
    a subband spread wider than the shared window read that subband directly,
    others share it; both give the oracle's sums."""
    from prestozl_amd import binding as B
    rng = np.random.default_rng(worklen + 7)
    last = rng.standard_normal((worklen, nsub)).astype(np.float32)
    cur = rng.standard_normal((worklen, nsub)).astype(np.float32)
    offs = rng.integers(0, worklen + 1, (numdms, nsub)).astype(np.int32)
    offs[:, ::3] = rng.integers(0, 40, (numdms, len(range(0, nsub, 3))))  # narrow subbands too
    dl, dc = B.DeviceArray(last.nbytes), B.DeviceArray(cur.nbytes)
    do, dout = B.DeviceArray(offs.nbytes), B.DeviceArray(numdms * worklen * 4)
    dl.upload(last)
    dc.upload(cur)
    do.upload(offs)
    B.get_dedisp(dc.ptr, dl.ptr, dout.ptr, worklen, nsub, numdms, do.ptr)
    assert np.array_equal(dout.download((numdms, worklen), np.float32), O.dedisp(cur, last, offs))


def test_dedisp_strided_rows_and_bad_offsets():
    from prestozl_amd import binding as B
    rng = np.random.default_rng(3)
    worklen, nsub, numdms, stride = 512, 8, 4, 1500
    last = rng.standard_normal((worklen, nsub)).astype(np.float32)
    cur = rng.standard_normal((worklen, nsub)).astype(np.float32)
    offs = _offsets(numdms, nsub, worklen, rng)
    dl, dc, do = B.DeviceArray(last.nbytes), B.DeviceArray(cur.nbytes), B.DeviceArray(offs.nbytes)
    dl.upload(last)
    dc.upload(cur)
    do.upload(offs)
    sentinel = np.full((numdms, stride), 7.5, np.float32)
    dout = B.DeviceArray(sentinel.nbytes)
    dout.upload(sentinel)
    B.get_dedisp(dc.ptr, dl.ptr, dout.ptr + 4 * 100, worklen, nsub, numdms, do.ptr, out_stride=stride)
    got = dout.download((numdms, stride), np.float32)
    want = sentinel.copy()
    want[:, 100:100 + worklen] = O.dedisp(cur, last, offs)
    assert np.array_equal(got, want)
    for bad in (worklen + 1, -1):
        o2 = offs.copy()
        o2[2, 3] = bad
        do.upload(o2)
        with pytest.raises(B.PzlError):
            B.get_dedisp(dc.ptr, dl.ptr, dout.ptr, worklen, nsub, numdms, do.ptr)


def test_dedisperse_series_equals_blockwise_oracle():
    """The whole-series block loop: output block k from blocks (k, k + 1), the
    last with a zero block, into one [numdms][nsamp] array."""
    from prestozl_amd import binding as B
    rng = np.random.default_rng(4)
    worklen, nblk, nsub, numdms = 2048, 5, 32, 12
    nsamp = worklen * nblk
    sub = rng.standard_normal((nsamp, nsub)).astype(np.float32)
    offs = _offsets(numdms, nsub, worklen, rng)
    ds, do, dout = B.DeviceArray(sub.nbytes), B.DeviceArray(offs.nbytes), B.DeviceArray(numdms * nsamp * 4)
    ds.upload(sub)
    do.upload(offs)
    B.dedisperse(ds.ptr, nsamp, nsub, worklen, numdms, do.ptr, dout.ptr)
    got = dout.download((numdms, nsamp), np.float32)
    zero = np.zeros((worklen, nsub), np.float32)
    for k in range(nblk):
        last = sub[k * worklen:(k + 1) * worklen]
        cur = sub[(k + 1) * worklen:(k + 2) * worklen] if k + 1 < nblk else zero
        assert np.array_equal(got[:, k * worklen:(k + 1) * worklen], O.dedisp(cur, last, offs)), k


def test_device_chain_subbands_to_candidates():
    """A sinusoid dispersed across 32 subbands with the delays of DM trial 9:
    dedispersion of 16 trials, one batched realfft and a search of the
    strongest trial's spectrum straight from device memory.  The trial and the
    frequency come back; the trials' spectra match the oracle's (dedispersion,
    then a double-precision realfft) to float FFT accuracy."""
    from prestozl_amd import binding as B
    rng = np.random.default_rng(9)
    worklen, nblk, nsub, numdms, dstar = 8192, 16, 32, 16, 9
    nsamp, dt = worklen * nblk, 6.4e-4
    offs = _offsets(numdms, nsub, worklen // 4, rng)
    r0 = 20345.25  # Fourier bin of the signal
    t = np.arange(nsamp + worklen)
    sub = (rng.standard_normal((nsamp, nsub)) * 2.0).astype(np.float32)
    for s in range(nsub):
        sub[:, s] += (0.5 * np.cos(2 * np.pi * r0 * (t[:nsamp] - offs[dstar, s]) / nsamp)).astype(np.float32)
    ds, do, dser = B.DeviceArray(sub.nbytes), B.DeviceArray(offs.nbytes), B.DeviceArray(numdms * nsamp * 4)
    ds.upload(sub)
    do.upload(offs)
    B.dedisperse(ds.ptr, nsamp, nsub, worklen, numdms, do.ptr, dser.ptr)
    series = dser.download((numdms, nsamp), np.float32)
    B.realfft_device(dser.ptr, nsamp, -1, nseries=numdms)
    spec = dser.download((numdms, nsamp), np.float32)
    for d in (0, dstar, numdms - 1):
        exact = O.realfft(series[d], -1, O.FP64)
        err = spec[d].astype(np.float64) - exact
        assert np.sqrt(np.mean(err ** 2) / np.mean(exact ** 2)) < 1e-6, d
    pw = spec.view(np.complex64)
    peak = np.abs(pw[:, int(r0)]) ** 2 + np.abs(pw[:, int(r0) + 1]) ** 2
    assert int(np.argmax(peak)) == dstar
    cfg = B.make_config(N=nsamp, dt=dt, numbins=nsamp // 2, zmax=10, wmax=0, numharm=2, sigma=3.0, wmaxP=False)
    plan = B.Plan(cfg)
    plan.set_spectrum_device(dser.ptr + dstar * nsamp * 4, nsamp // 2)
    cands, _ = plan.search()
    plan.close()
    best = max(cands, key=lambda c: c[1])
    print("best candidate: sigma %.1f, r %.2f, z %.2f, numharm %d" % (best[1], best[3], best[4], best[2]))
    assert abs(best[3] - r0) < 2.0 and abs(best[4]) <= 2.0
