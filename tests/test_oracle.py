"""Pinning the oracle (CPU only).

The oracle restates PRESTO C; each test checks one of its functions against
something that does not share its code: the reference's own known answers
(tests/test_presto_python.py:118-143), the reference's statistics / Fresnel /
median sources compiled into oracle/_ref (when the reference tree is present),
scipy/numpy library routines, closed forms and brute force.
"""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.special as sp
import scipy.stats as st

import pyoracle as O

REF_SO = os.path.join(os.path.dirname(O.__file__), "_ref", "libpresto_ref.so")


def ref_lib():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (the reference tree is not present)")
    L = C.CDLL(REF_SO)
    D, I = C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.fresnl.argtypes = [C.c_double, D, D]
    L.median.argtypes = [C.POINTER(C.c_float), C.c_int]
    L.median.restype = C.c_float
    for fn in ("cdfchi", "cdfchi2"):
        getattr(L, fn).argtypes = [I, D, D, D, D, I, D]
    for fn in ("cdfnor", "cdfnor2"):
        getattr(L, fn).argtypes = [I, D, D, D, D, D, I, D]
    return L


# ---------------- responses: the reference's known answers ----------------

def test_response_known_answers():
    """tests/test_presto_python.py:118-143 (numbetween 16, z=5, w=40)."""
    nb, z, w = 16, 5.0, 40.0
    bps = max(16, O.z_resp_halfwidth(z), O.w_resp_halfwidth(z, w))
    nn = nb * bps * 2
    for mode in (O.FP32, O.FP64):
        pr = np.abs(O.response("r", 0.0, nb, nn, mode=mode)) ** 2
        pz = np.abs(O.response("z", 0.0, nb, z, nn, mode=mode)) ** 2
        pw = np.abs(O.response("w", 0.0, nb, z, w, nn, mode=mode)) ** 2
        assert pr[nn // 2] == 1.0
        assert round(pz[nn // 2] - 0.227675, 6) == 0
        assert round(pw[nn // 2] - 0.019462, 6) == 0


def test_r_response_closed_form():
    """gen_r_response = sinc(pi x) e^{i pi x} sampled at x = numkern/(2 nb) - m/nb."""
    nb, nk = 2, 64
    resp = O.response("r", 0.0, nb, nk, mode=O.FP64)
    x = nk / (2.0 * nb) - np.arange(nk) / nb
    ref = np.sinc(x) * np.exp(1j * np.pi * x)
    assert np.abs(resp - ref).max() < 1e-14


@pytest.mark.parametrize("z", [-30.0, -6.0, 4.0, 12.0, 50.0])
def test_z_response_matches_chirp_dft(z):
    """Two independent formulas for an f-dot response: the Fresnel closed form
    (gen_z_response) and the DFT of a synthesised chirp with a vanishing
    f-dot-dot (gen_w_response with |w| just above its 1e-4 switch).  A sign,
    index or phase slip in either separates them by O(1)."""
    nb = 2
    nk = 2 * nb * O.z_resp_halfwidth(z)
    rz = O.response("z", 0.0, nb, z, nk, mode=O.FP64)
    rw = O.response("w", 0.0, nb, z, 2e-4, nk, mode=O.FP64)
    assert np.abs(rz - rw).max() < 0.02 * np.abs(rz).max()
    # Parseval: interbinned powers sum to ~numbetween for a unit-amplitude tone
    assert abs((np.abs(rz) ** 2).sum() / nb - 1.0) < 0.05


def test_w_response_is_dft_of_chirp():
    """gen_w_response: brute-force FP64 DFT of the same synthesised chirp."""
    nb, z, w, nk = 2, 4.0, 60.0, 48
    hw = O.w_resp_halfwidth(z, w)
    npts = O.lib().oracle_next2_to_n(6 * hw + 200 + nk // nb)
    t = np.arange(npts) / npts
    fbar = npts // 4
    f, fd, fdd = fbar - 0.5 * z + w / 12.0, (z - 0.5 * w) / 2.0, w / 6.0
    data = (2.0 / npts) * np.cos(2 * np.pi * (t * (t * (t * fdd + fd) + f)))
    k = fbar * nb - nk // 2 + np.arange(nk)
    n = npts * nb
    dft = np.exp(-2j * np.pi * np.outer(k, np.arange(npts)) / n) @ data
    resp = O.response("w", 0.0, nb, z, w, nk, mode=O.FP64)
    assert np.abs(resp - dft).max() < 1e-12 * np.abs(dft).max() * npts


@pytest.mark.parametrize("x", [0.0, 1e-3, 0.3, 0.9, 1.49, 1.5, 1.51, 2.2, 4.0, 9.5, 33.0, 400.0, -0.7, -6.1])
def test_fresnel_vs_scipy(x):
    s, c = O.fresnel(x)
    S, Cc = sp.fresnel(x)
    tol = 2e-15 if abs(x) < 30 else 1e-14  # cos/sin of pi x^2 / 2 lose bits at large x
    assert abs(s - S) < tol and abs(c - Cc) < tol


def test_fresnel_vs_reference_cephes():
    # Cephes returns exactly 1/2 above |x| = 36974 (fresnl.c:190-194); the
    # kernels never get there (|x| <= sqrt(2/|z|) * (numkern/4 + |z|/2) < 100).
    L = ref_lib()
    for x in np.concatenate([np.linspace(-8, 8, 401), [20.0, 100.0, 1000.0, 30000.0]]):
        s, c = O.fresnel(x)
        rs, rc = C.c_double(), C.c_double()
        L.fresnl(float(x), C.byref(rs), C.byref(rc))
        assert abs(s - rs.value) < 1e-14 and abs(c - rc.value) < 1e-14


# ---------------- FFTs ------------------------------------------------------

@pytest.mark.parametrize("n", [2, 3, 5, 128, 192, 384, 640, 768, 1280, 2048, 4096])
def test_cfft_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.abs(O.cfft(x, -1, O.FP64) - np.fft.fft(x)).max() < 1e-12 * n
    assert np.abs(O.cfft(x, 1, O.FP64) - np.fft.ifft(x) * n).max() < 1e-12 * n
    f32 = O.cfft(x, -1, O.FP32)
    assert np.abs(f32 - np.fft.fft(x)).max() < 2e-6 * np.abs(np.fft.fft(x)).max()


def test_realfft_packing():
    """fastffts.c realfft(-1): bins 1..n/2-1 of the DFT, bin 0 = (DC, Nyquist)."""
    rng = np.random.default_rng(7)
    for n in (16, 1024, 4096):
        x = rng.standard_normal(n)
        a = x.copy()
        O.lib().oracle_realfft_fwd_d(O.dptr(a), n)
        X = np.fft.rfft(x)
        packed = a[0::2] + 1j * a[1::2]
        assert abs(packed[0].real - X[0].real) < 1e-10 and abs(packed[0].imag - X[n // 2].real) < 1e-10
        assert np.abs(packed[1:] - X[1:n // 2]).max() < 1e-10


def test_realfft_roundtrip_known_answer():
    """The reference's own known answer (tests/test_presto_python.py:10-13):
    rfft(rfft(x, -1), 1) == x to 1e-6 at N = 20, in float; and at odd n / 2
    (n % 4 == 2), mixed-radix and power-of-two lengths, float and double."""
    rng = np.random.default_rng(11)
    for n in (20, 22, 30, 1000, 3000, 4096):
        x = rng.standard_normal(n)
        y = O.realfft(O.realfft(x, -1, O.FP32), 1, O.FP32)
        assert np.allclose(x, y, atol=1e-6 * np.sqrt(np.log2(n) / np.log2(20)) * 3), n
        y = O.realfft(O.realfft(x, -1, O.FP64), 1, O.FP64)
        assert np.abs(x - y).max() < 1e-12, n


def test_realfft_inverse_is_numpy_irfft():
    """realfft(+1) of an NR-packed spectrum is numpy's irfft of the unpacked
    one: the 2/n scale of an n/2-point inverse FFT is irfft's 1/n.  Bin n/4
    is where the reference flips signs around its unpacking loop (n % 4 == 0)."""
    rng = np.random.default_rng(12)
    for n in (20, 22, 64, 1000, 3000):
        X = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
        X[0] = X[0].real
        X[n // 2] = X[n // 2].real
        packed = np.empty(n)
        packed[0], packed[1] = X[0].real, X[n // 2].real
        packed[2::2], packed[3::2] = X[1:n // 2].real, X[1:n // 2].imag
        got = O.realfft(packed, 1, O.FP64)
        assert np.abs(got - np.fft.irfft(X, n)).max() < 1e-12, n


def test_realfft_closed_forms():
    """cos(2 pi f t / n) -> n/2 at bin f (real); sin -> -n/2 (imaginary); a
    constant c -> n c in bin 0's real part; (-1)^t -> n in bin 0's imaginary
    part (the Nyquist slot)."""
    for n in (24, 1000, 4096):
        t = np.arange(n)
        for f in (1, n // 4, n // 2 - 1):
            a = O.realfft(np.cos(2 * np.pi * f * t / n), -1, O.FP64)
            b = O.realfft(np.sin(2 * np.pi * f * t / n), -1, O.FP64)
            ea = np.zeros(n); ea[2 * f] = n / 2
            eb = np.zeros(n); eb[2 * f + 1] = -n / 2
            tol = 1e-12 * n  # the double trig recursion of the unpacking over n/4 steps
            assert np.abs(a - ea).max() < tol and np.abs(b - eb).max() < tol, (n, f)
        a = O.realfft(np.full(n, 0.75), -1, O.FP64)
        assert abs(a[0] - 0.75 * n) < tol and np.abs(a[1:]).max() < tol
        a = O.realfft((-1.0) ** t, -1, O.FP64)
        assert abs(a[1] - n) < tol and abs(a[0]) < tol and np.abs(a[2:]).max() < tol


# ---------------- dedispersion (prepsubband_kernels.cu:26-77) ----------------

def test_downsamp_closed_form():
    """A ramp in[t][s] = t nsub + s averaged over bs samples: the mean of bs
    consecutive t is (t' bs + (bs - 1) / 2) nsub + s (exact in float here); bs
    = 1 is the identity."""
    worklen, nsub = 50, 12
    for bs in (1, 2, 4, 5):
        t = np.arange(worklen * bs)[:, None]
        x = (t * nsub + np.arange(nsub)[None, :]).astype(np.float32)
        ds = O.downsamp(x, bs)
        tp = np.arange(worklen)[:, None]
        want = ((tp * bs + (bs - 1) / 2.0) * nsub + np.arange(nsub)[None, :]).astype(np.float32)
        assert np.array_equal(ds, want), bs


def _offsets(numdms, nsub, worklen, rng):
    # delays grow towards low frequency (subband 0) and vanish at the top
    # subband, as prepsubband_cu.c:456-466 builds them
    dmfac = rng.uniform(0.0, 1.0, numdms)
    f = np.linspace(0.6, 1.0, nsub)  # subband 0 lowest
    d = (1.0 / f ** 2 - 1.0 / f[-1] ** 2)
    return np.rint(dmfac[:, None] * d[None, :] / d.max() * (worklen - 1)).astype(np.int32)


def test_dedisp_impulse_across_the_block_boundary():
    """A pulse dispersed with the offsets of one DM (in the previous block for
    some subbands, the current for others) sums to nsub at its own DM and
    output time, and no other DM of different offsets reaches nsub there."""
    rng = np.random.default_rng(5)
    worklen, nsub, numdms = 400, 16, 9
    offs = _offsets(numdms, nsub, worklen, rng)
    dm, x0 = 4, 250
    last = np.zeros((worklen, nsub), np.float32)
    cur = np.zeros((worklen, nsub), np.float32)
    for s in range(nsub):
        t = x0 + offs[dm, s]
        (last if t < worklen else cur)[t % worklen if t >= worklen else t, s] = 1.0
    out = O.dedisp(cur, last, offs)
    assert out[dm, x0] == nsub
    assert out[dm].sum() == nsub  # every subband's impulse lands once in this DM's series
    for d in range(numdms):
        if not np.array_equal(offs[d], offs[dm]):
            assert out[d, x0] < nsub


def test_dedisp_constants_and_zero_offsets():
    rng = np.random.default_rng(6)
    worklen, nsub, numdms = 300, 10, 5
    offs = _offsets(numdms, nsub, worklen, rng)
    ones = np.ones((worklen, nsub), np.float32)
    assert np.array_equal(O.dedisp(ones, ones, offs), np.full((numdms, worklen), nsub, np.float32))
    last = rng.integers(-50, 50, (worklen, nsub)).astype(np.float32)
    cur = rng.integers(-50, 50, (worklen, nsub)).astype(np.float32)
    out = O.dedisp(cur, last, np.zeros((2, nsub), np.int32))
    assert np.array_equal(out[0], last.sum(axis=1)) and np.array_equal(out[1], out[0])
    out = O.dedisp(cur, last, np.full((1, nsub), worklen, np.int32))  # offset = worklen reads the current block
    assert np.array_equal(out[0], cur.sum(axis=1))


def test_dedisp_blocks_equal_the_whole_series():
    """prepsubband_cu.c feeds consecutive blocks as (last, current): block k of
    the output is sum_s series[s][k worklen + x + off_s] over the whole series
    (integer-valued data: exact in any order)."""
    rng = np.random.default_rng(7)
    worklen, nsub, numdms, nblk = 128, 8, 6, 5
    offs = _offsets(numdms, nsub, worklen, rng)
    series = rng.integers(-20, 20, ((nblk + 1) * worklen, nsub)).astype(np.float32)
    for k in range(nblk):
        last, cur = series[k * worklen:(k + 1) * worklen], series[(k + 1) * worklen:(k + 2) * worklen]
        out = O.dedisp(cur, last, offs)
        for d in range(numdms):
            idx = k * worklen + np.arange(worklen)[:, None] + offs[d][None, :]
            want = series[idx, np.arange(nsub)[None, :]].sum(axis=1)
            assert np.array_equal(out[d], want), (k, d)


# ---------------- statistics -------------------------------------------------

def test_median_is_lower_order_statistic():
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 8, 1024, 1025):
        a = rng.exponential(size=n).astype(np.float32)
        assert O.median(a) == np.sort(a)[(n - 1) // 2]


def test_median_vs_reference():
    L = ref_lib()
    rng = np.random.default_rng(4)
    for n in (5, 64, 1024, 2047):
        a = rng.exponential(size=n).astype(np.float32)
        b = a.copy()
        assert O.median(a) == L.median(b.ctypes.data_as(C.POINTER(C.c_float)), n)


@pytest.mark.parametrize("numsum", [1, 2, 4, 8, 16, 32])
def test_power_for_sigma_vs_scipy(numsum):
    for sigma, ntrials in ((2.0, 1e7), (5.0, 1.4e10), (8.0, 3e3)):
        p = O.power_for_sigma(sigma, numsum, ntrials)
        ref = st.chi2.isf(st.norm.sf(sigma) / ntrials, 2 * numsum) / 2.0
        assert abs(p - ref) < 1e-9 * ref


def _ref_power_for_sigma(L, sigma, numsum, numtrials):
    which, status = C.c_int(1), C.c_int(0)
    p, q, x, mean, sd, bound = (C.c_double(v) for v in (0, 0, sigma, 0, 1, 0))
    L.cdfnor(C.byref(which), C.byref(p), C.byref(q), C.byref(x), C.byref(mean), C.byref(sd), C.byref(status),
             C.byref(bound))
    q = C.c_double(q.value / numtrials)
    p = C.c_double(1.0 - q.value)
    which, df, status = C.c_int(2), C.c_double(2.0 * numsum), C.c_int(0)
    L.cdfchi(C.byref(which), C.byref(p), C.byref(q), C.byref(x), C.byref(df), C.byref(status), C.byref(bound))
    assert status.value == 0
    return 0.5 * x.value


def test_power_for_sigma_vs_reference_dcdflib():
    L = ref_lib()
    for numsum in (1, 2, 4, 8, 16):
        for sigma, ntrials in ((2.0, 1.2e7), (5.0, 1.4e10), (3.0, 5e8)):
            a = O.power_for_sigma(sigma, numsum, ntrials)
            b = _ref_power_for_sigma(L, sigma, numsum, ntrials)
            assert abs(a - b) < 2e-7 * b  # dcdflib's root finder stops at ~1e-8


def _ref_candidate_sigma(L, power, numsum, numtrials):
    """characteristics.c:647-668 composed from the reference's dcdflib."""
    chi2, dof = 2.0 * power, 2.0 * numsum
    if chi2 / dof > 15.0 or (dof > 150 and chi2 / dof > 6.0):
        return None  # asymptotic branch: pure formula, checked separately
    which, status = C.c_int(1), C.c_int(0)
    p, q, x, df, bound = (C.c_double(v) for v in (0, 0, chi2, dof, 0))
    L.cdfchi2(C.byref(which), C.byref(p), C.byref(q), C.byref(x), C.byref(df), C.byref(status), C.byref(bound))
    logp = np.log(q.value) + np.log(numtrials)
    if logp < -600 or logp > np.log(0.5):
        return None  # below: A&S formula (checked elsewhere); above: not a detection (sigma 0)
    q2 = np.exp(logp)
    which, p2, q2c, x2, mean, sd = C.c_int(2), C.c_double(1.0 - q2), C.c_double(q2), C.c_double(0), C.c_double(0), \
        C.c_double(1)
    L.cdfnor2(C.byref(which), C.byref(p2), C.byref(q2c), C.byref(x2), C.byref(mean), C.byref(sd), C.byref(status),
              C.byref(bound))
    return max(x2.value, 0.0)


def test_candidate_sigma_vs_reference_dcdflib():
    L = ref_lib()
    n = 0
    for numsum in (1, 2, 4, 8, 16):
        for power in np.linspace(numsum * 2.0, numsum * 15.0, 37):
            ref = _ref_candidate_sigma(L, power, numsum, 1.4e10)
            if ref is None:
                continue
            a = O.candidate_sigma(power, numsum, 1.4e10)
            assert abs(a - ref) < 1e-9 * max(ref, 1.0), (numsum, power, a, ref)
            n += 1
    assert n > 60


def test_candidate_sigma_inverts_power_for_sigma():
    for numsum in (1, 2, 4, 8, 16):
        for sigma in (2.0, 5.0, 7.5):
            p = O.power_for_sigma(sigma, numsum, 1e9)
            if 2 * p / (2 * numsum) <= 15.0:  # exact branch
                assert abs(O.candidate_sigma(p, numsum, 1e9) - sigma) < 1e-7


def test_candidate_sigma_vs_scipy():
    # exact branch (chi2/dof <= 15): gratio2's Poisson sum and the normal quantile
    for numsum in (1, 4, 16):
        for power in (numsum * 3.0, numsum * 6.0, numsum * 12.0):
            logp = st.chi2.logsf(2 * power, 2 * numsum)
            ref = st.norm.isf(np.exp(logp))
            assert abs(O.candidate_sigma(power, numsum, 1.0) - max(ref, 0)) < 1e-8
    # asymptotic branch (A&S 26.4.19 / 6.1.41): an expansion, close to the exact tail
    for numsum in (1, 4, 16):
        for power in (numsum * 20.0, numsum * 30.0, numsum * 45.0):
            logp = st.chi2.logsf(2 * power, 2 * numsum) + np.log(1e8)
            ref = st.norm.isf(np.exp(logp))
            assert np.isfinite(ref) and ref > 0
            assert abs(O.candidate_sigma(power, numsum, 1e8) - ref) < 2e-3 * ref


# ---------------- geometry: the closed forms of the search shape ----------

APPENDIX_A = [  # harm, zmax_h, wmax_h, numz, numw, fftlen, halfw, numrs(first rstep)
    ((1, 1), 200, 300, 201, 31, 2048, 153, 1436), ((2, 1), 100, 160, 101, 17, 1280, 85, 720),
    ((4, 1), 50, 80, 51, 9, 768, 50, 362), ((4, 3), 150, 220, 151, 23, 2048, 118, 1078),
    ((8, 1), 26, 40, 27, 5, 384, 33, 184), ((8, 3), 76, 120, 77, 13, 1024, 68, 540),
    ((8, 5), 126, 180, 127, 19, 2048, 101, 900), ((8, 7), 176, 260, 177, 27, 2048, 136, 1258),
    ((16, 1), 12, 20, 13, 3, 192, 24, 92), ((16, 3), 38, 60, 39, 7, 512, 42, 272),
    ((16, 5), 62, 100, 63, 11, 768, 59, 452), ((16, 7), 88, 140, 89, 15, 1024, 77, 632),
    ((16, 9), 112, 160, 113, 17, 1280, 92, 810), ((16, 11), 138, 200, 139, 21, 2048, 110, 988),
    ((16, 13), 162, 240, 163, 25, 2048, 127, 1170), ((16, 15), 188, 280, 189, 29, 2048, 145, 1348),
]


def test_config2_geometry_table():
    """Per-harmonic shapes at zmax 200 / wmax 300 / numharm 16 (from the
    sizing formulas accel_utils.c:137-149, 1134-1146, responses.c:49-84)."""
    obs = O.create_accelobs(O.make_params(1 << 24, 1e-4, 1 << 23, 200, 300, 16, 5.0))
    assert (obs.fftlen, obs.corr_uselen, obs.maxkernlen) == (2048, 1436, 612)
    assert O.num_rsteps(obs) == 11680
    ctx = None
    for (nh, hn), zmax, wmax, numz, numw, fftlen, halfw, numrs in APPENDIX_A:
        h = O.harm_info(obs, nh, hn)
        assert (h.zmax, h.wmax, h.numz, h.numw, h.fftlen, h.halfwidth) == (zmax, wmax, numz, numw, fftlen, halfw)
    # numrs of the first rstep follows accel_utils.c:1134-1146
    startr = obs.rlo
    for (nh, hn), *_, numrs in APPENDIX_A:
        hf = hn / nh
        drlo = np.rint(2 * startr * hf) / 2
        drhi = np.rint(2 * (startr + 717.5) * hf) / 2
        n = int((np.ceil(drhi) - np.floor(drlo)) * 2 + 1e-14) + 1
        if (nh, hn) == (1, 1):
            n = obs.corr_uselen
        elif n % 2:
            n = (n // 2 + 1) * 2
        assert n == numrs
    assert ctx is None


def test_powcut_config1():
    obs = O.create_accelobs(O.make_params(1 << 20, 64e-6, 1 << 19, 20, 0, 4, 2.0))
    assert [round(x, 2) for x in list(obs.powcut)[:3]] == [18.80, 21.21, 25.45]
    assert (obs.fftlen, obs.corr_uselen, obs.numz, obs.numw) == (2048, 1940, 21, 0)


# ---------------- correlation: brute force on a tiny case -----------------

def test_subharm_fderivs_vol_vs_direct_correlation():
    """One (w, z) row of the FP64 plane equals a direct circular correlation
    of the spread, normalised segment with the placed kernel."""
    rng = np.random.default_rng(11)
    nbins = 8192
    fft = (rng.standard_normal(nbins) + 1j * rng.standard_normal(nbins)).astype(np.complex64)
    fft[3000] += 40.0
    obs = O.create_accelobs(O.make_params(2 * nbins, 1e-3, nbins, 20, 40, 2, 3.0))
    ctx = O.Ctx(obs, O.FP64, 2)
    startr = obs.rlo + 2 * (obs.corr_uselen // 2)
    pows, _, _, meta = ctx.subharm(fft.view(np.float32), 1, 1, startr, startr + obs.corr_uselen // 2 - 0.5)
    h = O.harm_info(obs, 1, 1)
    N = h.fftlen
    lobin = int(meta[0]) - h.halfwidth
    seg = fft[lobin:lobin + N // 2].astype(np.complex128)
    pw = np.abs(seg) ** 2
    norm = 1.0 / np.sqrt(np.sort(pw)[(len(pw) - 1) // 2] / np.log(2.0))
    x = np.zeros(N, dtype=np.complex128)
    x[0::2] = seg * norm
    for wi, zi in ((0, 0), (h.numw // 2, h.numz // 2), (h.numw - 1, 3)):
        kfft = O.kernel(obs, 1, 1, zi, wi, O.FP64)
        kp = np.fft.ifft(kfft)  # the placed time-domain kernel
        # y[n] = sum_m x[(n+m) mod N] conj(kp[m])
        idx = (np.arange(N)[:, None] + np.arange(N)[None, :]) % N
        y = (x[idx] * np.conj(kp)[None, :]).sum(axis=1) * N
        off = 2 * h.halfwidth
        ref = np.abs(y[off:off + pows.shape[2]]) ** 2 / N ** 2
        assert np.abs(pows[wi, zi] - ref).max() < 1e-9 * max(ref.max(), 1.0)
    ctx.close()


# ---------------- committed reference vectors (tests/golden) ---------------

def _golden():
    import json
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.json")) as f:
        return json.load(f)


def test_golden_fresnel_and_median():
    g = _golden()
    for x, s, c in g["fresnl"]:
        os_, oc = O.fresnel(x)
        assert abs(os_ - s) < 1e-14 and abs(oc - c) < 1e-14, x
    for arr, m in g["median"]:
        assert O.median(np.array(arr, dtype=np.float32)) == np.float32(m)


def test_golden_power_for_sigma():
    for sigma, numsum, ntrials, p in _golden()["power_for_sigma"]:
        assert abs(O.power_for_sigma(sigma, numsum, ntrials) - p) < 2e-7 * p, (sigma, numsum)


def test_golden_candidate_sigma():
    g = _golden()["candidate_sigma"]
    assert len(g) > 40
    for power, numsum, ntrials, s in g:
        assert abs(O.candidate_sigma(power, numsum, ntrials) - s) < 1e-9 * max(s, 1.0), (power, numsum)


# ---------------- kernel placement (init_kernel, accel_utils.c:151-168) ------

@pytest.mark.parametrize("nh,hn,zi,wi", [(1, 1, 0, 0), (1, 1, 7, 3), (2, 1, 5, 1), (4, 3, 0, 2)])
def test_kernel_is_fft_of_wrap_placed_response(nh, hn, zi, wi):
    """The FFT'd kernel's inverse is the w/z response of numkern = 4 * halfwidth
    points with bin-zero response (index numkern/2) at index 0: the second
    half at the start, the first half at the end, zeros between
    (place_complex_kernel, corr_prep.c:92-114; forward FFT, sign -1)."""
    obs = O.create_accelobs(O.make_params(1 << 20, 1e-4, 1 << 19, 40, 100, 4, 3.0))
    h = O.harm_info(obs, nh, hn)
    z = -h.zmax + 2 * zi
    w = -h.wmax + 20 * wi
    hw = O.w_resp_halfwidth(float(z), float(w))
    numkern = 4 * hw
    resp = O.response("w", 0.0, 2, float(z), float(w), numkern, mode=O.FP64)
    kfft = O.kernel(obs, nh, hn, zi, wi, O.FP64)
    data = np.fft.ifft(kfft)
    N = h.fftlen
    half = numkern // 2
    exp = np.zeros(N, dtype=np.complex128)
    exp[:half] = resp[half:]
    exp[N - half:] = resp[:half]
    assert np.abs(data - exp).max() < 1e-12 * np.abs(resp).max()


# ---------------- harmonic summing and candidate insertion -----------------

def _nearest_even_half(x):
    """nearest multiple of 0.5, ties to the even multiple (rint of 2x)."""
    return np.round(2.0 * x) / 2.0


def _nearest_away(x, step):
    """nearest multiple of 'step', ties away from zero (NEAREST_INT)."""
    q = x / step
    return np.sign(q) * np.floor(np.abs(q) + 0.5) * step


def test_summed_plane_brute_force():
    """The oracle's summed fundamental plane equals, bit for bit, float32
    additions in the reference's order (stage 1..S, odd harmonics ascending,
    accelsearch.c:199-211) of each subharmonic plane read at the grid point
    nearest to h times the fundamental (r on the 0.5 grid with ties to even,
    z on the 2 grid and w on the 20 grid with ties away from zero)."""
    rng = np.random.default_rng(21)
    nbins = 1 << 15
    fft = (rng.standard_normal(nbins) + 1j * rng.standard_normal(nbins)).astype(np.complex64)
    for b in (4000, 9000, 12345):
        fft[b] += 30.0
    obs = O.create_accelobs(O.make_params(2 * nbins, 1e-3, nbins, 12, 40, 8, 2.0))
    ctx = O.Ctx(obs, O.FP32, 4)
    rstep = 5
    rbins = obs.corr_uselen // 2
    fullrlo = obs.rlo + rstep * rbins
    fullrhi = fullrlo + rbins - 0.5
    fund, _, _, fmeta = ctx.subharm(fft.view(np.float32), 1, 1, fullrlo, fullrhi)
    acc = fund[:, :, : obs.corr_uselen].astype(np.float32).copy()
    h0 = O.harm_info(obs, 1, 1)
    zf = -h0.zmax + 2.0 * np.arange(h0.numz)
    wf = -h0.wmax + 20.0 * np.arange(h0.numw)
    rf = fullrlo + 0.5 * np.arange(obs.corr_uselen)
    for stage in range(1, obs.numharmstages):
        for hn in range(1, 1 << stage, 2):
            nh = 1 << stage
            hf = hn / nh
            sub, rinds, zinds, meta = ctx.subharm(fft.view(np.float32), nh, hn, fullrlo, fullrhi)
            rlo_h, zlo_h, wlo_h = meta[0], meta[1], meta[2]
            exp_r = ((_nearest_even_half(rf * hf) - rlo_h) * 2).astype(int)
            exp_z = ((_nearest_away(zf * hf, 2.0) - zlo_h) / 2).astype(int)
            exp_w = ((_nearest_away(wf * hf, 20.0) - wlo_h) / 20).astype(int)
            assert np.array_equal(rinds[: obs.corr_uselen], exp_r), (nh, hn)
            assert np.array_equal(zinds[: h0.numz], exp_z), (nh, hn)
            acc = acc + sub[exp_w][:, exp_z][:, :, exp_r]
        # after each stage the oracle's running sum is the next stage's input;
        # the final one is checked below
    _, _, sums = ctx.search(fft.view(np.float32), rstep, rstep + 1, sum_rstep=rstep)
    assert sums.dtype == np.float32
    assert np.array_equal(sums, acc)
    ctx.close()


def test_insertion_known_answer():
    """A hit sequence worked by hand through insert_new_accelcand
    (accel_utils.c:382-480): replace a weaker neighbour, keep a stronger one,
    append past the tail, replace-and-remove, insert at the head."""
    seq = [(100.0, 5.0), (110.0, 4.0), (112.0, 6.0), (130.0, 3.0), (120.0, 7.0), (50.0, 2.0)]
    rr = [s[0] for s in seq]
    sg = [s[1] for s in seq]
    n = len(seq)
    out = O.insert_cands(rr, sg, [1.0] * n, [1] * n, [0.0] * n, [0.0] * n)
    assert [(c[3], c[1]) for c in out] == [(50.0, 2.0), (120.0, 7.0)]


def test_insertion_invariants():
    """Random hit streams: every candidate is one of the hits.  Hits more than
    15 bins apart all survive (in r order); hits within 15 bins of one another
    with rising sigma collapse onto the last; with falling sigma onto the first.
    (Two candidates can end up closer than 15 bins: a hit may overwrite the
    previous node while a stronger next node stays.)"""
    rng = np.random.default_rng(5)
    for trial in range(20):
        n = 2000
        rr = np.sort(rng.uniform(0, 3000, size=n)) if trial % 2 else rng.uniform(0, 3000, size=n)
        rr = np.round(rr * 8) / 8
        sg = rng.uniform(2, 9, size=n).astype(np.float32)
        out = O.insert_cands(rr, sg, np.ones(n), np.ones(n, dtype=np.int32), np.zeros(n), np.zeros(n))
        hits = set(zip(rr.tolist(), sg.tolist()))
        for c in out:
            assert (c[3], c[1]) in hits
    far = rng.permutation(np.arange(40) * 16.5 + 3.0)
    out = O.insert_cands(far, np.full(40, 5.0), np.ones(40), np.ones(40, dtype=np.int32), np.zeros(40),
                         np.zeros(40))
    assert [c[3] for c in out] == sorted(far.tolist())
    near = 500.0 + np.arange(10) * 1.25
    up = O.insert_cands(near, np.arange(10) + 3.0, np.ones(10), np.ones(10, dtype=np.int32), np.zeros(10),
                        np.zeros(10))
    assert [(c[3], c[1]) for c in up] == [(near[-1], 12.0)]
    down = O.insert_cands(near, 12.0 - np.arange(10), np.ones(10), np.ones(10, dtype=np.int32), np.zeros(10),
                          np.zeros(10))
    assert [(c[3], c[1]) for c in down] == [(near[0], 12.0)]
