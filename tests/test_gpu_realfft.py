"""realfft (fastffts.c:163-235) on the GPU against the oracle's restatement.

The GPU transform, like the reference's FFTW float transform, keeps the data
in float; the oracle's double-precision run (FP64) is the exact answer and its
float run (FP32) is the reference's own arithmetic.  Float FFT rounding grows
like sqrt(log2 n): the GPU's error against FP64, RMS over the bins relative to
the RMS bin, stays within 1e-6 at every size here (the oracle's FP32 run sits
at the same level), and no bin is off by more than 2e-5 of the largest bin.
The roundtrip is the reference's own known answer
(tests/test_presto_python.py:10-13: rfft(rfft(x, -1), 1) == x to 1e-6 at N =
20)."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

# n/2 = M: one workgroup up to 16448 points (odd and even M; radices 2, 3, 4,
# 5, 7, 11), four steps above: config 5's 2^18 samples, 3 x 2^17, 2^20,
# config 2's 2^24, and lengths choose_good_N picks (misc_utils.c:327-366):
# 2 x 1008 x 10 (M = 2^5 3^2 5 7) and 2 x 1056 x 1000 (M = 2^8 3 5^3 11)
SIZES = [2, 4, 14, 20, 22, 30, 4096, 6000, 20160, 32768, 1 << 18, 3 << 17, 2112000, 1 << 20, 1 << 24]


def _series(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.arange(n) / n
    return (rng.standard_normal(n) + 0.3 * np.cos(2 * np.pi * (0.1 * n * t + 40.0 * t * t))).astype(np.float32)


def _err(got, want):
    d = got.astype(np.float64) - want
    return np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(want * want)), np.abs(d).max() / np.abs(want).max()


@pytest.mark.parametrize("n", SIZES)
def test_forward_matches_oracle(n):
    from prestozl_amd import binding as B
    x = _series(n, n)
    got = B.realfft(x, -1)
    exact = O.realfft(x, -1, O.FP64)
    rms, mx = _err(got, exact)
    rms32, _ = _err(O.realfft(x, -1, O.FP32), exact) if n <= (1 << 20) else (float("nan"), 0)
    print("n %d: rms %.2e (oracle FP32 %.2e), max %.2e" % (n, rms, rms32, mx))
    assert rms < 1e-6 and mx < 2e-5, n


@pytest.mark.parametrize("n", [20, 22, 30, 4096, 20160, 1 << 18, 3 << 17, 2112000, 1 << 24])
def test_inverse_matches_oracle(n):
    """The inverse of a packed spectrum: bin n/4's sign flips (n % 4 == 0) and
    the 2/n scale."""
    from prestozl_amd import binding as B
    X = O.realfft(_series(n, n + 1), -1, O.FP64).astype(np.float32)
    got = B.realfft(X, 1)
    exact = O.realfft(X, 1, O.FP64)
    rms, mx = _err(got, exact)
    print("n %d: inverse rms %.2e, max %.2e" % (n, rms, mx))
    assert rms < 1e-6 and mx < 2e-5, n


@pytest.mark.parametrize("n", [20, 22, 6000, 2112000, 1 << 24])
def test_roundtrip_known_answer(n):
    from prestozl_amd import binding as B
    x = _series(n, 3 * n)
    y = B.realfft(B.realfft(x, -1), 1)
    tol = 1e-6 if n == 20 else 1e-6 * np.sqrt(np.log2(n) / np.log2(20)) * 4
    assert np.allclose(x, y, atol=tol), (n, np.abs(x - y).max())


def test_closed_forms_at_config2_length():
    """A cosine at bin f gives n/2 in bin f's real part, a sine -n/2 in its
    imaginary part, a constant c gives n c in bin 0's real part and (-1)^t n in
    its imaginary (Nyquist) part, at n = 2^24."""
    from prestozl_amd import binding as B
    n = 1 << 24
    t = np.arange(n, dtype=np.float64)
    f = 1234567
    x = (np.cos(2 * np.pi * f * t / n) + 0.5 * np.sin(2 * np.pi * (f + 10) * t / n) + 0.25
         + 0.125 * (-1.0) ** t).astype(np.float32)
    X = B.realfft(x, -1).view(np.complex64)
    tol = 2e-5 * n
    assert abs(X[f] - n / 2) < tol
    assert abs(X[f + 10] - (-0.5j * n / 2)) < tol
    assert abs(X[0].real - 0.25 * n) < tol and abs(X[0].imag - 0.125 * n) < tol
    rest = np.abs(X).copy()
    rest[[0, f, f + 10]] = 0
    assert rest.max() < tol


@pytest.mark.parametrize("n,nseries", [(20160, 3), (1 << 18, 5), (2112000, 2)])
def test_batched_series_equal_single_calls(n, nseries):
    """pzl_realfft_device_batch on series back to back in device memory gives
    each series exactly what one call on it gives."""
    import ctypes
    from prestozl_amd import binding as B
    B.lib()  # the library's HIP runtime is the one loaded below
    hip = ctypes.CDLL("libamdhip64.so")
    xs = np.stack([_series(n, 100 + i) for i in range(nseries)])
    nbytes = ctypes.c_size_t(xs.nbytes)
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), nbytes) == 0
    try:
        for isign, data in ((-1, xs), (1, np.stack([B.realfft(x, -1) for x in xs]))):
            assert hip.hipMemcpy(d, data.ctypes.data_as(ctypes.c_void_p), nbytes, 1) == 0
            B.realfft_device(d.value, n, isign, nseries=nseries)
            got = np.empty_like(data)
            assert hip.hipMemcpy(got.ctypes.data_as(ctypes.c_void_p), d, nbytes, 2) == 0
            for i in range(nseries):
                assert np.array_equal(got[i], B.realfft(data[i], isign)), (isign, i)
    finally:
        hip.hipFree(d)


def test_unsupported_lengths_are_refused():
    from prestozl_amd import binding as B
    for n in (21, 2 * 13 * 4096, 0):
        with pytest.raises(B.PzlError):
            B.realfft(np.zeros(n, np.float32), -1)


def test_realfft_command_dat_to_fft(tmp_path):
    """The realfft command: root.dat -> root.fft (forward) and back (inverse,
    from the suffix), equal to the library transform."""
    from prestozl_amd import binding as B
    exe = os.path.join(os.path.dirname(B.LIBPATH), "realfft")
    x = _series(1 << 18, 9)
    dat = tmp_path / "trial.dat"
    x.tofile(dat)
    r = subprocess.run([exe, str(dat)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    fft = np.fromfile(tmp_path / "trial.fft", dtype=np.float32)
    assert np.array_equal(fft, B.realfft(x, -1))
    os.remove(dat)
    r = subprocess.run([exe, "-delete", str(tmp_path / "trial.fft")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert not (tmp_path / "trial.fft").exists()
    back = np.fromfile(tmp_path / "trial.dat", dtype=np.float32)
    assert np.allclose(back, x, atol=1e-5)
