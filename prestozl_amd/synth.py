"""Synthetic jerk-search inputs in PRESTO's file formats.

* .fft: raw little-endian complex64, realfft packing -- bin 0 holds
  (DC, Nyquist), bins 1..N/2-1 the positive frequencies (fastffts.c:163-235,
  read_fft.c:4-47; numbins = filesize / 8, accel_utils.c:1763).
* .inf: the key = value text of writeinf (ioinf.c:254-350) for an artificial
  data set ("None (Artificial Data Set)", ioinf.c:8-11), which readinf
  (ioinf.c:121-250) parses without telescope fields.

Signals follow phi(t) = 2 pi (f0 t + fd t^2/2 + fdd t^3/6) with z = fd T^2 and
w = fdd T^3 (T = N dt), the Fourier-domain derivatives the search grids.
"""
import os

import numpy as np

ARTIFICIAL = "None (Artificial Data Set)"


def timeseries(N, dt, signals, seed, noise_sigma=1.0):
    """float32 time series: Gaussian noise plus sum of A cos(phi(t)).

    signals: iterable of (amp, f0_hz, z, w)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.standard_normal(N) * noise_sigma
    T = N * dt
    t = np.arange(N, dtype=np.float64) * dt
    for amp, f0, z, w in signals:
        fd = z / T ** 2
        fdd = w / T ** 3
        phase = 2.0 * np.pi * (f0 * t + 0.5 * fd * t * t + fdd * t * t * t / 6.0)
        x += amp * np.cos(phase)
    return x.astype(np.float32)


def realfft_packed(x, unit_mean_power=False):
    """Forward real FFT of x (float32) as PRESTO packs it: N/2 complex64 bins."""
    X = np.fft.rfft(x.astype(np.float64))
    N = len(x)
    out = np.empty(N // 2, dtype=np.complex128)
    out[0] = complex(X[0].real, X[N // 2].real)
    out[1:] = X[1:N // 2]
    if unit_mean_power:
        out /= np.sqrt(np.mean(np.abs(out[1:]) ** 2))
    return out.astype(np.complex64)


def write_inf(root, N, dt, notes="synthetic data"):
    lines = [
        " Data file name without suffix          =  %s" % os.path.basename(root),
        " Telescope used                         =  %s" % ARTIFICIAL,
        " Number of bins in the time series      =  %-11.0f" % N,
        " Width of each time series bin (sec)    =  %.15g" % dt,
        " Any breaks in the data? (1 yes, 0 no)  =  0",
        " Data analyzed by                       =  prestozl_amd.synth",
        " Any additional notes:",
        "    %s" % notes,
        "",
    ]
    with open(root + ".inf", "w") as f:
        f.write("\n".join(lines) + "\n")


def write_fft(root, fft):
    np.ascontiguousarray(fft, dtype=np.complex64).tofile(root + ".fft")


def read_inf(root):
    """(N, dt, telescope) from a .inf (fields in readinf order)."""
    vals = []
    with open(root + ".inf") as f:
        for line in f:
            if "=" in line:
                vals.append(line.split("=", 1)[1].strip())
    telescope = vals[1]
    if telescope != ARTIFICIAL:
        raise ValueError("only artificial-data .inf files are parsed here")
    return float(vals[2]), float(vals[3]), telescope


def read_fft(root):
    return np.fromfile(root + ".fft", dtype=np.complex64)


# ---- the benchmark inputs -------------------------------------------------

def config1(seed=1):
    """N=2^20, dt=64 us, noise + one sinusoid (A=0.05, 123.456 Hz, fdot=-2e-6 Hz/s)."""
    N, dt = 1 << 20, 64e-6
    T = N * dt
    z = -2e-6 * T * T
    x = timeseries(N, dt, [(0.05, 123.456, z, 0.0)], seed)
    return realfft_packed(x, unit_mean_power=True), N, dt, dict(zmax=20, wmax=0, numharm=4, sigma=2.0)


def random_signals(rng, n, zspan=150.0, wspan=250.0, amp=(0.02, 0.08), f=(50.0, 500.0)):
    return [(float(rng.uniform(*amp)), float(rng.uniform(*f)), float(rng.uniform(-zspan, zspan)),
             float(rng.uniform(-wspan, wspan))) for _ in range(n)]


def config2(seed=2, N=1 << 24):
    """N=2^24, dt=1e-4 s, noise + three jerked sinusoids; zmax 200, wmax 300, 16 harmonics."""
    dt = 1e-4
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    sigs = random_signals(rng, 3)
    x = timeseries(N, dt, sigs, seed)
    return realfft_packed(x), N, dt, dict(zmax=200, wmax=300, numharm=16, sigma=5.0, batchsize=8), sigs


def config3(seed=3, N=1 << 23):
    dt = 1e-4
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    sigs = random_signals(rng, 5)
    x = timeseries(N, dt, sigs, seed)
    return realfft_packed(x), N, dt, dict(zmax=200, wmax=500, numharm=16, sigma=5.0), sigs


def dm_trial(i, seed=5, N=262144, dt=6.4e-4):
    """DM-trial file i of 564: common sinusoid (37.2 Hz, z=40, w=-80) with a
    Gaussian amplitude profile in trial index peaking at 0.1 at trial 282."""
    amp = 0.1 * np.exp(-0.5 * ((i - 282) / 10.0) ** 2)
    x = timeseries(N, dt, [(amp, 37.2, 40.0, -80.0)], seed * 100000 + i)
    return realfft_packed(x), N, dt
