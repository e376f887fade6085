"""HIP path vs the oracle at the benchmark shapes themselves.

* config 2 at its full length (N = 2^24, zmax 200, wmax 300, 16 harmonics,
  sigma 5, 8 rsteps per launch): bands of 8 rsteps around each of the three
  injected signals;
* the served stream's shapes (50, 0, 8) in full and (200, 150 -> 160, 16) on a
  band, N = 2^21, and config 5's DM trial 282 (N = 2^18, 8 harmonics) in full;
* config 3 (N = 2^23, zmax 200, wmax 500) in its batched form: 32 rsteps per
  launch, the autotuner's kernel variants, full batches replayed from a
  captured hipGraph, on a band holding an injected signal.

Hit sets must be identical except for hits within TOL_POW of the threshold,
hit powers within TOL_POW (relative), and the candidate lists identical once
candidates within TOL_SIGMA of the threshold are removed from both (their
count is printed).  The search's own reference is PRESTO C's per-rstep loop
(accelsearch.c:183-223), restated by the oracle.
"""
import numpy as np
import pytest

import pyoracle as O
from paritylib import TOL_POW, compare_cands, compare_hits, injection_rstep

pytestmark = pytest.mark.gpu

NTHREADS = 16


def _plan(kw, fft):
    from prestozl_amd import binding as B
    plan = B.Plan(B.make_config(**kw))
    plan.set_spectrum(fft)
    plan.set_keep_hits(1 << 24)
    return plan


def _oracle(kw):
    okw = {k: v for k, v in kw.items() if k not in ("batchsize", "maxhits")}
    obs = O.create_accelobs(O.make_params(**okw))
    return obs, O.Ctx(obs, O.FP32, nthreads=NTHREADS)


def _check(tag, plan, ctx, obs, fft, lo, hi, sigma):
    cg, _ = plan.search(lo, hi, stats=False)
    hg, pg = plan.hits()
    co, ho, _ = ctx.search(fft, lo, hi, maxhits=1 << 24, hits_array=True)
    near, worst, common = compare_hits(plan, hg, pg, ho, list(obs.powcut))
    removed = compare_cands(cg, co, sigma)
    print("%s rsteps [%d, %d): %d gpu / %d oracle hits, %d near-threshold hit differences, hit power err %.3g; "
          "%d / %d candidates, %d within 1e-4 sigma of the threshold removed"
          % (tag, lo, hi, len(hg), len(ho), near, worst, len(cg), len(co), removed))
    assert common > 0 and len(co) > 0
    assert worst < TOL_POW
    return cg


@pytest.fixture(scope="module")
def config2_full():
    from prestozl_amd import synth
    fft, N, dt, opts, sigs = synth.config2()
    kw = dict(N=N, dt=dt, numbins=len(fft), zmax=200, wmax=300, numharm=16, sigma=5.0, wmaxP=True, batchsize=8)
    return fft, kw, sigs


@pytest.mark.parametrize("inj", [0, 1, 2])
def test_config2_full_length_bands(config2_full, inj):
    """N = 2^24, sigma 5, batchsize 8: 8 rsteps centred on injection 'inj',
    with every FFT length in one grouped launch (the lean group: the form the
    autotuner keeps at this shape) for two of the bands and one launch per
    length for the third."""
    from prestozl_amd import binding as B
    fft, kw, sigs = config2_full
    plan = _plan(kw, fft)
    if inj != 1:
        lean = [g for g, name in enumerate(B.corr_groups()) if "lean" in name][0]
        plan.set_corr_group(lean, True)
    obs, ctx = _oracle(kw)
    k = injection_rstep(obs, kw["N"], kw["dt"], sigs[inj])
    _check("config2 injection %d" % inj, plan, ctx, obs, fft, k - 4, k + 4, kw["sigma"])
    ctx.close()
    plan.close()


def _served(which):
    from prestozl_amd import synth
    if which == "c5":
        fft, N, dt = synth.dm_trial(282)
        return fft, dict(N=N, dt=dt, numbins=len(fft), zmax=200, wmax=300, numharm=8, sigma=5.0, wmaxP=True,
                         batchsize=8), (37.2, 40.0, -80.0)
    N, dt = 1 << 21, 1e-4
    sig = (0.05, 211.3, 30.0, -40.0)
    fft = synth.realfft_packed(synth.timeseries(N, dt, [sig], seed=41))
    zmax, wmax, nh = (50, 0, 8) if which == "c4a" else (200, 150, 16)
    return fft, dict(N=N, dt=dt, numbins=len(fft), zmax=zmax, wmax=wmax, numharm=nh, sigma=5.0, wmaxP=wmax > 0,
                     batchsize=8), sig[1:]


@pytest.mark.parametrize("which", ["c4a", "c4b", "c5"])
def test_served_shapes_hits_and_candidates(which):
    """(50, 0, 8) over the whole search; (200, 160, 16) on 16 rsteps around
    its injection; config 5's trial on its first 24 rsteps (the injected
    37.2 Hz signal sits in rstep 8)."""
    fft, kw, (f0, z, w) = _served(which)
    plan = _plan(kw, fft)
    obs, ctx = _oracle(kw)
    if which == "c4b":
        k = injection_rstep(obs, kw["N"], kw["dt"], (0.0, f0, z, w))
        lo, hi = k - 8, k + 8
    elif which == "c5":
        lo, hi = 0, 24
    else:
        lo, hi = 0, plan.geom.nrsteps
    cg = _check(which, plan, ctx, obs, fft, lo, hi, kw["sigma"])
    T = kw["N"] * kw["dt"]
    best = max(cg, key=lambda c: c[1])
    print("%s strongest candidate: r %.2f (injected mean r %.2f), z %.1f, w %.1f, sigma %.1f"
          % (which, best[3], f0 * T + z / 2 + w / 6, best[4], best[5], best[1]))
    ctx.close()
    plan.close()


def test_config3_batched_graph_band():
    """N = 2^23, zmax 200, wmax 500 (4096-point fundamental), sigma 5: 32
    rsteps per launch with the autotuner's variants and graph replay, on a
    band of 34 rsteps (one full replayed batch and a partial eager one),
    searched twice (capture, then replay of the captured graph)."""
    from prestozl_amd import synth
    fft, N, dt, opts, sigs = synth.config3()
    kw = dict(N=N, dt=dt, numbins=len(fft), zmax=200, wmax=500, numharm=16, sigma=5.0, wmaxP=True, batchsize=32)
    plan = _plan(kw, fft)
    plan.autotune(reps=1)
    plan.set_graph(True)
    obs, ctx = _oracle(kw)
    k = injection_rstep(obs, N, dt, sigs[0])
    lo, hi = k - 16, k + 18
    c1 = _check("config3 batched, graph capture", plan, ctx, obs, fft, lo, hi, kw["sigma"])
    c2, st = plan.search(lo, hi, stats=True)
    print("graph launches in the replayed search: %d" % st["graph_launches"])
    assert st["graph_launches"] >= 1
    assert c2 == c1
    ctx.close()
    plan.close()
