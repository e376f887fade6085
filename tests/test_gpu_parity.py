"""HIP path vs the oracle (PRESTO C restated) on the same synthetic inputs.

Tolerances (FP32 search, as the reference):
  * kernel bank: |K_gpu - K_ref| <= 2e-6 * max|K_ref| per kernel row;
  * powers and summed powers: |P_gpu - P_ref| / max(P_ref, 1) <= 1e-5, the
    floor 1 being the normalised noise power (median / ln 2 normalisation);
  * hits (rstep, stage, w, z, r): identical sets except cells whose power lies
    within 1e-5 (relative) of the stage threshold; candidates (r, z, w,
    numharm) identical except those within 1e-4 in sigma of the threshold.
"""
import numpy as np
import pytest

import pyoracle as O
from conftest import rel_err
from paritylib import compare_cands

pytestmark = pytest.mark.gpu

TOL_POW = 1e-5
TOL_KERN = 2e-6


def _plan(cfg_kwargs, fft):
    from prestozl_amd import binding as B
    cfg = B.make_config(**cfg_kwargs)
    plan = B.Plan(cfg)
    plan.set_spectrum(fft)
    return plan


def _oracle(cfg_kwargs, mode, nthreads=16):
    kw = dict(cfg_kwargs)
    kw.pop("batchsize", None)
    kw.pop("maxhits", None)
    obs = O.create_accelobs(O.make_params(**kw))
    return obs, O.Ctx(obs, mode=mode, nthreads=nthreads)


def _c1(config1_data, **over):
    fft, N, dt, opts = config1_data
    kw = dict(N=N, dt=dt, numbins=len(fft), zmax=opts["zmax"], wmax=opts["wmax"], numharm=opts["numharm"],
              sigma=opts["sigma"], wmaxP=True, batchsize=8)
    kw.update(over)
    return fft, kw


def _c2_small():
    """zmax 200 / wmax 300 / 16 harmonics at a reduced length (same per-rstep shapes)."""
    from prestozl_amd import synth
    fft, N, dt, opts, sigs = synth.config2(N=1 << 21)
    kw = dict(N=N, dt=dt, numbins=len(fft), zmax=200, wmax=300, numharm=16, sigma=5.0, wmaxP=True, batchsize=4)
    return fft, kw


def _c3_small():
    """zmax 200 / wmax 500 / 16 harmonics (4096-point fundamental) at a reduced length."""
    from prestozl_amd import synth
    fft, N, dt, opts, sigs = synth.config3(N=1 << 21)
    kw = dict(N=N, dt=dt, numbins=len(fft), zmax=200, wmax=500, numharm=16, sigma=5.0, wmaxP=True, batchsize=2)
    return fft, kw, sigs


def _c4_shapes(which):
    """Two of the served stream's shapes: (zmax, wmax, numharm) = (50, 0, 8)
    and (200, 150, 16) (wmax 150 rounds up to 160), N = 2^21, dt = 1e-4 s, one
    injected jerked sinusoid."""
    from prestozl_amd import synth
    N, dt = 1 << 21, 1e-4
    x = synth.timeseries(N, dt, [(0.05, 211.3, 30.0, -40.0)], seed=41)
    fft = synth.realfft_packed(x)
    zmax, wmax, nh = (50, 0, 8) if which == "c4a" else (200, 150, 16)
    kw = dict(N=N, dt=dt, numbins=len(fft), zmax=zmax, wmax=wmax, numharm=nh, sigma=5.0, wmaxP=wmax > 0,
              batchsize=4)
    return fft, kw


def _c5_trial():
    """Config 5's DM trial 282 (N = 2^18, dt = 6.4e-4 s; zmax 200, wmax 300, 8 harmonics)."""
    from prestozl_amd import synth
    fft, N, dt = synth.dm_trial(282)
    kw = dict(N=N, dt=dt, numbins=len(fft), zmax=200, wmax=300, numharm=8, sigma=5.0, wmaxP=True, batchsize=8)
    return fft, kw


def _shape(config1_data, which):
    if which == "c1":
        return _c1(config1_data)
    if which == "c2":
        return _c2_small()
    if which in ("c4a", "c4b"):
        return _c4_shapes(which)
    if which == "c5":
        return _c5_trial()
    fft, kw, _ = _c3_small()
    return fft, kw


def test_geometry_matches_oracle(config1_data):
    fft, kw = _c1(config1_data)
    plan = _plan(kw, fft)
    obs, _ = _oracle(kw, O.FP32, 1)
    g = plan.geom
    assert (g.fftlen, g.corr_uselen, g.numz, g.numw) == (obs.fftlen, obs.corr_uselen, obs.numz, obs.numw)
    assert g.nrsteps == O.num_rsteps(obs)
    assert list(g.powcut)[: g.numharmstages] == list(obs.powcut)[: obs.numharmstages]
    plan.close()


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_kernel_bank(config1_data, which):
    fft, kw = _c1(config1_data) if which == "c1" else _c2_small()
    plan = _plan(kw, fft)
    obs = O.create_accelobs(O.make_params(**{k: v for k, v in kw.items() if k not in ("batchsize", "maxhits")}))
    worst32 = worst64 = 0.0
    for idx, (nh, hn) in enumerate(O.harmonics(obs)):
        h = plan.harms[idx]
        assert (h.numharm, h.harmnum) == (nh, hn)
        rows = sorted({0, h.numw * h.numz - 1, (h.numw // 2) * h.numz + h.numz // 2, h.numz // 3})
        for row in rows:
            wi, zi = divmod(row, h.numz)
            kg = plan.kernel(idx, row)
            k64 = O.kernel(obs, nh, hn, zi, wi, O.FP64)
            k32 = O.kernel(obs, nh, hn, zi, wi, O.FP32)
            scale = np.abs(k64).max()
            worst64 = max(worst64, np.abs(kg - k64).max() / scale)
            worst32 = max(worst32, np.abs(kg - k32).max() / scale)
    print("kernel bank max err vs fp64 %.3g, vs fp32 %.3g" % (worst64, worst32))
    assert worst64 < TOL_KERN
    assert worst32 < 2 * TOL_KERN
    plan.close()


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_planes(config1_data, which):
    fft, kw = _c1(config1_data) if which == "c1" else _c2_small()
    plan = _plan(kw, fft)
    obs, ctx32 = _oracle(kw, O.FP32)
    _, ctx64 = _oracle(kw, O.FP64)
    nr = plan.geom.nrsteps
    rstep_bins = plan.geom.corr_uselen // 2
    hl = O.harmonics(obs)
    worst32 = worst64 = 0.0
    for rstep in sorted({0, 1, nr // 2, nr - 1}):
        startr = obs.rlo + rstep * rstep_bins
        for idx in range(len(hl)) if which == "c1" else (0, 1, 3, 8, 15):
            nh, hn = hl[idx]
            pg, rlo = plan.plane(rstep, idx)
            p32, _, _, meta = ctx32.subharm(fft, nh, hn, startr, startr + rstep_bins - 0.5)
            p64, _, _, _ = ctx64.subharm(fft, nh, hn, startr, startr + rstep_bins - 0.5)
            assert rlo == meta[0]
            assert pg.shape == p32.shape
            e32, _ = rel_err(pg, p32, floor=1.0)
            e64, _ = rel_err(pg, p64, floor=1.0)
            worst32, worst64 = max(worst32, e32), max(worst64, e64)
    print("plane max rel err vs fp32 %.3g vs fp64 %.3g" % (worst32, worst64))
    assert worst32 < TOL_POW
    assert worst64 < TOL_POW
    plan.close()


@pytest.mark.parametrize("which", ["c1", "c2", "c3", "c4a", "c4b", "c5"])
def test_summed_powers(config1_data, which):
    """Final summed fundamental plane vs the FP32 oracle (tolerance TOL_POW)
    and the FP64 correlation-and-sum oracle (max and RMS error reported)."""
    fft, kw = _shape(config1_data, which)
    plan = _plan(kw, fft)
    obs, ctx32 = _oracle(kw, O.FP32)
    _, ctx64 = _oracle(kw, O.FP64)
    nr = plan.geom.nrsteps
    for rstep in (0, nr // 3):
        sg = plan.sums(rstep)
        _, _, s32 = ctx32.search(fft, rstep, rstep + 1, sum_rstep=rstep)
        _, _, s64 = ctx64.search(fft, rstep, rstep + 1, sum_rstep=rstep)
        e, rms = rel_err(sg, s32, floor=1.0)
        e64, rms64 = rel_err(sg, s64, floor=1.0)
        print("%s rstep %d summed powers vs fp32: max rel err %.3g rms %.3g; vs fp64: max %.3g rms %.3g"
              % (which, rstep, e, rms, e64, rms64))
        assert e < TOL_POW
        assert e64 < TOL_POW
    plan.close()


def _compare_hits(plan, hits_gpu, pows_gpu, ohits, powcut):
    g = {tuple(int(x) for x in h): float(p) for h, p in zip(hits_gpu, pows_gpu)}
    o = {(h[0], h[1], h[2], h[3], h[4]): h[5] for h in ohits}
    only_g = set(g) - set(o)
    only_o = set(o) - set(g)
    near = 0
    for k in only_g:
        assert abs(g[k] - powcut[k[1]]) <= 1e-5 * powcut[k[1]], ("extra hit", k, g[k], powcut[k[1]])
        near += 1
    for k in only_o:
        assert abs(o[k] - powcut[k[1]]) <= 1e-5 * powcut[k[1]], ("missing hit", k, o[k], powcut[k[1]])
        near += 1
    worst = 0.0
    for k in set(g) & set(o):
        worst = max(worst, abs(g[k] - o[k]) / max(o[k], 1.0))
    return near, worst, len(set(g) & set(o))


def test_search_config1(config1_data):
    """Full config-1 search (N=2^20, zmax 20, wmax 0, numharm 4, sigma 2)."""
    fft, kw = _c1(config1_data)
    plan = _plan(kw, fft)
    obs, ctx32 = _oracle(kw, O.FP32)
    cg, st = plan.search()
    hg, pg = plan.hits()
    co, ho, _ = ctx32.search(fft)
    near, worst, common = _compare_hits(plan, hg, pg, ho, list(obs.powcut))
    print("config1: %d gpu hits, %d oracle hits, %d near-threshold diffs, hit power err %.3g; %d/%d cands"
          % (len(hg), len(ho), near, worst, len(cg), len(co)))
    assert worst < TOL_POW
    assert common > 0
    _assert_same_cands(cg, co, kw["sigma"])
    plan.close()


@pytest.mark.parametrize("option", ["nph", "flo_fhi", "numharm1", "lobin"])
def test_search_options(config1_data, option):
    """Search options the drop-in takes (accelsearch_cmd.h; create_accelobs,
    accel_utils.c:1636-2037), each over config 1 against the oracle: photon
    normalisation (-photon: powers / nph instead of the block median), an
    -flo/-fhi band, a single harmonic, and a spectrum file that starts at bin
    4000 (the .fft's lobin; the injected signal sits near bin 8286)."""
    fft, N, dt, opts = config1_data
    kw = dict(N=N, dt=dt, numbins=len(fft), zmax=opts["zmax"], wmax=opts["wmax"], numharm=opts["numharm"],
              sigma=opts["sigma"], wmaxP=True, batchsize=8)
    data = fft
    if option == "nph":
        kw["nph"] = 1.3
    elif option == "flo_fhi":
        kw.update(flo=50.0, fhi=300.0)
    elif option == "numharm1":
        kw.update(numharm=1, sigma=3.0)
    else:
        data = fft[4000:]
        kw.update(numbins=len(data), lobin=4000)
    plan = _plan(kw, data)
    obs, ctx32 = _oracle(kw, O.FP32)
    cg, _ = plan.search()
    hg, pg = plan.hits()
    co, ho, _ = ctx32.search(data)
    near, worst, common = _compare_hits(plan, hg, pg, ho, list(obs.powcut))
    print("%s: %d gpu hits, %d oracle hits, %d near-threshold diffs, hit power err %.3g; %d/%d cands"
          % (option, len(hg), len(ho), near, worst, len(cg), len(co)))
    assert worst < TOL_POW
    assert common > 0
    _assert_same_cands(cg, co, kw["sigma"])
    plan.close()


def test_search_config2_band():
    """Config-2 geometry (zmax 200, wmax 300, 16 harmonics) on a band of rsteps."""
    fft, kw = _c2_small()
    kw["sigma"] = 3.0  # more hits in a short band
    plan = _plan(kw, fft)
    obs, ctx32 = _oracle(kw, O.FP32)
    lo, hi = 60, 63  # holds the 210.7 Hz injection (r ~ 44190)
    cg, st = plan.search(lo, hi)
    hg, pg = plan.hits()
    co, ho, _ = ctx32.search(fft, lo, hi)
    near, worst, common = _compare_hits(plan, hg, pg, ho, list(obs.powcut))
    print("config2 band: %d gpu hits, %d oracle hits, %d near, err %.3g, cands %d/%d"
          % (len(hg), len(ho), near, worst, len(cg), len(co)))
    assert worst < TOL_POW
    _assert_same_cands(cg, co, kw["sigma"])
    plan.close()


def _assert_same_cands(cg, co, sigma):
    """(r, z, w, numharm) identical, power within TOL_POW and sigma within
    1e-4, once the candidates within 1e-4 sigma of the threshold are removed
    from both lists (their count is printed)."""
    removed = compare_cands(cg, co, sigma)
    print("candidates %d / %d, %d within 1e-4 sigma of the threshold removed" % (len(cg), len(co), removed))
    assert len(co) > 0


def test_search_config3_band():
    """Config-3 geometry (zmax 200, wmax 500: 4096-point fundamental) on the
    rsteps holding an injected signal."""
    fft, kw, sigs = _c3_small()
    kw["sigma"] = 3.0
    plan = _plan(kw, fft)
    obs, ctx32 = _oracle(kw, O.FP32)
    T = kw["N"] * kw["dt"]
    amp, f0, z, w = sigs[0]
    r_mean = f0 * T + z / 2.0 + w / 6.0  # mean Fourier frequency over the observation
    k = int((r_mean - obs.rlo) // (obs.corr_uselen // 2))
    lo, hi = max(k - 1, 0), k + 2
    cg, st = plan.search(lo, hi)
    hg, pg = plan.hits()
    co, ho, _ = ctx32.search(fft, lo, hi)
    near, worst, common = _compare_hits(plan, hg, pg, ho, list(obs.powcut))
    print("config3 band [%d, %d): %d gpu hits, %d oracle hits, %d near, err %.3g, cands %d/%d"
          % (lo, hi, len(hg), len(ho), near, worst, len(cg), len(co)))
    assert common > 0
    assert worst < TOL_POW
    _assert_same_cands(cg, co, kw["sigma"])
    plan.close()


def test_batch_size_invariance(config1_data):
    """Candidate list is independent of rsteps per launch (determinism)."""
    fft, kw = _c1(config1_data)
    res = []
    for bs in (1, 3, 8):
        kw["batchsize"] = bs
        plan = _plan(kw, fft)
        c, _ = plan.search(0, 60)
        res.append(c)
        plan.close()
    assert res[0] == res[1] == res[2]


@pytest.mark.parametrize("which", ["c1", "c2", "h32"])
def test_harmsum_forms_identical(config1_data, which):
    """Every harmonic-sum form (workgroup order x 1, 2 or 4 consecutive r per
    thread, with its wide loads and per-lane selection) gives the same hits,
    summed powers and candidates, bit for bit, as the one-r-per-thread form;
    that form is compared with the oracle by the tests above."""
    if which == "h32":
        fft, kw = _variant_shapes(config1_data)[0]  # six stages (numharm 32)
        lo, hi, rs = 0, 40, 20
    elif which == "c1":
        fft, kw = _c1(config1_data)
        lo, hi, rs = 0, -1, 100
    else:
        fft, kw = _c2_small()
        kw["sigma"] = 3.0
        lo, hi, rs = 58, 66, 61
    plan = _plan(kw, fft)
    ref = None
    for order in (0, 1):
        for rpt in (1, 2, 4):
            plan.set_sum_form(order, rpt)
            assert plan.sum_form() == (order, rpt)
            c, _ = plan.search(lo, hi)
            h, p = plan.hits()
            idx = np.lexsort(h.T[::-1])
            got = (c, h[idx], p[idx], plan.sums(rs))
            if ref is None:
                ref = got
                assert len(h) > 0
                continue
            print("%s order %d rpt %d: %d hits, %d candidates" % (which, order, rpt, len(h), len(c)))
            assert got[0] == ref[0]
            assert np.array_equal(got[1], ref[1])
            assert np.array_equal(got[2], ref[2])
            assert np.array_equal(got[3], ref[3])
    plan.close()


def test_fractional_rlo_every_sum_form(config1_data):
    """-rlo with a fractional value (kept as given, accel_utils.c:1923): every
    rstep starts at a half-integer r, so the subharmonic r indices come from the
    per-rstep tables instead of integer arithmetic.  Hits match the oracle and
    are identical for every harmonic-sum form."""
    fft, kw = _c1(config1_data, numharm=8, rlo=7000.5, rhi=12000.25)
    kw.pop("flo", None)
    plan = _plan(kw, fft)
    obs, ctx32 = _oracle(kw, O.FP32)
    assert obs.rlo != int(obs.rlo)
    co, ho, _ = ctx32.search(fft)
    ref = None
    for order, rpt in ((1, 1), (0, 2), (1, 4)):
        plan.set_sum_form(order, rpt)
        cg, _ = plan.search()
        hg, pg = plan.hits()
        near, worst, common = _compare_hits(plan, hg, pg, ho, list(obs.powcut))
        print("rlo %.2f, form (%d, %d): %d gpu hits, %d oracle hits, %d near, err %.3g"
              % (obs.rlo, order, rpt, len(hg), len(ho), near, worst))
        assert common > 0 and worst < TOL_POW
        _assert_same_cands(cg, co, kw["sigma"])
        idx = np.lexsort(hg.T[::-1])
        got = (cg, hg[idx], pg[idx])
        if ref is None:
            ref = got
        else:
            assert got[0] == ref[0]
            assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])
    plan.close()


def test_hit_overflow_rerun_identical(config1_data):
    """A batch whose hits overflow the hit slot is searched again one rstep at
    a time (the slot always holds one rstep's cells x stages): candidates and
    hits equal those of a plan whose slots hold whole batches.  Strong bins in
    one band of the spectrum (every third bin, well above the noise median)
    make nearly every cell there a hit."""
    fft, kw = _c1(config1_data)
    fft = fft.copy()
    lo_bin, hi_bin = 12000, 19000
    fft[lo_bin:hi_bin:3] *= 40.0
    lo, hi = 8, 24  # two batches of 8 rsteps over the band
    p1 = _plan(kw, fft)
    c1, _ = p1.search(lo, hi)
    h1, w1 = p1.hits()
    kw2 = dict(kw, maxhits=1)  # slot = one rstep's cells x stages
    p2 = _plan(kw2, fft)
    c2, _ = p2.search(lo, hi)
    h2, w2 = p2.hits()
    g = p1.geom
    per_rstep = g.corr_uselen * p1.harms[0].numz * max(p1.harms[0].numw, 1) * g.numharmstages
    per_batch = np.bincount((h1[:, 0] - lo) // kw["batchsize"])
    print("hits %d, per batch %s, slot of the small plan %d" % (len(h1), per_batch.tolist(), per_rstep))
    assert per_batch.max() > per_rstep  # the small plan's slot did overflow
    assert c1 == c2
    i1, i2 = np.lexsort(h1.T[::-1]), np.lexsort(h2.T[::-1])
    assert np.array_equal(h1[i1], h2[i2]) and np.array_equal(w1[i1], w2[i2])
    p1.close()
    p2.close()


def _variant_shapes(config1_data):
    from prestozl_amd import synth
    fft1, N1, dt1, _ = config1_data
    # numharm 32 at small z: FFT lengths 128 ... 2048 (six stages)
    a = (fft1, dict(N=N1, dt=dt1, numbins=len(fft1), zmax=4, wmax=0, numharm=32, sigma=3.0, wmaxP=True,
                    batchsize=4))
    fft3, N3, dt3, opts, _ = synth.config3(N=1 << 21)
    # zmax 200 / wmax 500: the 4096-point fundamental
    b = (fft3, dict(N=N3, dt=dt3, numbins=len(fft3), zmax=200, wmax=500, numharm=16, sigma=5.0, wmaxP=True,
                    batchsize=2))
    return [a, b]


@pytest.mark.parametrize("shape", [0, 1])
def test_corr_variants(config1_data, shape):
    """Every correlation-kernel variant of every FFT length gives the oracle's
    planes (FP64 correlation, tolerance TOL_POW)."""
    from prestozl_amd import binding as B
    fft, kw = _variant_shapes(config1_data)[shape]
    plan = _plan(kw, fft)
    obs, ctx64 = _oracle(kw, O.FP64)
    hl = O.harmonics(obs)
    nr = plan.geom.nrsteps
    rstep_bins = plan.geom.corr_uselen // 2
    rstep = nr // 2
    startr = obs.rlo + rstep * rstep_bins
    by_len = {}
    for idx in range(len(hl)):
        by_len.setdefault(plan.harms[idx].fftlen, []).append(idx)
    tested = 0
    for n, idxs in sorted(by_len.items()):
        # the largest and the smallest harmonic of the class
        pick = sorted({idxs[0], idxs[-1]})
        refs = {}
        for idx in pick:
            nh, hn = hl[idx]
            refs[idx] = ctx64.subharm(fft, nh, hn, startr, startr + rstep_bins - 0.5)[0]
        for v in range(B.lib().pzl_corr_num_variants(n)):
            plan.set_variant(n, v)
            for idx in pick:
                pg, _ = plan.plane(rstep, idx)
                e, _ = rel_err(pg, refs[idx], floor=1.0)
                name = B.lib().pzl_corr_variant_name(n, v).decode()
                print("fftlen %d variant %s harmonic %s: max rel err %.3g" % (n, name, hl[idx], e))
                assert e < TOL_POW, name
                tested += 1
    assert tested >= 2 * len(by_len)
    plan.close()


def _groups_for(plan):
    """(group, all) for every correlation group that holds the plan's shorter
    FFT lengths (all False) or every length (all True)."""
    from prestozl_amd import binding as B
    ok = []
    for every in (False, True):
        for g in range(len(B.corr_groups())):
            try:
                plan.set_corr_group(g, every)
            except RuntimeError:
                continue
            ok.append((g, every))
    plan.set_corr_group(-1)
    return ok


def _group_name(form):
    from prestozl_amd import binding as B
    g, every = form
    return "none" if g < 0 else ("every length: " if every else "shorter lengths: ") + B.corr_groups()[g]


@pytest.mark.parametrize("which", ["c2", "c3"])
def test_corr_groups(which):
    """Every correlation group, holding the shorter FFT lengths or every
    length in one launch, gives the oracle's planes (FP64 correlation,
    tolerance TOL_POW) for the largest and smallest harmonic of each length it
    holds, at the config-2 shape (2048-point fundamental: 192 .. 2048) and the
    config-3 shape (4096-point fundamental: the group holds 2048 and shorter,
    so only the shorter lengths are grouped)."""
    from prestozl_amd import binding as B
    fft, kw = _c2_small() if which == "c2" else _c3_small()[:2]
    plan = _plan(kw, fft)
    obs, ctx64 = _oracle(kw, O.FP64)
    hl = O.harmonics(obs)
    rstep = plan.geom.nrsteps // 2
    rstep_bins = plan.geom.corr_uselen // 2
    startr = obs.rlo + rstep * rstep_bins
    by_len = {}
    for idx in range(len(hl)):
        by_len.setdefault(plan.harms[idx].fftlen, []).append(idx)
    longest = max(by_len)
    forms = _groups_for(plan)
    assert any(not every for _, every in forms), "no group holds the shorter lengths %s" % sorted(by_len)
    if which == "c2":
        assert any(every for _, every in forms), "no group holds every length %s" % sorted(by_len)
    refs = {}
    tested = 0
    for form in forms:
        plan.set_corr_group(*form)
        for n, idxs in sorted(by_len.items()):
            if n == longest and not form[1]:
                continue
            for idx in sorted({idxs[0], idxs[-1]}):
                if idx not in refs:
                    nh, hn = hl[idx]
                    refs[idx] = ctx64.subharm(fft, nh, hn, startr, startr + rstep_bins - 0.5)[0]
                pg, _ = plan.plane(rstep, idx)
                e, _ = rel_err(pg, refs[idx], floor=1.0)
                print("%s fftlen %d harmonic %s: max rel err %.3g" % (_group_name(form), n, hl[idx], e))
                assert e < TOL_POW
                tested += 1
    assert tested >= 2 * len(forms)
    plan.close()


@pytest.mark.parametrize("which", ["c2", "c3"])
def test_search_band_every_group(which):
    """A band of the config-2 (3 rsteps at sigma 3, with the 210.7 Hz
    injection) and config-3 (4 rsteps) searches with one launch per FFT length
    and in every grouped form (the shorter lengths, or every length, in one
    launch of each group): hits and candidates as the oracle's."""
    from prestozl_amd import binding as B
    if which == "c2":
        fft, kw = _c2_small()
        kw["sigma"] = 3.0
        lo, hi = 60, 63
    else:
        fft, kw, _ = _c3_small()
        lo, hi = 100, 104
    plan = _plan(kw, fft)
    obs, ctx32 = _oracle(kw, O.FP32)
    co, ho, _ = ctx32.search(fft, lo, hi)
    for form in [(-1, False)] + _groups_for(plan):
        plan.set_corr_group(*form)
        cg, st = plan.search(lo, hi)
        hg, pg = plan.hits()
        near, worst, common = _compare_hits(plan, hg, pg, ho, list(obs.powcut))
        print("%s %s: %d gpu hits, %d oracle hits, %d near, err %.3g, grouped launches %d, cands %d/%d"
              % (which, _group_name(form), len(hg), len(ho), near, worst, st["group_launches"], len(cg), len(co)))
        assert worst < TOL_POW and common > 0
        assert (st["group_launches"] > 0) == (form[0] >= 0)
        _assert_same_cands(cg, co, kw["sigma"])
    plan.close()


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_graph_replay_identical(config1_data, which):
    """Batches replayed from captured graphs give the eager path's hits and
    candidates exactly."""
    fft, kw = _c1(config1_data) if which == "c1" else _c2_small()
    plan = _plan(kw, fft)
    lo, hi = (0, -1) if which == "c1" else (40, 80)  # more batches than hit slots: replays happen
    plan.set_graph(False)
    c_eager, _ = plan.search(lo, hi, stats=False)
    h_eager = plan.hits()
    plan.set_graph(True)
    c_graph, st = plan.search(lo, hi, stats=True)
    h_graph = plan.hits()
    c_again, _ = plan.search(lo, hi, stats=False)  # replay of the already captured graphs
    plan.close()
    print("graph launches %d, eager candidates %d" % (st["graph_launches"], len(c_eager)))
    assert st["graph_launches"] > 0
    assert c_graph == c_eager and c_again == c_eager
    assert np.array_equal(h_graph[1], h_eager[1])
    assert all(np.array_equal(a, b) for a, b in zip(h_graph[0], h_eager[0]))


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_cu_split_identical(config1_data, which):
    """Harmonic sums confined to some compute units (with the correlations on
    every CU, or kept off the sums' own) give the whole chip's hits and
    candidates exactly, eager and replayed; invalid counts are refused."""
    fft, kw = _c1(config1_data) if which == "c1" else _c2_small()
    plan = _plan(kw, fft)
    lo, hi = (0, -1) if which == "c1" else (40, 80)
    assert plan.cu_split() == (0, 0)
    c0, _ = plan.search(lo, hi, stats=False)
    h0 = plan.hits()
    for split in ((64, 0), (64, 32), (96, 32)):
        plan.set_cu_split(*split)
        assert plan.cu_split() == split
        for graph in (False, True):
            plan.set_graph(graph)
            c, _ = plan.search(lo, hi, stats=False)
            h = plan.hits()
            print("split %s graph %d: %d candidates" % (split, graph, len(c)))
            assert c == c0
            assert np.array_equal(h[1], h0[1])
            assert all(np.array_equal(a, b) for a, b in zip(h[0], h0[0]))
    for bad in ((64, 96), (-32, 0), (1 << 20, 0)):  # invalid on any device
        with pytest.raises(RuntimeError):
            plan.set_cu_split(*bad)
    plan.set_cu_split(0, 0)
    assert plan.search(lo, hi, stats=False)[0] == c0
    plan.close()


def test_hit_chunks_identical(config1_data):
    """Hit records travel to the host in chunks of pinned pool buffers: tiny
    chunks (a batch's records split over many buffers, copies and insertion
    jobs) give the default chunk's hits and candidates exactly; invalid sizes
    are refused."""
    fft, kw = _c1(config1_data)
    plan = _plan(kw, fft)
    c0, _ = plan.search(stats=False)
    h0 = plan.hits()
    assert len(h0[1]) > 1000
    for chunk in (97, 4096, 1 << 20):
        plan.set_hit_chunk(chunk)
        c, _ = plan.search(stats=False)
        h = plan.hits()
        print("chunk %d: %d hits, %d candidates" % (chunk, len(h[1]), len(c)))
        assert c == c0
        assert np.array_equal(h[1], h0[1])
        assert all(np.array_equal(a, b) for a, b in zip(h[0], h0[0]))
    for bad in (0, -5, 1 << 40):
        with pytest.raises(RuntimeError):
            plan.set_hit_chunk(bad)
    plan.close()


def test_tune_batchsize(config1_data):
    """The rsteps-per-launch autotuner returns one of the candidates, and the
    search it would configure gives the same candidates as batch size 8."""
    from prestozl_amd import binding as B
    fft, kw = _c1(config1_data)
    best, ms = B.tune_batchsize(B.make_config(**kw), fft, candidates=(2, 8, 16), band=48)
    print("batch-size tuning:", best, ms)
    assert best in (2, 8, 16) and all(v > 0 for v in ms.values())
    p8 = _plan(kw, fft)
    c8, st = p8.search()
    kw2 = dict(kw)
    kw2["batchsize"] = best
    pb = _plan(kw2, fft)
    for c in st["classes"]:  # the same kernel variants: identical arithmetic
        pb.set_variant(c["fftlen"], p8.variant(c["fftlen"]))
    pb.set_corr_group(*p8.corr_group())
    assert pb.search(stats=False)[0] == c8
    p8.close()
    pb.close()
