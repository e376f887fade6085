"""Candidate optimisation on the GPU (pzl_optimize / pzl_optimize_spectrum)
against the oracle's restatement of optimize_accelcand with harmonic
polishing (max_rzw_arr_harmonics, maximize_rzw.c:85-191; get_localpower3d /
get_derivs3d, characteristics.c:75-191).

Tolerances.  The GPU builds its responses in double precision where the
reference (and the oracle) use float FFTs, so interpolated amplitudes differ
at ~1e-6 relative:
  * the volume maximum (r, z, w): identical grid point;
  * power and local power: 1e-4 relative; phase: 1e-4 rad;
  * the finite differences of get_derivs3d (h = 0.003) amplify that: dpow
    within 2e-3 of the power, d2pow within 2% of |d2pow| + 0.05 pow, dphs
    within 2e-3 rad/bin, d2phs within 0.5 rad/bin^2.
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

N = 1 << 16


def _chirp(harmonics, noise, seed=5):
    t = np.arange(N, dtype=np.float64) / N
    x = np.zeros(N)
    for a, f0, z0, w in harmonics:
        x += a * np.cos(2 * np.pi * (f0 * t + 0.5 * z0 * t * t + w * t ** 3 / 6.0))
    x += np.random.Generator(np.random.PCG64(seed)).standard_normal(N) * noise
    return np.fft.rfft(x)[: N // 2].astype(np.complex64)


def _compare(g, o, label):
    ro, zo, wo, ders, pows = o
    assert (g["r"], g["z"], g["w"]) == pytest.approx((ro, zo, wo), abs=1e-9), label
    for h, (dg, do) in enumerate(zip(g["derivs"], ders)):
        p = do.pow
        assert dg["pow"] == pytest.approx(p, rel=1e-4), (label, h)
        assert dg["locpow"] == pytest.approx(do.locpow, rel=1e-4), (label, h)
        assert abs(dg["phs"] - do.phs) < 1e-4, (label, h)
        assert abs(dg["dpow"] - do.dpow) <= 2e-3 * p, (label, h)
        assert abs(dg["d2pow"] - do.d2pow) <= 0.02 * abs(do.d2pow) + 0.05 * p, (label, h)
        assert abs(dg["dphs"] - do.dphs) <= 2e-3, (label, h)
        assert abs(dg["d2phs"] - do.d2phs) <= 0.5, (label, h)
        assert g["pows"][h] == dg["pow"]


@pytest.mark.parametrize("numharm", [1, 2, 4, 8])
def test_optimize_matches_oracle_synthetic(numharm):
    from prestozl_amd import binding as B
    f0, z0, w = 1510.37, 3.1, 8.0
    comps = [(1.0 / (1 + 0.3 * k), f0 * (k + 1), z0 * (k + 1), w * (k + 1)) for k in range(numharm)]
    data = _chirp(comps, noise=5.0)
    r, z, ww = f0 + 0.5 * z0 + w / 6.0, z0 + 0.5 * w, w
    starts = [(np.round(r * numharm * 2) / 2 / numharm + dr, np.round(z * numharm / 2) * 2 / numharm,
               np.round(ww * numharm / 20) * 20 / numharm, numharm) for dr in (0.0, 0.5 / numharm)]
    got = B.optimize_spectrum(data, starts)
    for s, g in zip(starts, got):
        o = O.max_rzw_arr_harmonics(data, numharm, s[0], s[1], s[2])
        print("numharm %d start %s -> gpu (%.4f %.3f %.2f) oracle (%.4f %.3f %.2f) pow %.4g / %.4g"
              % (numharm, s[:3], g["r"], g["z"], g["w"], o[0], o[1], o[2], g["pows"][0], o[4][0]))
        _compare(g, o, s)


def test_optimize_search_candidates_config2_band():
    """Candidates of a zmax 200 / wmax 300 / 16-harmonic search band around an
    injection, optimised with the plan's resident spectrum."""
    from prestozl_amd import binding as B
    from prestozl_amd import synth
    fft, N2, dt, opts, sigs = synth.config2(N=1 << 21)
    kw = dict(N=N2, dt=dt, numbins=len(fft), zmax=200, wmax=300, numharm=16, sigma=3.0, wmaxP=True, batchsize=4)
    plan = B.Plan(B.make_config(**kw))
    plan.set_spectrum(fft)
    cands, _ = plan.search(60, 63)
    assert cands
    top = sorted(cands, key=lambda c: -c[1])[:3]
    sel = [(c[3], c[4], c[5], c[2]) for c in top]  # (r, z, w, numharm); the spectrum starts at bin 0
    got = plan.optimize(sel)
    plan.close()
    for s, g in zip(sel, got):
        o = O.max_rzw_arr_harmonics(fft, s[3], s[0], s[1], s[2])
        print("cand %s -> gpu (%.4f %.3f %.2f) oracle (%.4f %.3f %.2f)" % (s, g["r"], g["z"], g["w"], *o[:3]))
        _compare(g, o, s)


def test_optimize_rejects_bad_candidates():
    from prestozl_amd import binding as B
    data = _chirp([], noise=1.0)
    with pytest.raises(B.PzlError):
        B.optimize_spectrum(data, [(100.0, 0.0, 10.0, 0)])
    with pytest.raises(B.PzlError):
        B.optimize_spectrum(data, [(100.0, 0.0, 10.0, 40)])


@pytest.mark.parametrize("numharm", [1, 3])
def test_optimize_fdot_only_matches_oracle(numharm):
    """w = 0 searches: the downhill simplex in (r, z) (maximize_rz.c:141-220,
    amoeba.c), every candidate's steps batched on the GPU.  Both sides stop at
    a relative simplex spread of 1e-10 on function values that differ at
    ~1e-6, so their optima agree to the flatness of the peak: r within 2e-3
    bins, z within 2e-2; powers 1e-4; local powers 1e-2; phases 1e-2 rad; first derivatives
    within 1e-2 of the power (and of a radian), second within 2% + 0.05 pow
    (and 0.5 rad)."""
    from prestozl_amd import binding as B
    f0, z0 = 2210.3, 6.7
    comps = [(1.0 / (1 + 0.5 * k), f0 * (k + 1), z0 * (k + 1), 0.0) for k in range(numharm)]
    data = _chirp(comps, noise=0.3)
    r, z = f0 + 0.5 * z0, z0
    starts = [(np.round(r * numharm * 2) / 2 / numharm + dr, np.round(z * numharm / 2) * 2 / numharm, 0.0, numharm)
              for dr in (0.0, 0.5 / numharm)]
    got = B.optimize_spectrum(data, starts, fdot_only=True)
    for s, g in zip(starts, got):
        ro, zo, ders, pows = O.max_rz_arr_harmonics(data, numharm, s[0], s[1])
        print("numharm %d start %s -> gpu (%.5f %.4f) oracle (%.5f %.4f)" % (numharm, s[:2], g["r"], g["z"], ro, zo))
        assert abs(g["r"] - ro) < 2e-3 and abs(g["z"] - zo) < 2e-2 and g["w"] == 0.0
        for h, (dg, do) in enumerate(zip(g["derivs"], ders)):
            p = do.pow
            assert dg["pow"] == pytest.approx(p, rel=1e-4), h
            assert g["pows"][h] == pytest.approx(pows[h], rel=1e-4), h
            # the 20 local-power points sit on this strong harmonic's own
            # sidelobes, whose power changes by ~40 per bin near their nulls
            assert dg["locpow"] == pytest.approx(do.locpow, rel=1e-2), h
            assert abs(dg["phs"] - do.phs) < 1e-2, h
            assert abs(dg["dpow"] - do.dpow) <= 1e-2 * p, h
            assert abs(dg["d2pow"] - do.d2pow) <= 0.02 * abs(do.d2pow) + 0.05 * p, h
            assert abs(dg["dphs"] - do.dphs) <= 1e-2, h
            assert abs(dg["d2phs"] - do.d2phs) <= 0.5, h


def _compare_single(g, o, label):
    """One harmonic of -noharmpolish against the oracle: the grid point (or
    the simplex optimum), the raw maximum power, the derivatives."""
    ro, zo, wo, d, maxpow = o
    assert (g["r"], g["z"], g["w"]) == pytest.approx((ro, zo, wo), abs=1e-9), label
    assert g["pows"][0] == pytest.approx(maxpow, rel=1e-4), label
    dg, p = g["derivs"][0], d.pow
    assert dg["pow"] == pytest.approx(p, rel=1e-4), label
    assert dg["locpow"] == pytest.approx(d.locpow, rel=1e-4), label
    assert abs(dg["phs"] - d.phs) < 1e-4, label
    assert abs(dg["dpow"] - d.dpow) <= 2e-3 * p, label
    assert abs(dg["d2pow"] - d.d2pow) <= 0.02 * abs(d.d2pow) + 0.05 * p, label
    assert abs(dg["dphs"] - d.dphs) <= 2e-3, label
    assert abs(dg["d2phs"] - d.d2phs) <= 0.5, label


@pytest.mark.parametrize("numharm", [1, 4])
def test_maximize_rzw_matches_oracle(numharm):
    """-noharmpolish, w > 0: every harmonic of a candidate through max_rzw_arr
    (maximize_rzw.c:3-50) at (r, z, w) * (ii + 1); plus starts whose injection
    lies beyond the z or w range (the maximum on the range's edge)."""
    from prestozl_amd import binding as B
    f0, z0, w = 1510.37, 3.1, 8.0
    comps = [(1.0 / (1 + 0.3 * k), f0 * (k + 1), z0 * (k + 1), w * (k + 1)) for k in range(numharm)]
    data = _chirp(comps, noise=5.0)
    r, z, ww = f0 + 0.5 * z0 + w / 6.0, z0 + 0.5 * w, w
    rs, zs, ws = np.round(r * 2) / 2, np.round(z / 2) * 2, np.round(ww / 20) * 20
    pts = [(rs * (ii + 1), zs * (ii + 1), ws * (ii + 1)) for ii in range(numharm)]
    pts += [(r + 0.3, z - 3.0, ww), (r - 0.2, z, ww + 27.0), (12.4, 1.0, -10.0)]
    got = B.maximize_spectrum(data, pts)
    for s, g in zip(pts, got):
        o = O.max_rzw_arr(data, *s)
        print("start %s -> gpu (%.4f %.3f %.2f) %.5g oracle (%.4f %.3f %.2f) %.5g"
              % (s, g["r"], g["z"], g["w"], g["pows"][0], o[0], o[1], o[2], o[4]))
        _compare_single(g, o, s)


def test_maximize_rz_matches_oracle():
    """-noharmpolish, w = 0: max_rz_arr's simplex on each harmonic's raw power
    (maximize_rz.c:25-85); the optima agree to the flatness of the peak as the
    harmonic simplexes do, the returned raw powers to 1e-4."""
    from prestozl_amd import binding as B
    f0, z0 = 2210.3, 6.7
    comps = [(1.0 / (1 + 0.5 * k), f0 * (k + 1), z0 * (k + 1), 0.0) for k in range(3)]
    data = _chirp(comps, noise=0.3)
    r, z = f0 + 0.5 * z0, z0
    rs, zs = np.round(r * 2) / 2, np.round(z / 2) * 2
    pts = [(rs * (ii + 1), zs * (ii + 1), 0.0) for ii in range(3)]
    got = B.maximize_spectrum(data, pts, fdot_only=True)
    for s, g in zip(pts, got):
        ro, zo, d, pw = O.max_rz_arr(data, s[0], s[1])
        print("start %s -> gpu (%.5f %.4f) %.6g oracle (%.5f %.4f) %.6g" % (s[:2], g["r"], g["z"], g["pows"][0], ro,
                                                                          zo, pw))
        assert abs(g["r"] - ro) < 2e-3 and abs(g["z"] - zo) < 2e-2 and g["w"] == 0.0
        assert g["pows"][0] == pytest.approx(pw, rel=1e-4)
        dg = g["derivs"][0]
        assert dg["pow"] == pytest.approx(d.pow, rel=1e-4)
        assert dg["locpow"] == pytest.approx(d.locpow, rel=1e-2)
        assert abs(dg["phs"] - d.phs) < 1e-2


def test_maximize_search_candidates_config2_band():
    """The harmonics of a zmax 200 / wmax 300 / 16-harmonic search's top
    candidates, each on its own (-noharmpolish), with the plan's resident
    spectrum (pzl_maximize)."""
    from prestozl_amd import binding as B
    from prestozl_amd import synth
    fft, N2, dt, opts, sigs = synth.config2(N=1 << 21)
    kw = dict(N=N2, dt=dt, numbins=len(fft), zmax=200, wmax=300, numharm=16, sigma=3.0, wmaxP=True, batchsize=4)
    plan = B.Plan(B.make_config(**kw))
    plan.set_spectrum(fft)
    cands, _ = plan.search(60, 63)
    assert cands
    top = sorted(cands, key=lambda c: -c[1])[:2]
    pts = [(c[3] * (ii + 1), c[4] * (ii + 1), c[5] * (ii + 1)) for c in top for ii in range(c[2])]
    arr = B._opt_array([(r, z, w, 1) for r, z, w in pts])
    B.check(B.lib().pzl_maximize(plan.h, arr, len(pts)))
    got = B._opt_result(arr, len(pts))
    plan.close()
    for s, g in zip(pts, got):
        _compare_single(g, O.max_rzw_arr(fft, *s), s)


def test_maximize_rejects_several_harmonics():
    from prestozl_amd import binding as B
    data = _chirp([], noise=1.0)
    arr = B._opt_array([(100.0, 0.0, 10.0, 2)])
    with pytest.raises(B.PzlError):
        B.check(B.lib().pzl_maximize_spectrum(0, B._fptr(data.view(np.float32)), len(data), arr, 1))
