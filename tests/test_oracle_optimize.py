"""Pinning the oracle's restatement of the candidate optimisation
(accelsearch_CPU2 -> optimize_accelcand -> max_rzw_arr_harmonics,
get_localpower3d, get_derivs3d, corr_rzw_vol, rzw_interp; calc_props) against
what the mathematics fixes, on noise-free and lightly noisy synthetic chirps:

* interpolation at a chirp's true (r, z, w) recovers its coherent amplitude
  N A / 2 (the positive-frequency DFT amplitude of A cos);
* the correlation volume (FFT route, rzwinterp.c:3) equals direct
  interpolation (rzwinterp.c:145) at the same points;
* the harmonic-summed volume maximum lies within one grid step of an injected
  multi-harmonic signal;
* a constant-amplitude signal has purity 1 and centroid 1/2 (Middleditch,
  Deich and Kulkarni 1993; calc_props, characteristics.c:193).

Signal convention: phase 2 pi (f0 t + z0 t^2 / 2 + w t^3 / 6) over t in
[0, 1) gives PRESTO's (r, z, w) = (f0 + z0 / 2 + w / 6, z0 + w / 2, w): the
average Fourier frequency and f-dot over the observation.
"""
import numpy as np
import pytest

import pyoracle as O

N = 1 << 16


def chirp(harmonics, noise=0.0, seed=5):
    """harmonics: (amp, f0, z0, w) per component (t in units of T)."""
    t = np.arange(N, dtype=np.float64) / N
    x = np.zeros(N)
    for a, f0, z0, w in harmonics:
        x += a * np.cos(2 * np.pi * (f0 * t + 0.5 * z0 * t * t + w * t ** 3 / 6.0))
    if noise:
        x += np.random.Generator(np.random.PCG64(seed)).standard_normal(N) * noise
    X = np.fft.rfft(x)[: N // 2]
    return X.astype(np.complex64)


def presto_rzw(f0, z0, w):
    return f0 + 0.5 * z0 + w / 6.0, z0 + 0.5 * w, w


@pytest.mark.parametrize("z0,w", [(0.0, 0.0), (30.0, 0.0), (-20.0, 90.0), (12.0, -150.0)])
def test_rzw_interp_recovers_coherent_amplitude(z0, w):
    A, f0 = 1.0, 5000.25
    data = chirp([(A, f0, z0, w)])
    r, z, ww = presto_rzw(f0, z0, w)
    khw = O.w_resp_halfwidth_acc(z, ww, True)
    amp = abs(O.rzw_interp(data, r, z, ww, khw))
    assert amp == pytest.approx(N * A / 2, rel=5e-3)
    # and clearly less a quarter bin away
    assert abs(O.rzw_interp(data, r + 0.25, z, ww, khw)) < 0.95 * amp


def test_corr_volume_equals_direct_interpolation():
    """Same kernel half-width on both routes (the volume's, from the extreme z
    and w of its range): they differ only by FFT rounding and by the window
    edge (the correlation kernel spans [r - khw, r + khw), the direct sum the
    integer bins [int(r) - khw, int(r) + khw)).  An output index off by one
    step (1/8 bin) is off by ~0.4 of the peak."""
    f0, z0, w = 3000.0, 18.0, 60.0
    data = chirp([(1.0, f0, z0, w)])
    r, z, ww = presto_rzw(f0, z0, w)
    numbetween, numz, numw = 8, 5, 5
    zlo, zhi, wlo, whi = z - 4, z + 4, ww - 40, ww + 40
    khw = O.w_resp_halfwidth_acc(zhi if abs(zhi) > abs(zlo) else zlo, whi if abs(whi) > abs(wlo) else wlo, False)
    fftlen = 1 << int(np.ceil(np.log2(numbetween * (2 * khw + 30))))
    startbin = int(r) - 1
    vol, nextbin = O.corr_rzw_vol(data, numbetween, startbin, zlo, zhi, numz, wlo, whi, numw, fftlen)
    assert vol.shape[2] == fftlen - 2 * khw * numbetween
    assert nextbin == startbin + vol.shape[2] // numbetween
    peak = np.abs(vol).max()
    assert peak == pytest.approx(N / 2, rel=2e-3)  # the grid holds the injection
    err = {0: 0.0, 1: 0.0}
    for i in range(numw):
        for j in range(numz):
            zz, wv = zlo + j * (zhi - zlo) / (numz - 1), wlo + i * (whi - wlo) / (numw - 1)
            for k in range(3 * numbetween):
                for shift in err:
                    direct = O.rzw_interp(data, startbin + (k + shift) / numbetween, zz, wv, khw)
                    err[shift] = max(err[shift], abs(vol[i, j, k] - direct) / peak)
    assert err[0] < 2e-3, err
    assert err[1] > 0.1, err


@pytest.mark.parametrize("numharm", [1, 2, 4])
def test_harmonic_volume_maximum_finds_injection(numharm):
    f0, z0, w = 1510.37, 3.1, 8.0  # fundamental
    comps = [(1.0 / (1 + 0.3 * k), f0 * (k + 1), z0 * (k + 1), w * (k + 1)) for k in range(numharm)]
    data = chirp(comps, noise=5.0)
    r, z, ww = presto_rzw(f0, z0, w)
    # a start as the search grid would give it: r on a half-bin grid, z, w on their grids (fundamental units)
    rs = np.round(r * numharm * 2) / 2 / numharm
    zs = np.round(z * numharm / 2) * 2 / numharm
    ws = np.round(ww * numharm / 20) * 20 / numharm
    ro, zo, wo, ders, pows = O.max_rzw_arr_harmonics(data, numharm, rs, zs, ws)
    assert abs(ro - r) <= 1.0 / (8 * numharm) + 1e-9
    assert abs(zo - z) <= (2.0 / 4) / numharm + 0.15
    assert abs(wo - ww) <= (20.0 / 4) / numharm + 2.0
    assert len(ders) == numharm and all(p > 0 for p in pows)
    assert ders[0].pow == pows[0] and ders[0].locpow > 0


@pytest.mark.parametrize("z0,w", [(0.0, 0.0), (25.0, -60.0)])
def test_constant_amplitude_signal_has_unit_purity_and_half_centroid(z0, w):
    f0 = 7000.0
    data = chirp([(0.5, f0, z0, w)], noise=0.5)  # purity error 1/sqrt(10 P) ~ 3e-3
    r, z, ww = presto_rzw(f0, z0, w)
    lp = O.localpower3d(data, r, z, ww)
    d = O.derivs3d(data, r, z, ww, lp)
    p = O.calc_props(d, r, z, ww)
    assert p.pur == pytest.approx(1.0, abs=0.01)
    assert p.cen == pytest.approx(0.5, abs=0.005)
    assert p.pow == pytest.approx(d.pow / lp, rel=1e-6)
    assert p.sig == pytest.approx(np.sqrt(2 * p.pow - np.log(np.pi * p.pow)), rel=1e-5)


def test_local_power_of_noise():
    """Noise of variance s^2 per sample has mean power N s^2 per bin; the
    estimate averages 20 bins (exponential powers: ~22% scatter)."""
    data = chirp([], noise=3.0, seed=11)
    vals = [O.localpower3d(data, r, 10.0, 40.0) / (N * 9.0) for r in (2000.3, 4000.7, 9000.1, 12000.5)]
    assert np.mean(vals) == pytest.approx(1.0, rel=0.25)


def test_calc_props_refuses_zero_power():
    d = O.RDerivs(0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    with pytest.raises(ValueError):
        O.calc_props(d, 1.0, 0.0, 0.0)


@pytest.mark.parametrize("numharm", [1, 3])
def test_rz_simplex_finds_injection(numharm):
    """w = 0 searches optimise with a downhill simplex (maximize_rz.c:141,
    amoeba.c): continuous, so it lands on the injected (r, z) far inside the
    search grid's steps; the power there is the coherent power (N A / 2)^2."""
    f0, z0 = 2210.3, 6.7
    comps = [(1.0 / (1 + 0.5 * k), f0 * (k + 1), z0 * (k + 1), 0.0) for k in range(numharm)]
    data = chirp(comps, noise=0.3)
    r, z = f0 + 0.5 * z0, z0
    rs = np.round(r * numharm * 2) / 2 / numharm
    zs = np.round(z * numharm / 2) * 2 / numharm
    ro, zo, ders, pows = O.max_rz_arr_harmonics(data, numharm, rs, zs)
    p = O.calc_props(ders[0], ro, zo, 0.0)
    # within the statistical errors (Middleditch et al.; these are pessimistic
    # here: the local power holds the strong signal's own sidelobes) and far
    # inside the search grid's steps
    assert abs(ro - r) < min(3 * p.rerr, 2e-3)
    assert abs(zo - z) < min(3 * p.zerr, 2e-2)
    assert pows[0] == pytest.approx((N / 2) ** 2, rel=2e-2)
    assert ders[0].pow == pytest.approx(pows[0], rel=1e-2)
    assert p.pur == pytest.approx(1.0, abs=0.01)


def _single_grid_brute_force(data, rin, zin, win):
    """max_rzw_arr's grid by direct interpolation: r = int(rin) - 1 + k / 8
    (k < 24), z = zin - 2 + j / 2, w = win - 20 + 5 i (i, j < 9), with the
    kernel half-width of the grid's extreme z and w (LOWACC)."""
    zlo, zhi, wlo, whi = zin - 2.0, zin + 2.0, win - 20.0, win + 20.0
    khw = O.w_resp_halfwidth_acc(zhi if abs(zlo) < abs(zhi) else zlo, whi if abs(wlo) < abs(whi) else wlo, False)
    sb = int(rin) - 1
    p = np.zeros((9, 9, 24))
    for i in range(9):
        for j in range(9):
            for k in range(24):
                a = O.rzw_interp(data, sb + k / 8.0, zlo + j * (zhi - zlo) / 8, wlo + i * (whi - wlo) / 8, khw)
                p[i, j, k] = abs(a) ** 2
    return p, sb


@pytest.mark.parametrize("dz,dw", [(0.3, -4.0), (3.0, 0.0), (-0.6, 31.0)])
def test_single_harmonic_volume_is_its_grid_maximum(dz, dw):
    """max_rzw_arr (maximize_rzw.c:3): the raw-power maximum of a 9 x 9 x 24
    grid over z +- 2, w +- 20 and three bins from int(r) - 1, found here by
    brute force over direct interpolations.  An injection beyond the z (or w)
    range puts the maximum on the range's edge (a 17-point, +- 4 grid would
    follow it out); the power is raw (N A / 2)^2-scale, not local-power
    normalised; the derivatives carry the local power of the input point."""
    f0, z0, w = 4321.2, 7.0, 40.0
    data = chirp([(1.0, f0, z0, w)], noise=2.0, seed=21)
    r, z, ww = presto_rzw(f0, z0, w)
    rin, zin, win = r + 0.3, z - dz, ww - dw
    ro, zo, wo, d, maxpow = O.max_rzw_arr(data, rin, zin, win)
    p, sb = _single_grid_brute_force(data, rin, zin, win)
    i, j, k = np.unravel_index(np.argmax(p), p.shape)
    srt = np.sort(p.ravel())
    assert maxpow == pytest.approx(p.max(), rel=3e-3)
    if srt[-2] < 0.995 * srt[-1]:
        assert (ro, zo, wo) == pytest.approx((sb + k / 8.0, zin - 2 + j / 2.0, win - 20 + 5.0 * i), abs=1e-9)
    else:
        assert p[int(round((wo - win + 20) / 5)), int(round((zo - zin + 2) * 2)), int(round((ro - sb) * 8))] \
            >= 0.995 * p.max()
    if abs(dz) > 2.0:
        assert zo == pytest.approx(zin + np.sign(dz) * 2.0)
    elif abs(dw) > 20.0:
        assert wo == pytest.approx(win + np.sign(dw) * 20.0)
    else:
        assert abs(ro - r) <= 0.125 + 1e-9 and abs(zo - z) <= 0.5 + 0.2
    assert maxpow > 0.5 * (N / 2) ** 2
    assert d.locpow == np.float32(O.localpower3d(data, rin, zin, win))
    assert d.locpow != np.float32(O.localpower3d(data, ro, zo, wo))


@pytest.mark.parametrize("dr,dz", [(0.2, 0.7), (-0.4, -1.3)])
def test_single_harmonic_simplex_is_raw_power_peak(dr, dz):
    """max_rz_arr (maximize_rz.c:25): the simplex of max_rz_arr_harmonics on
    the raw power of one harmonic lands on the same peak, and returns that
    peak's raw power (the coherent (N A / 2)^2, not a local-power ratio) with
    the derivatives at the peak."""
    f0, z0 = 3310.6, -5.2
    data = chirp([(1.0, f0, z0, 0.0)], noise=0.3, seed=4)
    r, z = f0 + 0.5 * z0, z0
    ro, zo, d, pw = O.max_rz_arr(data, r + dr, z + dz)
    rh, zh, dh, ph = O.max_rz_arr_harmonics(data, 1, r + dr, z + dz)
    assert abs(ro - r) < 2e-3 and abs(zo - z) < 2e-2
    assert abs(ro - rh) < 1e-5 and abs(zo - zh) < 1e-4
    assert pw == pytest.approx((N / 2) ** 2, rel=2e-2)
    assert pw == pytest.approx(ph[0], rel=1e-6)
    assert d.locpow == np.float32(O.localpower3d(data, ro, zo, 0.0))
    assert d.pow == pytest.approx(pw, rel=1e-2)
