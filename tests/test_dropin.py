"""The accel.h drop-in API (libaccelsearch_lib.so) through ctypes, as the
reference's callers bind it (accelsearch_cu.c:22-34,
bin/accelsearch_pipeline_cu.py:51-233).

CPU tests: command-line parsing, create_accelobs / create_subharminfos
against the oracle's geometry, the candidate-list operations against the
oracle's restatement.  GPU tests: the PRESTO C per-rstep loop
(accelsearch.c:183-223) over subharm_fderivs_vol / add_ffdotpows /
search_ffdotpows, and the accelsearch command-line driver end to end.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import dropin as D
import pyoracle as O
from prestozl_amd import binding, synth

DRIVER = os.path.join(os.path.dirname(binding.__file__), "lib", "accelsearch")


def parse(args):
    n, av = D.argv(["accelsearch"] + args)
    return D.api().parseCmdline(n, av).contents, av


def test_parse_defaults():
    cmd, _av = parse(["x.fft"])
    # defaults of accelsearch_cmd.c:20-96
    assert (cmd.batchsizeP, cmd.batchsize) == (b"\x01", 8)
    assert (cmd.numharmP, cmd.numharm) == (b"\x01", 8)
    assert (cmd.zmaxP, cmd.zmax) == (b"\x01", 200)
    assert (cmd.wmaxP, cmd.wmax) == (b"\x00", 0)
    assert cmd.sigma == pytest.approx(2.0)
    assert (cmd.floP, cmd.flo, cmd.fhiP, cmd.fhi) == (b"\x01", 1.0, b"\x01", 10000.0)
    assert (cmd.rloP, cmd.rhiP) == (b"\x00", b"\x00")
    assert cmd.argc == 1 and cmd.argv[0] == b"x.fft"
    assert cmd.full_cmd_line == b"accelsearch x.fft"


def test_parse_values():
    cmd, _av = parse(["-zmax", "51", "-wmax", "150", "-numharm", "16", "-sigma", "5.0", "-batchsize", "12",
                      "-rlo", "100", "-noharmremove", "-ncpus", "4", "f.fft"])
    assert cmd.zmax == 51 and cmd.wmaxP == b"\x01" and cmd.wmax == 150
    assert cmd.numharm == 16 and cmd.sigma == pytest.approx(5.0) and cmd.batchsize == 12
    assert cmd.rloP == b"\x01" and cmd.rlo == 100.0 and cmd.floP == b"\x01"  # -flo stays on (default)
    assert cmd.noharmremoveP == b"\x01" and cmd.ncpus == 4
    assert cmd.argc == 1 and cmd.argv[0] == b"f.fft"


@pytest.mark.parametrize("args", [["-numharm", "64"], ["-zmax", "1201"], ["-sigma", "0.5"], ["-wmax", "-20"],
                                  ["-bogus"]])
def test_parse_rejects(args):
    r = subprocess.run([DRIVER] + args + ["x.fft"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "option" in r.stdout


def open_obs(root, args, usemmap=1):
    cmd, av = parse(args + [root + ".fft"])
    obs, idata = D.AccelObs(), D.InfoData()
    D.api().create_accelobs(C.byref(obs), C.byref(idata), C.byref(cmd), usemmap)
    return obs, idata, cmd, av


GEOM_CASES = [
    # (log2 N, dt, zmax, wmax, numharm, sigma)
    (20, 64e-6, 20, 0, 4, 2.0),
    (21, 1e-4, 200, 300, 16, 5.0),
    (19, 6.4e-4, 200, 300, 8, 3.0),
    (20, 1e-4, 50, None, 8, 2.0),
    (21, 1e-4, 51, 150, 2, 6.0),
]


@pytest.mark.parametrize("case", GEOM_CASES)
def test_create_accelobs_matches_oracle(tmp_path, case):
    lg, dt, zmax, wmax, numharm, sigma = case
    N = 1 << lg
    fft = np.zeros(N // 2, dtype=np.complex64)
    root = D.setup_files(tmp_path, fft, N, dt)
    args = ["-zmax", str(zmax), "-numharm", str(numharm), "-sigma", str(sigma)]
    if wmax is not None:
        args += ["-wmax", str(wmax)]
    obs, idata, cmd, _av = open_obs(root, args)
    ref = O.create_accelobs(O.make_params(N, dt, N // 2, zmax, wmax or 0, numharm, sigma,
                                          wmaxP=wmax is not None))
    assert idata.N == N and idata.dt == pytest.approx(dt, rel=1e-15)
    for f in ("N", "numbins", "highestbin", "maxkernlen", "corr_uselen", "fftlen", "numharmstages", "numz",
              "numw"):
        assert getattr(obs, f) == getattr(ref, f), f
    for f in ("T", "rlo", "rhi", "zlo", "zhi", "wlo", "whi"):
        assert getattr(obs, f) == getattr(ref, f), f
    for s in range(obs.numharmstages):
        assert obs.powcut[s] == ref.powcut[s]
        assert obs.numindep[s] == ref.numindep[s]
    assert obs.inmem == 0 and obs.dr == 0.5 and obs.numbetween == 2
    assert obs.candnm.decode().startswith(root + "_ACCEL_")
    # the .fft is mapped, not copied
    assert obs.mmap_file > 0 and obs.fft[5].r == 0.0

    shis = D.api().create_subharminfos(C.byref(obs))
    for stage in range(obs.numharmstages):
        harms = [1] if stage == 0 else list(range(1, 1 << stage, 2))
        for j in harms:
            sh = shis[stage][0 if stage == 0 else j - 1]
            h = O.harm_info(ref, 1 << stage, j)
            assert (sh.numharm, sh.harmnum, sh.zmax, sh.wmax) == (1 << stage, j, h.zmax, h.wmax)
            assert sh.numkern_zdim == h.numz and sh.numkern_wdim == h.numw
            assert sh.numkern == h.numz * h.numw
            k0 = sh.kern[0][0]
            assert (k0.z, k0.w) == (-h.zmax, -h.wmax)
            assert k0.fftlen == h.fftlen and k0.kern_half_width == h.halfwidth
            assert k0.numgoodbins == h.fftlen - 2 * 2 * h.halfwidth
            kl = sh.kern[h.numw - 1][h.numz - 1]
            assert (kl.z, kl.w) == (h.zmax, h.wmax)
            assert not kl.data  # the FFT'd kernels live in GPU memory
    D.api().free_subharminfos(C.byref(obs), shis)
    D.api().free_accelobs(C.byref(obs))


def test_get_fourier_amplitudes(tmp_path):
    N = 1 << 14
    fft = (np.arange(N // 2) + 1j * np.arange(N // 2) * 0.5).astype(np.complex64)
    root = D.setup_files(tmp_path, fft, N, 1e-3)
    obs, _i, _c, _av = open_obs(root, ["-zmax", "10", "-numharm", "2"])
    for lobin, n in [(-40, 100), (0, 64), (1000, 512), (N // 2 - 30, 100), (N // 2 + 5, 10)]:
        p = D.api().get_fourier_amplitudes(lobin, n, C.byref(obs))
        got = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(2 * n,))
        got = got[0::2] + 1j * got[1::2]
        exp = np.zeros(n, dtype=np.complex64)
        for i in range(n):
            b = lobin + i
            if 0 <= b < N // 2:
                exp[i] = fft[b]
        np.testing.assert_array_equal(got, exp)
    D.api().free_accelobs(C.byref(obs))


def random_cands(rng, n):
    """Candidates with harmonic families, near-misses and sigma ties."""
    base = rng.uniform(50, 5000, size=n // 3)
    rs = list(base)
    for r in base[: n // 3]:
        kind = rng.integers(0, 4)
        if kind == 0:
            rs.append(r * rng.integers(2, 17) + rng.uniform(-1.4, 1.4))
        elif kind == 1:
            p, q = O.HARMONIC_RATIOS[rng.integers(len(O.HARMONIC_RATIOS))]
            rs.append(r * p / q + rng.uniform(-1.6, 1.6))
        elif kind == 2:
            rs.append(r / rng.integers(2, 17) + rng.uniform(-1.4, 1.4))
        else:
            rs.append(rng.uniform(50, 5000))
    rs = rs[:n]
    sig = rng.choice([3.0, 4.5, 6.25, 9.0], size=len(rs)) + np.round(rng.uniform(0, 2, size=len(rs)), 1)
    pw = rng.choice([10.0, 20.0, 30.0], size=len(rs))
    return [(float(pw[i]), float(sig[i]), int(rng.choice([1, 2, 4, 8, 16])), float(rs[i]),
             float(rng.integers(-50, 50) * 2), float(rng.integers(-5, 5) * 20)) for i in range(len(rs))]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_sort_and_eliminate_harmonics(seed):
    rng = np.random.default_rng(seed)
    cands = random_cands(rng, 300)
    lst = D.make_list(cands)
    lst = D.api().sort_accelcands(lst)
    got_sorted = D.read_list(lst)
    f32 = lambda c: (float(np.float32(c[0])), float(np.float32(c[1])), c[2], c[3], c[4], c[5])
    exp_sorted = O.sort_cands([f32(c) for c in cands])
    assert got_sorted == exp_sorted
    n = C.c_int(len(got_sorted))
    lst = D.api().eliminate_harmonics(lst, C.byref(n))
    got = D.read_list(lst)
    exp = O.eliminate_harmonics(exp_sorted)
    assert got == exp and n.value == len(exp)
    assert len(exp) < len(exp_sorted)
    D.free_list(lst)


def test_eliminate_oracle_pinned():
    """Known answer worked by hand, plus the defining invariant checked by
    brute force: no survivor is related to an earlier survivor, and every
    removed candidate is related to an earlier survivor."""
    c = lambda s, r: (10.0, s, 1, r, 0.0, 0.0)
    cands = [c(9, 100.0), c(8, 200.0), c(7, 150.0), c(6, 137.0), c(5, 33.6), c(4, 41.0), c(3, 1600.9),
             c(2, 274.0)]
    # 200 = 2x100, 150 = 3/2x100, 33.6 ~ 100/3 (|33.33-33.6| < 1.5), 1600.9 ~ 16x100;
    # 41.0 = 2/5 x 100 + 1.0; 274 = 2 x 137 (removed by the survivor 137)
    assert [x[3] for x in O.eliminate_harmonics(cands)] == [100.0, 137.0]
    rng = np.random.default_rng(7)
    cands = O.sort_cands(random_cands(rng, 200))
    out = O.eliminate_harmonics(cands)
    kept = set(id(x) for x in out)
    for i, a in enumerate(out):
        for b in out[i + 1:]:
            assert not O.harmonically_related(a[3], b[3])
    for x in cands:
        if id(x) in kept:
            continue
        pos = cands.index(x)
        assert any(O.harmonically_related(s[3], x[3]) for s in out if cands.index(s) < pos)


def test_sigma_wrappers():
    L = D.api()
    for sigma, numsum, ntr in [(5.0, 16, 1e9), (2.0, 1, 1e5), (8.0, 4, 3e7)]:
        p = L.power_for_sigma(sigma, numsum, ntr)
        assert p == pytest.approx(O.power_for_sigma(sigma, numsum, ntr), rel=1e-12)
        # candidate_sigma switches to an asymptotic series above chi2/dof = 15, so
        # the round trip is only close; the oracle gives the same value
        assert L.candidate_sigma(p, numsum, ntr) == pytest.approx(O.candidate_sigma(p, numsum, ntr), abs=1e-9)
        assert L.candidate_sigma(p, numsum, ntr) == pytest.approx(sigma, abs=1e-3)


def test_candidate_sigma_dense_grid():
    """The significance the GPU gives every hit (pzl_stats.h, here its host
    build behind accel.h's candidate_sigma) against the oracle's restatement of
    the reference (pinned to the reference's dcdflib) on a dense grid: from the
    detection threshold of each harmonic count to strong signals, both the
    Poisson-sum and the asymptotic chi-square branches, the inverse normal tail
    started from AS 241.  The stored sigma is a float: it must come out the
    same float."""
    L = D.api()
    n = same = 0
    worst = 0.0
    for numsum in (1, 2, 4, 8, 16, 32):
        for ntr in (1.0, 3.3e5, 1.4e10):
            p0 = O.power_for_sigma(2.0, numsum, ntr)
            for power in np.concatenate([p0 * np.linspace(1.0, 4.0, 700), p0 * np.logspace(0.61, 2.5, 120)]):
                a = L.candidate_sigma(float(power), numsum, ntr)
                b = O.candidate_sigma(float(power), numsum, ntr)
                worst = max(worst, abs(a - b) / max(b, 1.0))
                same += np.float32(a) == np.float32(b)
                n += 1
    print("%d points: max rel diff %.3g, same float %d" % (n, worst, same))
    assert worst < 1e-12
    assert same >= n - 2


# --------------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_presto_loop_through_plane_api(tmp_path, config1_data):
    """accelsearch.c:183-223 driven through subharm_fderivs_vol, add_ffdotpows
    and search_ffdotpows: the candidate list equals the oracle's search of the
    same rsteps."""
    fft, N, dt, opts = config1_data
    root = D.setup_files(tmp_path, fft, N, dt)
    obs, _i, cmd, _av = open_obs(root, ["-zmax", "20", "-wmax", "0", "-numharm", "4", "-sigma", "2.0"])
    L = D.api()
    shis = L.create_subharminfos(C.byref(obs))
    rstep = obs.corr_uselen * 0.5
    lo, hi = 5, 12
    cands = C.POINTER(D.GSList)()
    for k in range(lo, hi):
        startr = obs.rlo + k * rstep
        lastr = startr + rstep - 0.5
        fund = L.subharm_fderivs_vol(1, 1, startr, lastr, C.byref(shis[0][0]), C.byref(obs))
        cands = L.search_ffdotpows(fund, 1, C.byref(obs), cands)
        for stage in range(1, obs.numharmstages):
            h2s = 1 << stage
            for harm in range(1, h2s, 2):
                sub = L.subharm_fderivs_vol(h2s, harm, startr, lastr, C.byref(shis[stage][harm - 1]),
                                            C.byref(obs))
                L.add_ffdotpows(fund, sub, h2s, harm)
                L.free_ffdotpows(sub)
            cands = L.search_ffdotpows(fund, h2s, C.byref(obs), cands)
        L.free_ffdotpows(fund)
    got = D.read_list(cands)
    ref_obs = O.create_accelobs(O.make_params(N, dt, len(fft), 20, 0, 4, 2.0, wmaxP=True))
    ctx = O.Ctx(ref_obs, O.FP32, nthreads=8)
    exp, _hits, _ = ctx.search(fft, lo, hi)
    ctx.close()
    print("plane API: %d candidates, oracle %d" % (len(got), len(exp)))
    assert [(c[2], c[3], c[4], c[5]) for c in got] == [(c[2], c[3], c[4], c[5]) for c in exp]
    for g, e in zip(got, exp):
        assert g[0] == pytest.approx(e[0], rel=1e-5)
        assert g[1] == pytest.approx(e[1], abs=1e-4)
    D.free_list(cands)
    L.free_subharminfos(C.byref(obs), shis)
    L.free_accelobs(C.byref(obs))


FOURIERPROPS = np.dtype([("r", "<f8"), ("rerr", "<f4"), ("z", "<f8"), ("zerr", "<f4"), ("w", "<f8"), ("werr", "<f4")] +
                        [(k, "<f4") for k in ("pow", "powerr", "sig", "rawpow", "phs", "phserr", "cen", "cenerr",
                                              "pur", "purerr", "locpow")], align=True)  # presto.h:171-189


def _accel_tables(path):
    """Rows of the fundamentals and harmonics tables of an ACCEL text file."""
    lines = open(path).read().split("\n")
    tables, cur, in_table = [], [], False
    for ln in lines:
        if ln.startswith("----"):
            in_table, cur = True, []
            continue
        if in_table:
            if not ln.strip():
                tables.append(cur)
                in_table = False
            else:
                cur.append(ln)
    return tables


@pytest.mark.gpu
def test_driver_end_to_end(tmp_path, config1_data):
    """The accelsearch driver (CPU1 -> GPU -> CPU2) on a config-1 file (w = 0):
    its sorted, harmonic-filtered list equals the oracle's, and every
    candidate's .cand record holds the fundamental's properties after the
    (r, z) simplex optimisation (maximize_rz.c:141-220), as the oracle's."""
    fft, N, dt, opts = config1_data
    root = D.setup_files(tmp_path, fft, N, dt)
    r = subprocess.run([DRIVER, "-zmax", "20", "-wmax", "0", "-numharm", "4", "-sigma", "2.0", root + ".fft"],
                       capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    base = root + "_ACCEL_20"  # wmax 0: numw = 0, no _JERK_ (accel_utils.c:1873-1893)
    props = np.fromfile(base + ".cand", dtype=FOURIERPROPS)
    fund, harm = _accel_tables(base)[:2]
    numharm = [int(row.split()[4]) for row in fund]
    assert "FFT 'w'" not in open(base).read()  # no jerk columns
    ref_obs = O.create_accelobs(O.make_params(N, dt, len(fft), 20, 0, 4, 2.0, wmaxP=True))
    ctx = O.Ctx(ref_obs, O.FP32, nthreads=8)
    exp, _hits, _ = ctx.search(fft)
    ctx.close()
    exp = O.eliminate_harmonics(O.sort_cands(exp))
    print("driver: %d candidates, oracle %d" % (len(props), len(exp)))
    assert len(props) == len(exp) == len(fund) and len(harm) == sum(numharm)
    assert [(h, round(p["r"], 4), round(p["z"], 4), round(p["w"], 4)) for h, p in zip(numharm, props)] == \
           [(c[2], round(c[3], 4), round(c[4], 4), round(c[5], 4)) for c in exp]
    for p, e in zip(props, exp):
        ro, zo, ders, pows = O.max_rz_arr_harmonics(fft, e[2], e[3], e[4])
        q = O.calc_props(ders[0], e[3], e[4], e[5])
        # the two simplexes stop at optima that agree to the flatness of the
        # summed peak (test_gpu_optimize.py); a noise-dominated fundamental's
        # power moves by up to ~1e-3 over that distance
        assert p["pow"] == pytest.approx(q.pow, rel=5e-3) and p["sig"] == pytest.approx(q.sig, rel=5e-3)
        assert (p["rerr"], p["zerr"]) == (np.float32(0.5 / e[2]), np.float32(2.0 / e[2]))


@pytest.mark.gpu
def test_driver_jerk_search_outputs(tmp_path):
    """CPU2 of a w > 0 search: each candidate optimised on the GPU (harmonic
    polishing), its fundamental's Fourier properties in the binary .cand file
    (fourierprops records) and the ACCEL text tables.  The .cand records equal
    the oracle's calc_props of its optimisation of the same candidates (the
    tolerances of tests/test_gpu_optimize.py)."""
    from prestozl_amd import synth
    N, dt = 1 << 20, 1e-4
    sigs = [(0.06, 123.4, 8.0, 30.0), (0.04, 2 * 123.4, 16.0, 60.0)]
    x = synth.timeseries(N, dt, sigs, seed=21)
    fft = synth.realfft_packed(x)
    root = D.setup_files(tmp_path, fft, N, dt, name="jerk")
    r = subprocess.run([DRIVER, "-zmax", "20", "-wmax", "40", "-numharm", "2", "-sigma", "3.0", root + ".fft"],
                       capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout + r.stderr
    base = root + "_ACCEL_20_JERK_40"
    props = np.fromfile(base + ".cand", dtype=FOURIERPROPS)
    assert FOURIERPROPS.itemsize == 88 and len(props) > 0
    fund, harm = _accel_tables(base)[:2]
    assert len(fund) == len(props)
    numharm = [int(row.split()[4]) for row in fund]
    assert len(harm) == sum(numharm)
    assert "Freq Deriv" in open(base).read() and "FFT 'w'" in open(base).read()
    for p, nh in zip(props, numharm):
        ro, zo, wo, ders, pows = O.max_rzw_arr_harmonics(fft, nh, p["r"], p["z"], p["w"])
        e = O.calc_props(ders[0], p["r"], p["z"], p["w"])
        print("cand r %.3f z %.2f w %.2f nh %d: pow %.4f / %.4f  pur %.4f / %.4f" %
              (p["r"], p["z"], p["w"], nh, p["pow"], e.pow, p["pur"], e.pur))
        assert (p["rerr"], p["zerr"], p["werr"]) == (np.float32(0.5 / nh), np.float32(2.0 / nh), np.float32(20.0 / nh))
        for k in ("pow", "rawpow", "sig", "locpow"):
            assert p[k] == pytest.approx(getattr(e, k), rel=1e-4), k
        assert abs(p["phs"] - e.phs) < 1e-4
        assert p["pur"] == pytest.approx(e.pur, rel=3e-2)
        assert abs(p["cen"] - e.cen) < 1e-3


@pytest.mark.gpu
def test_driver_noharmpolish_outputs(tmp_path):
    """CPU2 with -noharmpolish (accel_utils.c:610-652): every harmonic of
    every candidate maximised on its own, max_rzw_arr at (r, z, w) * (ii + 1).
    The .cand records hold the fundamental's properties from its own maximum,
    and each harmonics-table row's 'r' is that harmonic's own maximum, as the
    oracle's max_rzw_arr gives them."""
    from prestozl_amd import synth
    N, dt = 1 << 20, 1e-4
    sigs = [(0.06, 123.4, 8.0, 30.0), (0.04, 2 * 123.4, 16.0, 60.0)]
    x = synth.timeseries(N, dt, sigs, seed=21)
    fft = synth.realfft_packed(x)
    root = D.setup_files(tmp_path, fft, N, dt, name="nhp")
    r = subprocess.run([DRIVER, "-zmax", "20", "-wmax", "40", "-numharm", "2", "-sigma", "3.0", "-noharmpolish",
                        root + ".fft"], capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout + r.stderr
    base = root + "_ACCEL_20_JERK_40"
    props = np.fromfile(base + ".cand", dtype=FOURIERPROPS)
    fund, harm = _accel_tables(base)[:2]
    numharm = [int(row.split()[4]) for row in fund]
    assert len(props) == len(fund) > 0 and len(harm) == sum(numharm)
    hrows = iter(harm)
    for p, nh in zip(props, numharm):
        for ii in range(nh):
            ro, zo, wo, d, maxpow = O.max_rzw_arr(fft, p["r"] * (ii + 1), p["z"] * (ii + 1), p["w"] * (ii + 1))
            tok = next(hrows).split()
            t0 = 1 if ii == 0 else 0  # the candidate number opens its first row
            assert int(tok[t0]) == ii + 1
            txt = tok[t0 + 4].split("(")[0]  # printed to the digits of its error
            hr = float(txt)
            dec = len(txt.split(".")[1]) if "." in txt else 0
            print("cand r %.3f harm %d: table r %s, oracle %.4f" % (p["r"], ii + 1, txt, ro))
            assert abs(hr - ro) <= 0.5 * 10.0 ** -dec + 1e-9
            if ii == 0:
                e = O.calc_props(d, p["r"], p["z"], p["w"])
                for k in ("pow", "rawpow", "sig", "locpow"):
                    assert p[k] == pytest.approx(getattr(e, k), rel=1e-4), k


@pytest.mark.gpu
def test_driver_noharmpolish_fdot_only(tmp_path, config1_data):
    """-noharmpolish in a w = 0 search (config 1): max_rz_arr per harmonic
    (maximize_rz.c:25-85).  Each .cand record holds the fundamental's
    properties at its own simplex optimum, as the oracle's max_rz_arr gives
    them (the simplexes agree to the flatness of the peak: 5e-3 in power)."""
    fft, N, dt, opts = config1_data
    root = D.setup_files(tmp_path, fft, N, dt, name="nhp0")
    r = subprocess.run([DRIVER, "-zmax", "20", "-wmax", "0", "-numharm", "4", "-sigma", "2.0", "-noharmpolish",
                        root + ".fft"], capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout + r.stderr
    base = root + "_ACCEL_20"
    props = np.fromfile(base + ".cand", dtype=FOURIERPROPS)
    fund, harm = _accel_tables(base)[:2]
    numharm = [int(row.split()[4]) for row in fund]
    assert len(props) == len(fund) > 0 and len(harm) == sum(numharm)
    for p in props[:12]:
        ro, zo, d, pw = O.max_rz_arr(fft, p["r"], p["z"])
        q = O.calc_props(d, p["r"], p["z"], 0.0)
        assert p["pow"] == pytest.approx(q.pow, rel=5e-3) and p["rawpow"] == pytest.approx(q.rawpow, rel=5e-3)


PIPELINE = os.path.join(os.path.dirname(binding.__file__), "lib", "accelsearch_pipeline")


def _same_shape_files(dirpath, n=5):
    """n spectra of one shape (N = 2^20, dt = 1e-4 s), each with its own
    jerked sinusoid, written as obs<k>.fft/.inf into dirpath."""
    os.makedirs(dirpath, exist_ok=True)
    names = []
    for k in range(n):
        N, dt = 1 << 20, 1e-4
        x = synth.timeseries(N, dt, [(0.05 + 0.01 * k, 90.0 + 37.0 * k, 6.0 - 3.0 * k, 20.0 * (k - 2))], seed=500 + k)
        D.setup_files(dirpath, synth.realfft_packed(x), N, dt, name="obs%d" % k)
        names.append("obs%d.fft" % k)
    return names


@pytest.mark.gpu
def test_pipeline_threads_equal_serial(tmp_path):
    """The reference pipeline's thread overlap (bin/accelsearch_pipeline_cu.py:
    296-407): CPU1, GPU and CPU2 of different files of the same shape run at
    once on three threads of one process.  Every file's .cand records and
    ACCEL text equal those of the command-line driver run on it alone."""
    opts = ["-zmax", "20", "-wmax", "40", "-numharm", "4", "-sigma", "3.0"]
    serial, piped = str(tmp_path / "serial"), str(tmp_path / "piped")
    names = _same_shape_files(serial)
    _same_shape_files(piped)
    for nm in names:
        r = subprocess.run([DRIVER] + opts + [nm], capture_output=True, text=True, timeout=300, cwd=serial)
        assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([PIPELINE] + opts + ["--"] + names, capture_output=True, text=True, timeout=300, cwd=piped)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout + r.stderr
    ncands = 0
    for nm in names:
        base = nm[:-4] + "_ACCEL_20_JERK_40"
        for suffix in ("", ".cand"):
            a = open(os.path.join(serial, base + suffix), "rb").read()
            b = open(os.path.join(piped, base + suffix), "rb").read()
            assert a == b, (nm, suffix)
        ncands += len(np.fromfile(os.path.join(piped, base + ".cand"), dtype=np.uint8)) // 88
    print("%d files, %d candidates, outputs identical to the serial runs" % (len(names), ncands))
    assert ncands > len(names)


@pytest.mark.gpu
def test_spectrum_read_whole_usemmap0(tmp_path):
    """create_accelobs(..., usemmap = 0) reads the spectrum whole instead of
    mapping it; the GPU stage and CPU2's candidate optimisation work from that
    copy, and the outputs equal the command-line driver's (which maps)."""
    opts = ["-zmax", "20", "-wmax", "40", "-numharm", "4", "-sigma", "3.0"]
    a_dir, b_dir = str(tmp_path / "mapped"), str(tmp_path / "read")
    nm = _same_shape_files(a_dir, 1)[0]
    _same_shape_files(b_dir, 1)
    r = subprocess.run([DRIVER] + opts + [nm], capture_output=True, text=True, timeout=300, cwd=a_dir)
    assert r.returncode == 0, r.stdout + r.stderr
    L = D.api()
    L.accelsearch_GPU.restype = C.c_int
    L.accelsearch_GPU.argtypes = [D.AccelObs, C.POINTER(C.POINTER(D.SubharmInfo)), C.POINTER(C.POINTER(D.GSList)),
                                  C.POINTER(D.Cmdline)]
    L.accelsearch_CPU2.argtypes = [C.POINTER(C.POINTER(D.GSList)), C.POINTER(D.AccelObs), C.POINTER(D.InfoData),
                                   C.POINTER(D.Cmdline)]
    cwd = os.getcwd()
    os.chdir(b_dir)
    try:
        n, av = D.argv(["accelsearch"] + opts + [nm])
        cmd = L.parseCmdline(n, av)
        obs, idata = D.AccelObs(), D.InfoData()
        L.create_accelobs(C.byref(obs), C.byref(idata), cmd, 0)
        assert obs.mmap_file == 0 and bool(obs.fft)
        shi = L.create_subharminfos(C.byref(obs))
        cands = C.POINTER(D.GSList)()
        assert L.accelsearch_GPU(obs, shi, C.byref(cands), cmd) == 0
        L.accelsearch_CPU2(C.byref(cands), C.byref(obs), C.byref(idata), cmd)
    finally:
        os.chdir(cwd)
    base = nm[:-4] + "_ACCEL_20_JERK_40"
    for suffix in ("", ".cand"):
        assert open(os.path.join(a_dir, base + suffix), "rb").read() == \
            open(os.path.join(b_dir, base + suffix), "rb").read(), suffix
