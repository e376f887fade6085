"""Comparison helpers for the GPU-vs-oracle parity tests.

Hits are compared as sets of scan-order keys (rstep, stage, w, z, r packed
the way the search encodes them); a hit present on one side only must lie
within TOL_POW (relative) of its stage's threshold.  Candidate lists are
compared after removing, from both, the candidates whose significance lies
within TOL_SIGMA of the search threshold: a hit that near the threshold may
pass on one side and not the other, and it can only create or displace a
candidate of that same significance.  The number removed is returned so the
tests can print it.
"""
import numpy as np

TOL_POW = 1e-5
TOL_SIGMA = 1e-4


def pack_keys(rstep, stage, w, z, r, S, W, Z, R):
    rstep, stage, w, z, r = (np.asarray(a, dtype=np.int64) for a in (rstep, stage, w, z, r))
    return (((rstep * S + stage) * W + w) * Z + z) * R + r


def gpu_hit_keys(plan, hits):
    g = plan.geom
    S, W, Z, R = g.numharmstages, max(plan.harms[0].numw, 1), plan.harms[0].numz, g.corr_uselen
    return pack_keys(hits[:, 0], hits[:, 1], hits[:, 2], hits[:, 3], hits[:, 4], S, W, Z, R), (S, W, Z, R)


def oracle_hit_keys(ohits, dims):
    S, W, Z, R = dims
    if isinstance(ohits, np.ndarray):
        return (pack_keys(ohits["rstep"], ohits["stage"], ohits["wi"], ohits["zi"], ohits["ri"], S, W, Z, R),
                np.asarray(ohits["power"], dtype=np.float64))
    a = np.asarray([h[:6] for h in ohits], dtype=np.float64).reshape(-1, 6)
    return pack_keys(a[:, 0], a[:, 1], a[:, 2], a[:, 3], a[:, 4], S, W, Z, R), a[:, 5]


def compare_hits(plan, hits_gpu, pows_gpu, ohits, powcut, tol=TOL_POW):
    """(near-threshold differences, worst relative power error on the common
    hits, common hits).  Asserts that every hit on one side only lies within
    tol of its stage's threshold."""
    kg, dims = gpu_hit_keys(plan, np.asarray(hits_gpu, dtype=np.int64).reshape(-1, 5))
    ko, po = oracle_hit_keys(ohits, dims)
    pg = np.asarray(pows_gpu, dtype=np.float64)
    ig, io = np.argsort(kg, kind="stable"), np.argsort(ko, kind="stable")
    kg, pg, ko, po = kg[ig], pg[ig], ko[io], po[io]
    assert np.all(np.diff(kg) > 0) and np.all(np.diff(ko) > 0), "duplicate hit keys"
    S, W, Z, R = dims
    cut = np.asarray(powcut, dtype=np.float64)
    near = 0
    for keys, pows, other, what in ((kg, pg, ko, "extra"), (ko, po, kg, "missing")):
        only = ~np.isin(keys, other, assume_unique=True)
        if only.any():
            stage = (keys[only] // (R * Z * W)) % S
            dev = np.abs(pows[only] - cut[stage]) / cut[stage]
            bad = np.nonzero(dev > tol)[0]
            assert bad.size == 0, ("%s hit away from the threshold" % what, int(keys[only][bad[0]]),
                                   float(pows[only][bad[0]]), float(cut[stage][bad[0]]))
            near += int(only.sum())
    common, cg_idx, co_idx = np.intersect1d(kg, ko, assume_unique=True, return_indices=True)
    worst = 0.0
    if common.size:
        worst = float(np.max(np.abs(pg[cg_idx] - po[co_idx]) / np.maximum(po[co_idx], 1.0)))
    return near, worst, int(common.size)


def compare_cands(cg, co, sigma_thr, tol_sigma=TOL_SIGMA, tol_pow=TOL_POW):
    """Asserts the candidate lists (power, sigma, numharm, r, z, w) are equal
    once the candidates within tol_sigma of the threshold are removed from
    both; returns how many were removed."""
    def far(c):
        return abs(c[1] - sigma_thr) > tol_sigma

    fg = [c for c in cg if far(c)]
    fo = [c for c in co if far(c)]
    removed = (len(cg) - len(fg)) + (len(co) - len(fo))

    def key(c):
        return (round(c[3], 6), round(c[4], 6), round(c[5], 6), c[2])

    kg, ko = [key(c) for c in fg], [key(c) for c in fo]
    if kg != ko:
        sg, so = set(kg), set(ko)
        raise AssertionError("candidate lists differ: %d only on the GPU (%s), %d only in the oracle (%s)"
                             % (len(sg - so), sorted(sg - so)[:4], len(so - sg), sorted(so - sg)[:4]))
    for a, b in zip(fg, fo):
        assert abs(a[0] - b[0]) <= tol_pow * max(b[0], 1.0), (a, b)
        assert abs(a[1] - b[1]) <= tol_sigma * max(b[1], 1.0), (a, b)
    return removed


def injection_rstep(obs, N, dt, sig):
    """The rstep holding an injected signal's mean Fourier frequency."""
    amp, f0, z, w = sig
    r_mean = f0 * N * dt + z / 2.0 + w / 6.0
    return int((r_mean - obs.rlo) // (obs.corr_uselen // 2))
