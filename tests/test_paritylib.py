"""The parity comparators themselves (CPU): a planted difference far from the
threshold must fail, one at the threshold must be counted and tolerated."""
import numpy as np
import pytest

from paritylib import compare_cands, compare_hits, pack_keys


class _Harm:
    numw, numz = 3, 5


class _Geom:
    numharmstages, corr_uselen = 2, 7


class _Plan:
    geom = _Geom()
    harms = [_Harm()]


def _hits():
    rng = np.random.default_rng(3)
    h = np.unique(np.stack([rng.integers(0, 4, 40), rng.integers(0, 2, 40), rng.integers(0, 3, 40),
                            rng.integers(0, 5, 40), rng.integers(0, 7, 40)], 1), axis=0)
    p = rng.uniform(20.0, 50.0, len(h)).astype(np.float32)
    return h, p


def _oracle_list(h, p):
    return [tuple(int(x) for x in hh) + (float(pp), 0.0) for hh, pp in zip(h, p)]


def test_keys_follow_scan_order():
    S, W, Z, R = 2, 3, 5, 7
    k = pack_keys([1, 0, 0], [0, 1, 0], [0, 0, 2], [0, 0, 0], [0, 0, 6], S, W, Z, R)
    assert list(k) == [((1 * S + 0) * W * Z) * R, ((0 * S + 1) * W * Z) * R, (2 * Z) * R + 6]


def test_identical_hits():
    h, p = _hits()
    near, worst, common = compare_hits(_Plan(), h, p, _oracle_list(h, p), [10.0, 10.0])
    assert (near, worst, common) == (0, 0.0, len(h))


def test_hit_at_threshold_is_counted():
    h, p = _hits()
    p2 = p.copy()
    p2[0] = 30.0
    cut = [30.0 * (1 + 1e-6), 30.0 * (1 + 1e-6)]  # hit 0 sits at the threshold
    near, worst, common = compare_hits(_Plan(), h[1:], p2[1:], _oracle_list(h, p2), cut)
    assert near == 1 and common == len(h) - 1


def test_hit_away_from_threshold_fails():
    h, p = _hits()
    with pytest.raises(AssertionError):
        compare_hits(_Plan(), h[1:], p[1:], _oracle_list(h, p), [10.0, 10.0])


def test_power_error_reported():
    h, p = _hits()
    p2 = p.copy()
    p2[3] *= 1.001
    _, worst, _ = compare_hits(_Plan(), h, p2, _oracle_list(h, p), [10.0, 10.0])
    assert abs(worst - 0.001 * p[3] / p[3]) < 1e-6


def _cands():
    return [(40.0, 7.5, 4, 1000.5, 2.0, 0.0), (30.0, 5.00005, 2, 2000.0, -4.0, 20.0),
            (60.0, 9.1, 8, 3000.25, 8.0, -40.0)]


def test_candidates_near_threshold_removed():
    c = _cands()
    assert compare_cands(c, c, 5.0) == 2
    assert compare_cands(c, [c[0], c[2]], 5.0) == 1


def test_candidates_differing_fail():
    c = _cands()
    with pytest.raises(AssertionError):
        compare_cands(c, [c[0], c[1]], 5.0)
    moved = list(c)
    moved[2] = (60.0, 9.1, 8, 3000.5, 8.0, -40.0)
    with pytest.raises(AssertionError):
        compare_cands(c, moved, 5.0)
