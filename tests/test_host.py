"""Host logic of the product library, without a GPU.

* every entry point include/*.h declares is exported by libaccelsearch_lib.so;
* the search geometry and thresholds equal the oracle's (create_accelobs);
* the candidate list (insert_new_accelcand semantics) equals the oracle's
  literal linked-list walk on random and adversarial hit sequences.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as O
from prestozl_amd import binding as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"^\s*#[^\n]*", "", text, flags=re.M)
        text = re.sub(r"typedef struct[^;]*?\{.*?\}[^;]*;", "", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof", "matrix_3d_index"):
                names.add(name)
    return names


def test_library_exports_declared_symbols():
    assert os.path.exists(B.LIBPATH), "build first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIBPATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(n for n in declared_functions() if n not in exported)
    assert not missing, missing


CONFIGS = [
    (1 << 20, 64e-6, 20, 0, 4, 2.0, True),
    (1 << 24, 1e-4, 200, 300, 16, 5.0, True),
    (1 << 23, 1e-4, 200, 500, 16, 5.0, True),
    (1 << 21, 1e-4, 50, 0, 8, 5.0, True),
    (1 << 22, 1e-4, 200, 150, 16, 5.0, True),
    (262144, 6.4e-4, 200, 300, 8, 5.0, True),
    (1 << 20, 1e-4, 37, 55, 2, 3.5, True),   # zmax/wmax rounded up
    (1 << 20, 1e-4, 12, 0, 1, 2.0, False),   # no -wmax: acceleration search
]


@pytest.mark.parametrize("cfg", CONFIGS)
def test_geometry_equals_oracle(cfg):
    N, dt, zmax, wmax, nh, sig, wmaxP = cfg
    obs = O.create_accelobs(O.make_params(N, dt, N // 2, zmax, wmax, nh, sig, wmaxP=wmaxP))
    g = B.geometry(B.make_config(N, dt, N // 2, zmax, wmax, nh, sig, wmaxP=wmaxP))
    assert (g.fftlen, g.corr_uselen, g.maxkernlen, g.numz, g.numw, g.numharmstages) == \
        (obs.fftlen, obs.corr_uselen, obs.maxkernlen, obs.numz, obs.numw, obs.numharmstages)
    assert (g.rlo, g.rhi, g.highestbin) == (obs.rlo, obs.rhi, obs.highestbin)
    assert g.nrsteps == O.num_rsteps(obs)
    assert list(g.numindep)[: g.numharmstages] == list(obs.numindep)[: obs.numharmstages]
    assert list(g.powcut)[: g.numharmstages] == list(obs.powcut)[: obs.numharmstages]


def test_correlation_lengths_fit_the_segment_transform():
    """Every correlation length the library supports is at most 4096 points:
    a segment holds half of it, and the segment transform's registers are
    sized for 2048 points.  The fundamental length of every configured shape
    is supported (including config 3's 4096)."""
    L = B.lib()
    L.pzl_corr_supported.argtypes = [C.c_int]
    L.pzl_corr_supported.restype = C.c_int
    supported = [n for n in range(1, 16385) if L.pzl_corr_supported(n)]
    assert supported and max(supported) == 4096, supported[-5:]
    for N, dt, zmax, wmax, nh, sig, wmaxP in CONFIGS:
        g = B.geometry(B.make_config(N, dt, N // 2, zmax, wmax, nh, sig, wmaxP=wmaxP))
        assert L.pzl_corr_supported(g.fftlen), g.fftlen


def _insert_both(rr, sig, pw, nh, z, w):
    L = B.lib()
    n = len(rr)
    out = (B.Cand * max(n, 1))()
    nout = C.c_longlong()
    L.pzl_insert_cands.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_longlong,
                                   C.POINTER(B.Cand), C.c_longlong, C.POINTER(C.c_longlong)]
    arrs = [np.ascontiguousarray(a, dtype=t) for a, t in
            ((rr, np.float64), (sig, np.float32), (pw, np.float32), (nh, np.int32), (z, np.float64), (w, np.float64))]
    ptrs = [a.ctypes.data_as(C.POINTER(C.c_double if a.dtype == np.float64 else
                                       C.c_float if a.dtype == np.float32 else C.c_int)) for a in arrs]
    assert L.pzl_insert_cands(*ptrs, n, out, n, C.byref(nout)) == 0
    prod = [(out[i].power, out[i].sigma, out[i].numharm, out[i].r, out[i].z, out[i].w) for i in range(nout.value)]
    orac = O.insert_cands(*arrs)
    return prod, orac


@pytest.mark.parametrize("seed", range(12))
def test_candidate_list_equals_linked_list(seed):
    rng = np.random.default_rng(seed)
    n = 4000
    # clustered hits as a search produces them: runs of increasing r within
    # rows, restarting, with a few strong clusters and ties in sigma
    centres = rng.uniform(0, 3000, size=25)
    rr = (rng.choice(centres, size=n) + rng.normal(0, 20, size=n))
    rr = np.round(rr * 4) / 4
    sig = np.round(rng.uniform(2, 12, size=n), 1)  # many exact ties
    pw = sig * 3
    nh = rng.choice([1, 2, 4, 8], size=n)
    z = rng.uniform(-100, 100, size=n)
    w = rng.uniform(-300, 300, size=n)
    prod, orac = _insert_both(rr, sig, pw, nh, z, w)
    assert prod == orac


def test_candidate_list_out_of_order_overwrite():
    """An overwrite of the previous node can leave the list unsorted (the last
    node stays below it); the walk must then follow the literal order."""
    rr = [100.0, 110.0, 120.0, 112.0, 111.0, 118.0, 95.0, 126.0, 119.5, 113.0]
    sig = [5.0, 9.0, 4.0, 6.0, 10.0, 3.0, 8.0, 7.0, 11.0, 2.0]
    n = len(rr)
    prod, orac = _insert_both(rr, sig, [1.0] * n, [1] * n, [0.0] * n, [0.0] * n)
    assert prod == orac
    rng = np.random.default_rng(99)
    for _ in range(200):
        m = 60
        rr = np.round(rng.uniform(0, 80, size=m) * 2) / 2
        sig = np.round(rng.uniform(1, 6, size=m), 0)
        prod, orac = _insert_both(rr, sig, sig, [1] * m, [0.0] * m, [0.0] * m)
        assert prod == orac


@pytest.mark.parametrize("fn,val,err,exp", [
    # one-digit error (output.c:6): rounded at the error's decade
    (1, 1.2345, 0.012, " 1.23(1)"),
    (1, 1234.5, 12.0, " 1.23(1)x10^3"),
    (1, -0.5, 0.03, "-0.50(3)"),
    (1, 3.14159, 0.096, " 3.1(1)"),        # error 0.096 rounds to 0.1: one decade up
    # two-digit error (output.c:114)
    (2, 1.25, 0.012, " 1.250(12)"),
    (2, 1234.5, 12.0, " 1235(12)"),
    (2, 0.5, 0.5, " 0.50(50)"),
])
def test_nice_output_known_answers(fn, val, err, exp):
    """value(error) strings as in scientific journals (hand-worked)."""
    L = B.lib()
    f = getattr(L, "nice_output_%d" % fn)
    f.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_int]
    buf = C.create_string_buffer(80)
    assert f(buf, val, err, 0) == len(exp)
    assert buf.value.decode() == exp
    f(buf, val, err, 16)  # right-justified
    assert buf.value.decode() == exp.rjust(16)
