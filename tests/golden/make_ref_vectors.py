"""Writes ref_vectors.json: outputs of the reference's own C sources
(dcdflib.c, ipmpar.c, fresnl.c, median.c compiled unmodified into
oracle/_ref/libpresto_ref.so by `make -C oracle ref`) on fixed inputs.

    python tests/golden/make_ref_vectors.py

The values pin the oracle's Fresnel integrals, lower median,
power_for_sigma (characteristics.c:670-705) and candidate_sigma
(characteristics.c:647-668, exact branch) where the reference tree and
oracle/_ref are absent.  power_for_sigma and candidate_sigma are composed
from the dcdflib calls exactly as characteristics.c makes them.
"""
import ctypes as C
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpresto_ref.so")


def lib():
    L = C.CDLL(REF_SO)
    D, I = C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.fresnl.argtypes = [C.c_double, D, D]
    L.median.argtypes = [C.POINTER(C.c_float), C.c_int]
    L.median.restype = C.c_float
    for fn in ("cdfchi", "cdfchi2"):
        getattr(L, fn).argtypes = [I, D, D, D, D, I, D]
    for fn in ("cdfnor", "cdfnor2"):
        getattr(L, fn).argtypes = [I, D, D, D, D, D, I, D]
    return L


def power_for_sigma(L, sigma, numsum, numtrials):
    which, status = C.c_int(1), C.c_int(0)
    p, q, x, mean, sd, bound = (C.c_double(v) for v in (0, 0, sigma, 0, 1, 0))
    L.cdfnor(C.byref(which), C.byref(p), C.byref(q), C.byref(x), C.byref(mean), C.byref(sd), C.byref(status),
             C.byref(bound))
    q = C.c_double(q.value / numtrials)
    p = C.c_double(1.0 - q.value)
    which, df, status = C.c_int(2), C.c_double(2.0 * numsum), C.c_int(0)
    L.cdfchi(C.byref(which), C.byref(p), C.byref(q), C.byref(x), C.byref(df), C.byref(status), C.byref(bound))
    assert status.value == 0
    return 0.5 * x.value


def candidate_sigma(L, power, numsum, numtrials):
    chi2, dof = 2.0 * power, 2.0 * numsum
    if chi2 / dof > 15.0 or (dof > 150 and chi2 / dof > 6.0):
        return None
    which, status = C.c_int(1), C.c_int(0)
    p, q, x, df, bound = (C.c_double(v) for v in (0, 0, chi2, dof, 0))
    L.cdfchi2(C.byref(which), C.byref(p), C.byref(q), C.byref(x), C.byref(df), C.byref(status), C.byref(bound))
    logp = math.log(q.value) + math.log(numtrials)
    if logp < -600 or logp > math.log(0.5):
        return None
    q2 = math.exp(logp)
    which, p2, q2c, x2, mean, sd = C.c_int(2), C.c_double(1.0 - q2), C.c_double(q2), C.c_double(0), \
        C.c_double(0), C.c_double(1)
    L.cdfnor2(C.byref(which), C.byref(p2), C.byref(q2c), C.byref(x2), C.byref(mean), C.byref(sd), C.byref(status),
              C.byref(bound))
    return max(x2.value, 0.0)


def main():
    L = lib()
    out = {"source": "reference dcdflib.c, ipmpar.c, fresnl.c, median.c via oracle/_ref"}
    xs = [-7.5, -2.0, -0.3, 0.0, 0.1, 0.7, 1.0, 1.6, 2.5, 3.3, 5.0, 8.0, 12.5, 20.0, 47.0, 100.0]
    fr = []
    for x in xs:
        s, c = C.c_double(), C.c_double()
        L.fresnl(x, C.byref(s), C.byref(c))
        fr.append([x, s.value, c.value])
    out["fresnl"] = fr
    med = []
    state = 12345
    for n in (1, 2, 5, 8, 17, 64):
        a = []
        for _ in range(n):
            state = (state * 1103515245 + 12345) % (1 << 31)
            a.append(float(C.c_float(state / float(1 << 31) * 10.0).value))
        arr = (C.c_float * n)(*a)
        med.append([a, float(L.median(arr, n))])
    out["median"] = med
    pfs = []
    for numsum in (1, 2, 4, 8, 16, 32):
        for sigma, ntrials in ((2.0, 1.2e7), (5.0, 1.4e10), (3.0, 5e8), (6.5, 2.0e9)):
            pfs.append([sigma, numsum, ntrials, power_for_sigma(L, sigma, numsum, ntrials)])
    out["power_for_sigma"] = pfs
    cs = []
    for numsum in (1, 2, 4, 8, 16):
        for k in range(25):
            power = numsum * (2.0 + 13.0 * k / 24.0)
            v = candidate_sigma(L, power, numsum, 1.4e10)
            if v is not None:
                cs.append([power, numsum, 1.4e10, v])
    out["candidate_sigma"] = cs
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote %d fresnl, %d median, %d power_for_sigma, %d candidate_sigma vectors" %
          (len(fr), len(med), len(pfs), len(cs)))


if __name__ == "__main__":
    main()
