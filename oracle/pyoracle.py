"""ctypes binding of the oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker for the HIP implementation; the product package never imports
it.  See oracle/oracle.h for what each function restates.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FP32, FP64 = 0, 1


class Params(C.Structure):
    _fields_ = [("N", C.c_double), ("dt", C.c_double), ("numbins", C.c_longlong),
                ("lobin", C.c_longlong), ("zmax", C.c_int), ("wmax", C.c_int),
                ("wmaxP", C.c_int), ("numharm", C.c_int), ("sigma", C.c_float),
                ("floP", C.c_int), ("fhiP", C.c_int), ("rloP", C.c_int), ("rhiP", C.c_int),
                ("flo", C.c_double), ("fhi", C.c_double), ("rlo", C.c_double),
                ("rhi", C.c_double), ("nph", C.c_double)]


class Obs(C.Structure):
    _fields_ = [("N", C.c_longlong), ("numbins", C.c_longlong), ("lobin", C.c_longlong),
                ("highestbin", C.c_longlong), ("maxkernlen", C.c_int), ("corr_uselen", C.c_int),
                ("fftlen", C.c_int), ("numharmstages", C.c_int), ("numz", C.c_int),
                ("numw", C.c_int), ("dt", C.c_double), ("T", C.c_double), ("rlo", C.c_double),
                ("rhi", C.c_double), ("zlo", C.c_double), ("zhi", C.c_double),
                ("wlo", C.c_double), ("whi", C.c_double), ("sigma", C.c_float),
                ("powcut", C.c_float * 6), ("numindep", C.c_longlong * 6), ("nph", C.c_double),
                ("zmax_cmd", C.c_int), ("wmax_cmd", C.c_int)]


class HarmInfo(C.Structure):
    _fields_ = [("numharm", C.c_int), ("harmnum", C.c_int), ("zmax", C.c_int),
                ("wmax", C.c_int), ("numz", C.c_int), ("numw", C.c_int), ("fftlen", C.c_int),
                ("halfwidth", C.c_int)]


class Cand(C.Structure):
    _fields_ = [("power", C.c_float), ("sigma", C.c_float), ("numharm", C.c_int),
                ("r", C.c_double), ("z", C.c_double), ("w", C.c_double)]


class Hit(C.Structure):
    _fields_ = [("rstep", C.c_longlong), ("stage", C.c_int), ("wi", C.c_int), ("zi", C.c_int),
                ("ri", C.c_int), ("power", C.c_float), ("sigma", C.c_float)]


class RDerivs(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("pow", "phs", "dpow", "dphs", "d2pow", "d2phs", "locpow")]


class FourierProps(C.Structure):
    _fields_ = [("r", C.c_double), ("rerr", C.c_float), ("z", C.c_double), ("zerr", C.c_float),
                ("w", C.c_double), ("werr", C.c_float)] + \
               [(k, C.c_float) for k in ("pow", "powerr", "sig", "rawpow", "phs", "phserr", "cen", "cenerr",
                                         "pur", "purerr", "locpow")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = C.CDLL(path)
        P = C.POINTER
        L.oracle_cfft_f.argtypes = [P(C.c_float), C.c_int, C.c_int]
        L.oracle_cfft_d.argtypes = [P(C.c_double), C.c_int, C.c_int]
        L.oracle_realfft_fwd_f.argtypes = [P(C.c_float), C.c_long]
        L.oracle_realfft_fwd_d.argtypes = [P(C.c_double), C.c_long]
        L.oracle_realfft_inv_f.argtypes = [P(C.c_float), C.c_long]
        L.oracle_realfft_inv_d.argtypes = [P(C.c_double), C.c_long]
        L.oracle_downsamp.argtypes = [P(C.c_float), P(C.c_float), C.c_int, C.c_int, C.c_int]
        L.oracle_dedisp.argtypes = [P(C.c_float), P(C.c_float), P(C.c_float), C.c_int, C.c_int, C.c_int,
                                    P(C.c_int)]
        for fn in ("oracle_next2_to_n", "oracle_next_good_fftlen", "oracle_fftlen_from_kernwidth"):
            getattr(L, fn).argtypes = [C.c_int]
            getattr(L, fn).restype = C.c_int
        L.oracle_z_resp_halfwidth.argtypes = [C.c_double]
        L.oracle_w_resp_halfwidth.argtypes = [C.c_double, C.c_double]
        L.oracle_gen_r_response.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.oracle_gen_z_response.argtypes = [C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_void_p]
        L.oracle_gen_w_response.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double, C.c_int,
                                            C.c_int, C.c_void_p]
        L.oracle_fresnel.argtypes = [C.c_double, P(C.c_double), P(C.c_double)]
        for fn in ("oracle_power_for_sigma", "oracle_candidate_sigma"):
            getattr(L, fn).argtypes = [C.c_double, C.c_int, C.c_double]
            getattr(L, fn).restype = C.c_double
        L.oracle_chi2_logp2.argtypes = [C.c_double, C.c_double]
        L.oracle_chi2_logp2.restype = C.c_double
        L.oracle_equivalent_gaussian_sigma2.argtypes = [C.c_double]
        L.oracle_equivalent_gaussian_sigma2.restype = C.c_double
        L.oracle_normal_sf.argtypes = [C.c_double]
        L.oracle_normal_sf.restype = C.c_double
        L.oracle_median.argtypes = [P(C.c_float), C.c_int]
        L.oracle_median.restype = C.c_float
        L.oracle_create_accelobs.argtypes = [P(Params), P(Obs)]
        L.oracle_harm_info_get.argtypes = [P(Obs), C.c_int, C.c_int, P(HarmInfo)]
        L.oracle_kernel.argtypes = [P(Obs), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.oracle_ctx_create.argtypes = [P(Obs), C.c_int, C.c_int]
        L.oracle_ctx_create.restype = C.c_void_p
        L.oracle_ctx_free.argtypes = [C.c_void_p]
        L.oracle_subharm_fderivs_vol.argtypes = [C.c_void_p, P(C.c_float), C.c_int, C.c_int,
                                                 C.c_double, C.c_double, C.c_void_p,
                                                 P(C.c_uint16), P(C.c_uint16), P(C.c_longlong)]
        L.oracle_search.argtypes = [C.c_void_p, P(C.c_float), C.c_longlong, C.c_longlong,
                                    P(Cand), C.c_longlong, P(Hit), C.c_longlong, P(C.c_longlong),
                                    C.c_longlong, C.c_void_p, P(C.c_int)]
        L.oracle_search.restype = C.c_longlong
        L.oracle_num_rsteps.argtypes = [P(Obs)]
        L.oracle_num_rsteps.restype = C.c_longlong
        L.oracle_insert_cands.argtypes = [P(C.c_double), P(C.c_float), P(C.c_float), P(C.c_int),
                                          P(C.c_double), P(C.c_double), C.c_longlong, P(Cand), C.c_longlong]
        L.oracle_insert_cands.restype = C.c_longlong
        L.oracle_z_resp_halfwidth_acc.argtypes = [C.c_double, C.c_int]
        L.oracle_w_resp_halfwidth_acc.argtypes = [C.c_double, C.c_double, C.c_int]
        L.oracle_rz_interp.argtypes = [P(C.c_float), C.c_int, C.c_double, C.c_double, C.c_int, P(C.c_float)]
        L.oracle_rzw_interp.argtypes = [P(C.c_float), C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                        P(C.c_float)]
        L.oracle_get_localpower3d.argtypes = [P(C.c_float), C.c_int, C.c_double, C.c_double, C.c_double]
        L.oracle_get_localpower3d.restype = C.c_double
        L.oracle_get_derivs3d.argtypes = [P(C.c_float), C.c_int, C.c_double, C.c_double, C.c_double,
                                          C.c_double, P(RDerivs)]
        L.oracle_corr_rzw_vol.argtypes = [P(C.c_float), C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                          C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, P(C.c_int),
                                          P(C.c_float)]
        L.oracle_corr_rzw_vol.restype = C.c_int
        L.oracle_max_rzw_arr_harmonics.argtypes = [P(C.c_float), C.c_int, C.c_int, C.c_double, C.c_double,
                                                   C.c_double, P(C.c_double), P(C.c_double), P(C.c_double),
                                                   P(RDerivs), P(C.c_double)]
        L.oracle_calc_props.argtypes = [P(RDerivs), C.c_double, C.c_double, C.c_double, P(FourierProps)]
        L.oracle_max_rz_arr_harmonics.argtypes = [P(C.c_float), C.c_int, C.c_int, C.c_double, C.c_double,
                                                  P(C.c_double), P(C.c_double), P(RDerivs), P(C.c_double)]
        L.oracle_calc_props.restype = C.c_int
        L.oracle_max_rzw_arr.argtypes = [P(C.c_float), C.c_int, C.c_double, C.c_double, C.c_double,
                                         P(C.c_double), P(C.c_double), P(C.c_double), P(RDerivs)]
        L.oracle_max_rzw_arr.restype = C.c_double
        L.oracle_max_rz_arr.argtypes = [P(C.c_float), C.c_int, C.c_double, C.c_double, P(C.c_double),
                                        P(C.c_double), P(RDerivs)]
        L.oracle_max_rz_arr.restype = C.c_double
        _lib = L
    return _lib


def _c64(data):
    a = np.ascontiguousarray(data)
    if a.dtype == np.complex64:
        a = a.view(np.float32)
    assert a.dtype == np.float32
    return a


def w_resp_halfwidth_acc(z, w, highacc=False):
    return lib().oracle_w_resp_halfwidth_acc(z, w, int(highacc))


def rzw_interp(data, r, z, w, khw):
    """Complex Fourier interpolation at (r, z, w) (rzwinterp.c:145)."""
    a = _c64(data)
    out = np.zeros(2, dtype=np.float32)
    lib().oracle_rzw_interp(fptr(a), len(a) // 2, r, z, w, khw, fptr(out))
    return complex(out[0], out[1])


def localpower3d(data, r, z, w):
    a = _c64(data)
    return lib().oracle_get_localpower3d(fptr(a), len(a) // 2, r, z, w)


def derivs3d(data, r, z, w, locpow):
    a = _c64(data)
    d = RDerivs()
    lib().oracle_get_derivs3d(fptr(a), len(a) // 2, r, z, w, locpow, C.byref(d))
    return d


def corr_rzw_vol(data, numbetween, startbin, zlo, zhi, numz, wlo, whi, numw, fftlen):
    """(volume [numw][numz][numgoodbins] complex64, nextbin) (rzwinterp.c:3)."""
    a = _c64(data)
    out = np.zeros(2 * numw * numz * fftlen, dtype=np.float32)
    nb = C.c_int(0)
    n = lib().oracle_corr_rzw_vol(fptr(a), len(a) // 2, numbetween, startbin, zlo, zhi, numz, wlo, whi, numw,
                                  fftlen, C.byref(nb), fptr(out))
    return out[: 2 * numw * numz * n].view(np.complex64).reshape(numw, numz, n), nb.value


def max_rzw_arr_harmonics(data, numharm, r, z, w):
    """(r, z, w, [RDerivs per harmonic], powers) (maximize_rzw.c:85)."""
    a = _c64(data)
    ro, zo, wo = C.c_double(), C.c_double(), C.c_double()
    ders = (RDerivs * numharm)()
    pows = (C.c_double * numharm)()
    lib().oracle_max_rzw_arr_harmonics(fptr(a), len(a) // 2, numharm, r, z, w, C.byref(ro), C.byref(zo),
                                       C.byref(wo), ders, pows)
    return ro.value, zo.value, wo.value, list(ders), list(pows)


def max_rz_arr_harmonics(data, numharm, r, z):
    """(r, z, [RDerivs per harmonic], powers) of a w = 0 search (maximize_rz.c:141)."""
    a = _c64(data)
    ro, zo = C.c_double(), C.c_double()
    ders = (RDerivs * numharm)()
    pows = (C.c_double * numharm)()
    lib().oracle_max_rz_arr_harmonics(fptr(a), len(a) // 2, numharm, r, z, C.byref(ro), C.byref(zo), ders, pows)
    return ro.value, zo.value, list(ders), list(pows)


def max_rzw_arr(data, r, z, w):
    """(r, z, w, RDerivs, raw power) of one harmonic's 9 x 9 x 24 volume
    (maximize_rzw.c:3)."""
    a = _c64(data)
    ro, zo, wo, d = C.c_double(), C.c_double(), C.c_double(), RDerivs()
    p = lib().oracle_max_rzw_arr(fptr(a), len(a) // 2, r, z, w, C.byref(ro), C.byref(zo), C.byref(wo), C.byref(d))
    return ro.value, zo.value, wo.value, d, p


def max_rz_arr(data, r, z):
    """(r, z, RDerivs, raw power) of one harmonic's simplex (maximize_rz.c:25)."""
    a = _c64(data)
    ro, zo, d = C.c_double(), C.c_double(), RDerivs()
    p = lib().oracle_max_rz_arr(fptr(a), len(a) // 2, r, z, C.byref(ro), C.byref(zo), C.byref(d))
    return ro.value, zo.value, d, p


def calc_props(derivs, r, z, w):
    p = FourierProps()
    if lib().oracle_calc_props(C.byref(derivs), r, z, w, C.byref(p)) != 0:
        raise ValueError("calc_props: power or local power out of bounds")
    return p


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(N, dt, numbins, zmax, wmax, numharm, sigma, wmaxP=True, flo=1.0, fhi=10000.0,
                rlo=None, rhi=None, lobin=0, nph=0.0):
    p = Params()
    p.N, p.dt, p.numbins, p.lobin = float(N), float(dt), int(numbins), int(lobin)
    p.zmax, p.wmax, p.wmaxP, p.numharm, p.sigma = int(zmax), int(wmax), int(bool(wmaxP)), int(numharm), float(sigma)
    if rlo is not None:
        p.floP, p.rloP, p.rlo = 0, 1, float(rlo)
    else:
        p.floP, p.flo = 1, float(flo)
    if rhi is not None:
        p.fhiP, p.rhiP, p.rhi = 0, 1, float(rhi)
    else:
        p.fhiP, p.fhi = 1, float(fhi)
    p.nph = float(nph)
    return p


def create_accelobs(params):
    obs = Obs()
    rc = lib().oracle_create_accelobs(C.byref(params), C.byref(obs))
    if rc != 0:
        raise ValueError("oracle_create_accelobs failed: %d" % rc)
    return obs


def harm_info(obs, numharm, harmnum):
    h = HarmInfo()
    lib().oracle_harm_info_get(C.byref(obs), numharm, harmnum, C.byref(h))
    return h


def harmonics(obs):
    """(numharm, harmnum) in the reference's stage order (accelsearch.c:199-211)."""
    out = [(1, 1)]
    for stage in range(1, obs.numharmstages):
        for harm in range(1, 1 << stage, 2):
            out.append((1 << stage, harm))
    return out


def kernel(obs, numharm, harmnum, zi, wi, mode=FP32):
    h = harm_info(obs, numharm, harmnum)
    out = np.zeros(2 * h.fftlen, dtype=np.float64 if mode == FP64 else np.float32)
    lib().oracle_kernel(C.byref(obs), numharm, harmnum, zi, wi, mode, out.ctypes.data)
    return out[0::2] + 1j * out[1::2]


class Ctx:
    def __init__(self, obs, mode=FP32, nthreads=1):
        self.obs, self.mode = obs, mode
        self.h = lib().oracle_ctx_create(C.byref(obs), mode, nthreads)

    def close(self):
        if self.h:
            lib().oracle_ctx_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def subharm(self, fft, numharm, harmnum, fullrlo, fullrhi):
        """subharm_fderivs_vol: returns (powers[w][z][r], rinds, zinds, meta)."""
        obs = self.obs
        h = harm_info(obs, numharm, harmnum)
        cap = h.numw * h.numz * (obs.corr_uselen + 8)
        pows = np.zeros(cap, dtype=np.float64 if self.mode == FP64 else np.float32)
        rinds = np.zeros(obs.corr_uselen, dtype=np.uint16)
        zinds = np.zeros(max(obs.numz, 1), dtype=np.uint16)
        meta = np.zeros(5, dtype=np.int64)
        numrs = lib().oracle_subharm_fderivs_vol(
            self.h, fptr(fft), numharm, harmnum, fullrlo, fullrhi, pows.ctypes.data,
            rinds.ctypes.data_as(C.POINTER(C.c_uint16)), zinds.ctypes.data_as(C.POINTER(C.c_uint16)),
            meta.ctypes.data_as(C.POINTER(C.c_longlong)))
        pows = pows[: h.numw * h.numz * numrs].reshape(h.numw, h.numz, numrs)
        return pows, rinds, zinds, meta

    def search(self, fft, rstep_lo=0, rstep_hi=-1, maxcands=1 << 16, maxhits=1 << 20, sum_rstep=-1,
               hits_array=False):
        """oracle_search over rsteps [rstep_lo, rstep_hi): (candidates, hits,
        summed plane of sum_rstep).  hits: a list of (rstep, stage, wi, zi, ri,
        power, sigma) tuples, or with hits_array a structured numpy array of
        those fields.  More than maxhits hits raise an error."""
        obs = self.obs
        cands = (Cand * maxcands)()
        hits = (Hit * maxhits)()
        nh = C.c_longlong(0)
        h = harm_info(obs, 1, 1)
        sums = None
        snumrs = C.c_int(0)
        if sum_rstep >= 0:
            sums = np.zeros(h.numw * h.numz * obs.corr_uselen,
                            dtype=np.float64 if self.mode == FP64 else np.float32)
        n = lib().oracle_search(self.h, fptr(fft), rstep_lo, rstep_hi, cands, maxcands, hits, maxhits,
                                C.byref(nh), sum_rstep, sums.ctypes.data if sums is not None else None,
                                C.byref(snumrs))
        clist = [(cands[i].power, cands[i].sigma, cands[i].numharm, cands[i].r, cands[i].z, cands[i].w)
                 for i in range(min(n, maxcands))]
        if hits_array:
            if nh.value > maxhits:
                raise RuntimeError("oracle search: %d hits, above maxhits %d" % (nh.value, maxhits))
            dt = np.dtype([(name, np.dtype(ct)) for name, ct in Hit._fields_], align=True)
            assert dt.itemsize == C.sizeof(Hit)
            hlist = np.frombuffer(hits, dtype=dt, count=nh.value).copy()
        else:
            hlist = [(hits[i].rstep, hits[i].stage, hits[i].wi, hits[i].zi, hits[i].ri, hits[i].power,
                      hits[i].sigma) for i in range(min(nh.value, maxhits))]
        if sums is not None:
            sums = sums.reshape(h.numw, h.numz, obs.corr_uselen)
        return clist, hlist, sums


def num_rsteps(obs):
    return lib().oracle_num_rsteps(C.byref(obs))


def power_for_sigma(sigma, numsum, numtrials):
    return lib().oracle_power_for_sigma(sigma, numsum, numtrials)


def candidate_sigma(power, numsum, numtrials):
    return lib().oracle_candidate_sigma(power, numsum, numtrials)


def fresnel(x):
    s, c = C.c_double(), C.c_double()
    lib().oracle_fresnel(x, C.byref(s), C.byref(c))
    return s.value, c.value


def median(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return lib().oracle_median(fptr(a), len(a))


def realfft(x, isign, mode=FP32):
    """realfft (fastffts.c:163-235) of n floats in place semantics: returns the
    transformed copy (isign -1: NR-packed spectrum as n floats; +1: inverse,
    scaled by 2/n)."""
    dt = np.float64 if mode == FP64 else np.float32
    a = np.array(x, dtype=dt, copy=True)
    fn = {(-1, FP32): "oracle_realfft_fwd_f", (-1, FP64): "oracle_realfft_fwd_d",
          (1, FP32): "oracle_realfft_inv_f", (1, FP64): "oracle_realfft_inv_d"}[(isign, mode)]
    getattr(lib(), fn)(dptr(a) if mode == FP64 else fptr(a), len(a))
    return a


def downsamp(data, bs):
    """DownsampDataArray (prepsubband_kernels.cu:26-46): data [worklen*bs][nsub]
    float32 -> [worklen][nsub]."""
    a = np.ascontiguousarray(data, dtype=np.float32)
    worklen, nsub = a.shape[0] // bs, a.shape[1]
    out = np.empty((worklen, nsub), dtype=np.float32)
    lib().oracle_downsamp(fptr(out), fptr(a), worklen, bs, nsub)
    return out


def dedisp(cur, last, offsets):
    """getDedisp (prepsubband_kernels.cu:48-77): cur, last [worklen][nsub],
    offsets [numdms][nsub] -> [numdms][worklen]."""
    c = np.ascontiguousarray(cur, dtype=np.float32)
    l = np.ascontiguousarray(last, dtype=np.float32)
    o = np.ascontiguousarray(offsets, dtype=np.int32)
    worklen, nsub = c.shape
    out = np.empty((o.shape[0], worklen), dtype=np.float32)
    lib().oracle_dedisp(fptr(c), fptr(l), fptr(out), worklen, nsub, o.shape[0], o.ctypes.data_as(C.POINTER(C.c_int)))
    return out


def cfft(x, sign, mode=FP32):
    if mode == FP64:
        a = np.empty(2 * len(x), dtype=np.float64)
        a[0::2], a[1::2] = x.real, x.imag
        lib().oracle_cfft_d(dptr(a), len(x), sign)
    else:
        a = np.empty(2 * len(x), dtype=np.float32)
        a[0::2], a[1::2] = x.real, x.imag
        lib().oracle_cfft_f(fptr(a), len(x), sign)
    return a[0::2] + 1j * a[1::2]


def response(kind, *args, mode=FP32):
    """kind in r/z/w; args as the reference: r:(roffset, numbetween, numkern),
    z:(roffset, numbetween, z, numkern), w:(roffset, numbetween, z, w, numkern)."""
    numkern = args[-1]
    out = np.zeros(2 * numkern, dtype=np.float64 if mode == FP64 else np.float32)
    L = lib()
    if kind == "r":
        L.oracle_gen_r_response(args[0], args[1], numkern, mode, out.ctypes.data)
    elif kind == "z":
        L.oracle_gen_z_response(args[0], args[1], args[2], numkern, mode, out.ctypes.data)
    else:
        L.oracle_gen_w_response(args[0], args[1], args[2], args[3], numkern, mode, out.ctypes.data)
    return out[0::2] + 1j * out[1::2]


def w_resp_halfwidth(z, w):
    return lib().oracle_w_resp_halfwidth(z, w)


def z_resp_halfwidth(z):
    return lib().oracle_z_resp_halfwidth(z)


def insert_cands(rr, sigma, power, numharm, z, w):
    """insert_new_accelcand over hits in order -> [(power, sigma, numharm, r, z, w)]."""
    n = len(rr)
    rr = np.ascontiguousarray(rr, dtype=np.float64)
    sigma = np.ascontiguousarray(sigma, dtype=np.float32)
    power = np.ascontiguousarray(power, dtype=np.float32)
    numharm = np.ascontiguousarray(numharm, dtype=np.int32)
    z = np.ascontiguousarray(z, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    cap = max(n, 1)
    out = (Cand * cap)()
    k = lib().oracle_insert_cands(dptr(rr), fptr(sigma), fptr(power), numharm.ctypes.data_as(C.POINTER(C.c_int)),
                                  dptr(z), dptr(w), n, out, cap)
    return [(out[i].power, out[i].sigma, out[i].numharm, out[i].r, out[i].z, out[i].w) for i in range(k)]


def sort_cands(cands):
    """sort_accelcands (accel_utils.c:361-380): decreasing sigma, ties by
    decreasing power; g_slist_sort is a stable merge sort."""
    return sorted(cands, key=lambda c: (-c[1], -c[0]))


# eliminate_harmonics (accel_utils.c:482-571): the ratios p/q, besides the
# integer multiples and fractions 1..16, for which a later candidate counts as
# harmonically related to an earlier one (|r_earlier * p / q - r_later| < 1.5).
HARMONIC_RATIOS = ((3, 2), (5, 2), (2, 3), (4, 3), (5, 3), (3, 4), (5, 4), (2, 5), (3, 5), (4, 5), (5, 6),
                   (2, 7), (3, 7), (4, 7), (3, 8), (5, 8), (2, 9), (3, 10), (2, 11), (3, 11), (2, 13),
                   (3, 13), (2, 15))


def harmonically_related(r_cur, r_other, tooclose=1.5, maxharm=16):
    for ii in range(1, maxharm + 1):
        if abs(r_cur / ii - r_other) < tooclose:
            return True
    for ii in range(1, maxharm + 1):
        if abs(r_cur * ii - r_other) < tooclose:
            return True
    return any(abs(r_cur * p / q - r_other) < tooclose for p, q in HARMONIC_RATIOS)


def eliminate_harmonics(cands):
    """eliminate_harmonics (accel_utils.c:482-571) over a sorted list: for each
    surviving candidate in order, drop every later one related to it."""
    out = list(cands)
    i = 0
    while i < len(out) - 1:
        cur = out[i][3]
        out = out[: i + 1] + [c for c in out[i + 1:] if not harmonically_related(cur, c[3])]
        i += 1
    return out
