"""ctypes binding of the C ABI in include/pzl_accel.h (libaccelsearch_lib.so).

The library is built in-tree by prestozl_amd.build (hipcc, gfx950).  Loading
fails loudly when it is missing: there is no CPU fallback for the search.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libaccelsearch_lib.so")

PZL_ERR_OVERFLOW = 1
PZL_ERR_TIMEOUT = -5


class Config(C.Structure):
    _fields_ = [("N", C.c_double), ("dt", C.c_double), ("numbins", C.c_longlong),
                ("lobin", C.c_longlong), ("zmax", C.c_int), ("wmax", C.c_int),
                ("wmaxP", C.c_int), ("numharm", C.c_int), ("sigma", C.c_float),
                ("floP", C.c_int), ("fhiP", C.c_int), ("rloP", C.c_int), ("rhiP", C.c_int),
                ("flo", C.c_double), ("fhi", C.c_double), ("rlo", C.c_double), ("rhi", C.c_double),
                ("nph", C.c_double), ("batchsize", C.c_int), ("maxhits", C.c_longlong)]


class Geometry(C.Structure):
    _fields_ = [("N", C.c_longlong), ("numbins", C.c_longlong), ("lobin", C.c_longlong),
                ("highestbin", C.c_longlong), ("nrsteps", C.c_longlong),
                ("maxkernlen", C.c_int), ("corr_uselen", C.c_int), ("fftlen", C.c_int),
                ("numharmstages", C.c_int), ("numz", C.c_int), ("numw", C.c_int),
                ("numharm_kernels", C.c_int),
                ("dt", C.c_double), ("T", C.c_double), ("rlo", C.c_double), ("rhi", C.c_double),
                ("zlo", C.c_double), ("zhi", C.c_double), ("wlo", C.c_double), ("whi", C.c_double),
                ("sigma", C.c_float), ("powcut", C.c_float * 6), ("numindep", C.c_longlong * 6),
                ("bank_bytes", C.c_longlong), ("plane_bytes_per_rstep", C.c_longlong),
                ("device_bytes", C.c_longlong)]


class HarmInfo(C.Structure):
    _fields_ = [("numharm", C.c_int), ("harmnum", C.c_int), ("stage", C.c_int), ("zmax", C.c_int),
                ("wmax", C.c_int), ("numz", C.c_int), ("numw", C.c_int), ("fftlen", C.c_int),
                ("halfwidth", C.c_int), ("max_numrs", C.c_int)]


class Cand(C.Structure):
    _fields_ = [("power", C.c_float), ("sigma", C.c_float), ("numharm", C.c_int),
                ("r", C.c_double), ("z", C.c_double), ("w", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("ms_wall", C.c_double), ("ms_gpu", C.c_double), ("ms_segments", C.c_double),
                ("ms_corr", C.c_double), ("ms_sum", C.c_double), ("ms_corr_max_class", C.c_double),
                ("ms_host", C.c_double), ("corr_launches", C.c_longlong), ("nhits", C.c_longlong),
                ("nrsteps", C.c_longlong), ("cells", C.c_double), ("bytes_alg", C.c_double),
                ("bytes_alg_corr", C.c_double), ("nclass", C.c_int), ("class_fftlen", C.c_int * 16),
                ("class_ms", C.c_double * 16), ("class_launches", C.c_longlong * 16),
                ("class_bytes", C.c_double * 16), ("ms_graph", C.c_double), ("graph_launches", C.c_longlong),
                ("class_bank_bytes", C.c_double * 16), ("class_flops", C.c_double * 16),
                ("ms_corr_group", C.c_double), ("group_launches", C.c_longlong),
                ("ms_wait_count", C.c_double), ("ms_wait_pool", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("class_")}
        d["classes"] = [dict(fftlen=self.class_fftlen[i], ms=self.class_ms[i], launches=self.class_launches[i],
                             bytes=self.class_bytes[i], bank_bytes=self.class_bank_bytes[i],
                             flops=self.class_flops[i]) for i in range(self.nclass)]
        return d


MAX_NUMHARM = 32


class RDerivs(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("pow", "phs", "dpow", "dphs", "d2pow", "d2phs", "locpow")]


class OptCand(C.Structure):
    _fields_ = [("r", C.c_double), ("z", C.c_double), ("w", C.c_double), ("numharm", C.c_int),
                ("fdot_only", C.c_int), ("ropt", C.c_double), ("zopt", C.c_double), ("wopt", C.c_double),
                ("derivs", RDerivs * MAX_NUMHARM), ("pows", C.c_double * MAX_NUMHARM)]


class RequestStats(C.Structure):
    _fields_ = [("ms_queue", C.c_double), ("ms_run", C.c_double), ("ms_total", C.c_double),
                ("t_submit_ms", C.c_double), ("t_admit_ms", C.c_double), ("t_end_ms", C.c_double),
                ("footprint", C.c_longlong), ("nhits", C.c_longlong), ("cells", C.c_double),
                ("plan_reused", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [("budget", C.c_longlong), ("admitted_bytes", C.c_longlong), ("peak_admitted_bytes", C.c_longlong),
                ("resident_bytes", C.c_longlong), ("queued", C.c_longlong), ("completed", C.c_longlong),
                ("pooled_plans", C.c_longlong), ("inflight", C.c_int), ("peak_inflight", C.c_int),
                ("peak_resident_bytes", C.c_longlong), ("timed_out", C.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


class PzlError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__("pzl error %d: %s" % (rc, msg))
        self.rc = rc


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libaccelsearch_lib.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIBPATH)
        P = C.POINTER
        L.pzl_last_error.restype = C.c_char_p
        L.pzl_device_count.argtypes = [P(C.c_int)]
        L.pzl_device_memory.argtypes = [C.c_int, P(C.c_longlong), P(C.c_longlong)]
        L.pzl_realfft.argtypes = [P(C.c_float), C.c_longlong, C.c_int]
        L.pzl_realfft_device.argtypes = [C.c_int, C.c_void_p, C.c_longlong, C.c_int]
        L.pzl_realfft_device_batch.argtypes = [C.c_int, C.c_void_p, C.c_longlong, C.c_int, C.c_int]
        L.pzl_realfft_device_noalloc.argtypes = [C.c_int, C.c_void_p, C.c_longlong, C.c_int, C.c_int, C.c_void_p]
        L.pzl_downsamp_data_array.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.pzl_get_dedisp.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.pzl_get_dedisp_strided.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                             C.c_void_p, C.c_longlong]
        L.pzl_dedisperse.argtypes = [C.c_void_p, C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.pzl_geometry_from_config.argtypes = [P(Config), P(Geometry)]
        L.pzl_plan_create.argtypes = [P(Config), C.c_int, P(C.c_void_p)]
        L.pzl_plan_destroy.argtypes = [C.c_void_p]
        L.pzl_plan_destroy.restype = None
        L.pzl_plan_geometry.argtypes = [C.c_void_p, P(Geometry)]
        L.pzl_plan_device_bytes.argtypes = [C.c_void_p, P(C.c_longlong)]
        L.pzl_plan_harm.argtypes = [C.c_void_p, C.c_int, P(HarmInfo)]
        L.pzl_set_spectrum.argtypes = [C.c_void_p, P(C.c_float), C.c_longlong]
        L.pzl_set_spectrum_device.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong]
        L.pzl_search.argtypes = [C.c_void_p, C.c_longlong, C.c_longlong, P(Cand), C.c_longlong,
                                 P(C.c_longlong), P(Stats)]
        L.pzl_get_hits.argtypes = [C.c_void_p, P(C.c_longlong), P(C.c_float), C.c_longlong, P(C.c_longlong)]
        L.pzl_get_kernel.argtypes = [C.c_void_p, C.c_int, C.c_int, P(C.c_float)]
        L.pzl_compute_plane.argtypes = [C.c_void_p, C.c_longlong, C.c_int, P(C.c_float), C.c_longlong,
                                        P(C.c_int), P(C.c_longlong)]
        L.pzl_compute_sums.argtypes = [C.c_void_p, C.c_longlong, P(C.c_float), C.c_longlong]
        L.pzl_autotune.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int]
        L.pzl_server_create.argtypes = [C.c_int, C.c_longlong, C.c_int, P(C.c_void_p)]
        L.pzl_server_submit.argtypes = [C.c_void_p, P(Config), P(C.c_float), C.c_longlong, P(C.c_longlong)]
        L.pzl_server_submit_device.argtypes = [C.c_void_p, P(Config), C.c_void_p, C.c_longlong, P(C.c_longlong)]
        L.pzl_server_wait.argtypes = [C.c_void_p, C.c_longlong, P(Cand), C.c_longlong, P(C.c_longlong),
                                      P(RequestStats)]
        L.pzl_server_poll.argtypes = [C.c_void_p, P(C.c_longlong), C.c_int, P(C.c_int)]
        L.pzl_server_stats.argtypes = [C.c_void_p, P(ServerStats)]
        L.pzl_server_submit_deadline.argtypes = [C.c_void_p, P(Config), C.c_void_p, C.c_int, C.c_longlong,
                                                 C.c_double, P(C.c_longlong)]
        L.pzl_plan_set_deadline.argtypes = [C.c_void_p, C.c_double]
        L.pzl_plan_set_keep_hits.argtypes = [C.c_void_p, C.c_longlong]
        L.pzl_server_destroy.argtypes = [C.c_void_p]
        L.pzl_server_destroy.restype = None
        L.pzl_plan_variant.argtypes = [C.c_void_p, C.c_int, P(C.c_int)]
        L.pzl_plan_set_graph.argtypes = [C.c_void_p, C.c_int]
        L.pzl_plan_set_overlap.argtypes = [C.c_void_p, C.c_int]
        L.pzl_plan_set_cu_split.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.pzl_plan_set_hit_chunk.argtypes = [C.c_void_p, C.c_longlong]
        L.pzl_plan_cu_split.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.pzl_tune_batchsize.argtypes = [P(Config), C.c_int, P(C.c_float), C.c_longlong, P(C.c_int), C.c_int,
                                         C.c_longlong, P(C.c_int), P(C.c_double)]
        L.pzl_plan_set_variant.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.pzl_plan_corr_group.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int)]
        L.pzl_plan_set_corr_group.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.pzl_corr_group_name.argtypes = [C.c_int]
        L.pzl_corr_group_name.restype = C.c_char_p
        L.pzl_plan_sum_form.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int)]
        L.pzl_optimize.argtypes = [C.c_void_p, P(OptCand), C.c_longlong]
        L.pzl_optimize_spectrum.argtypes = [C.c_int, P(C.c_float), C.c_longlong, P(OptCand), C.c_longlong]
        L.pzl_maximize.argtypes = [C.c_void_p, P(OptCand), C.c_longlong]
        L.pzl_maximize_spectrum.argtypes = [C.c_int, P(C.c_float), C.c_longlong, P(OptCand), C.c_longlong]
        L.pzl_plan_set_sum_form.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.pzl_corr_num_variants.argtypes = [C.c_int]
        L.pzl_corr_variant_name.argtypes = [C.c_int, C.c_int]
        L.pzl_corr_variant_name.restype = C.c_char_p
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise PzlError(rc, lib().pzl_last_error().decode())
    return rc


def make_config(N, dt, numbins, zmax, wmax, numharm, sigma, wmaxP=True, flo=1.0, fhi=10000.0, rlo=None,
                rhi=None, lobin=0, nph=0.0, batchsize=8, maxhits=0):
    c = Config()
    c.N, c.dt, c.numbins, c.lobin = float(N), float(dt), int(numbins), int(lobin)
    c.zmax, c.wmax, c.wmaxP, c.numharm, c.sigma = int(zmax), int(wmax), int(bool(wmaxP)), int(numharm), float(sigma)
    if rlo is not None:
        c.floP, c.rloP, c.rlo = 0, 1, float(rlo)
    else:
        c.floP, c.flo = 1, float(flo)
    if rhi is not None:
        c.fhiP, c.rhiP, c.rhi = 0, 1, float(rhi)
    else:
        c.fhiP, c.fhi = 1, float(fhi)
    c.nph, c.batchsize, c.maxhits = float(nph), int(batchsize), int(maxhits)
    return c


RUNTIME_BYTES_PER_SEARCH = 32 << 20  # PZL_RUNTIME_BYTES_PER_SEARCH (pzl_accel.h)


def realfft(x, isign=-1):
    """realfft (fastffts.c:163-235) of n float32 values on the GPU: returns the
    transformed copy (isign -1: the NR-packed spectrum as n floats, view it as
    complex64 for the n/2 bins; +1: the inverse, scaled by 2/n)."""
    a = np.array(x, dtype=np.float32, copy=True)
    check(lib().pzl_realfft(a.ctypes.data_as(C.POINTER(C.c_float)), len(a), int(isign)))
    return a


class DeviceArray:
    """A device allocation of the library's HIP runtime (test and bench
    plumbing for the device-pointer entry points)."""

    def __init__(self, nbytes):
        self._hip = C.CDLL("libamdhip64.so")
        lib()
        self.nbytes = int(nbytes)
        self._p = C.c_void_p()
        if self._hip.hipMalloc(C.byref(self._p), C.c_size_t(max(self.nbytes, 1))) != 0:
            raise PzlError(-2, "hipMalloc of %d bytes failed" % self.nbytes)

    @property
    def ptr(self):
        return self._p.value

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        if self._hip.hipMemcpy(C.c_void_p(self.ptr + offset), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1):
            raise PzlError(-2, "hipMemcpy to the device failed")

    def download(self, shape, dtype, offset=0):
        out = np.empty(shape, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        if self._hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset), C.c_size_t(out.nbytes), 2):
            raise PzlError(-2, "hipMemcpy from the device failed")
        return out

    def free(self):
        if self._p.value:
            self._hip.hipFree(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def downsamp_data_array(d_out, d_in, worklen, bs, nsub):
    check(lib().pzl_downsamp_data_array(C.c_void_p(d_out), C.c_void_p(d_in), worklen, bs, nsub))


def corr_groups():
    """Names of the correlation groups (grouped launches of several FFT lengths)."""
    L = lib()
    return [L.pzl_corr_group_name(g).decode() for g in range(L.pzl_corr_num_groups())]


def get_dedisp(d_cur, d_last, d_out, worklen, nsub, numdms, d_offsets, out_stride=None):
    if out_stride is None:
        check(lib().pzl_get_dedisp(C.c_void_p(d_cur), C.c_void_p(d_last), C.c_void_p(d_out), worklen, nsub, numdms,
                                   C.c_void_p(d_offsets)))
    else:
        check(lib().pzl_get_dedisp_strided(C.c_void_p(d_cur), C.c_void_p(d_last), C.c_void_p(d_out), worklen, nsub,
                                           numdms, C.c_void_p(d_offsets), int(out_stride)))


def dedisperse(d_subbands, nsamp, nsub, worklen, numdms, d_offsets, d_out):
    check(lib().pzl_dedisperse(C.c_void_p(d_subbands), int(nsamp), nsub, worklen, numdms, C.c_void_p(d_offsets),
                               C.c_void_p(d_out)))


def realfft_device(ptr, n, isign=-1, device=0, nseries=1):
    """The same in place on device memory at 'ptr': nseries series of n floats
    each, back to back."""
    check(lib().pzl_realfft_device_batch(device, C.c_void_p(ptr), int(n), int(nseries), int(isign)))


def realfft_device_noalloc(ptr, n, isign=-1, device=0, nseries=1, stream=None):
    """pzl_realfft_device_noalloc: the transforms of realfft_device with the
    library's fixed workspace (n * nseries <= 2^22), enqueued on 'stream' (a
    hipStream_t as an int, None: the null stream) without waiting."""
    check(lib().pzl_realfft_device_noalloc(device, C.c_void_p(ptr), int(n), int(nseries), int(isign),
                                           C.c_void_p(stream)))


def device_memory(device=0):
    """(free, total) bytes of the device's memory."""
    f, t = C.c_longlong(), C.c_longlong()
    check(lib().pzl_device_memory(device, C.byref(f), C.byref(t)))
    return f.value, t.value


def bank_cache_clear():
    """Free every kernel bank no plan holds."""
    check(lib().pzl_bank_cache_clear())


def geometry(cfg):
    g = Geometry()
    check(lib().pzl_geometry_from_config(C.byref(cfg), C.byref(g)))
    return g


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Plan:
    """A search plan: geometry + the kernel bank resident on one GPU."""

    def __init__(self, cfg, device=0):
        self.cfg = cfg
        self.h = C.c_void_p()
        check(lib().pzl_plan_create(C.byref(cfg), device, C.byref(self.h)))
        self.geom = Geometry()
        check(lib().pzl_plan_geometry(self.h, C.byref(self.geom)))
        self.harms = []
        i = 0
        while True:
            hi = HarmInfo()
            if lib().pzl_plan_harm(self.h, i, C.byref(hi)) != 0:
                break
            self.harms.append(hi)
            i += 1

    def close(self):
        if self.h:
            lib().pzl_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_bytes(self):
        """Device memory the plan holds (its allocations and its kernel bank)."""
        v = C.c_longlong()
        check(lib().pzl_plan_device_bytes(self.h, C.byref(v)))
        return v.value

    def geometry(self):
        """The plan's geometry; device_bytes carries its actual sort scratch."""
        g = Geometry()
        check(lib().pzl_plan_geometry(self.h, C.byref(g)))
        return g

    def set_spectrum(self, fft):
        """fft: complex64 array (numbins) or float32 interleaved."""
        a = np.ascontiguousarray(fft)
        if a.dtype == np.complex64:
            a = a.view(np.float32)
        assert a.dtype == np.float32
        check(lib().pzl_set_spectrum(self.h, _fptr(a), len(a) // 2))

    def autotune(self, reps=3):
        """Time every correlation variant per FFT length on this plan's data and
        keep the fastest; returns the report lines."""
        buf = C.create_string_buffer(1 << 16)
        check(lib().pzl_autotune(self.h, int(reps), buf, len(buf)))
        return buf.value.decode().splitlines()

    def set_deadline(self, ms):
        """Searches stop launching batches 'ms' milliseconds from now and fail
        with PZL_ERR_TIMEOUT (ms <= 0: no deadline)."""
        check(lib().pzl_plan_set_deadline(self.h, float(ms)))

    def set_graph(self, on=True):
        check(lib().pzl_plan_set_graph(self.h, int(bool(on))))

    def set_cu_split(self, sum_cus, sum_only_cus=0):
        """Harmonic sums on the first sum_cus compute units; segments and
        correlations on all but the first sum_only_cus (0: every kernel on
        every compute unit)."""
        check(lib().pzl_plan_set_cu_split(self.h, int(sum_cus), int(sum_only_cus)))

    def cu_split(self):
        """(sum_cus, sum_only_cus)."""
        a, b = C.c_int(0), C.c_int(0)
        check(lib().pzl_plan_cu_split(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_overlap(self, on=2):
        """0 / False: one stream; 1: harmonic sums on their own stream,
        segments and correlations in one; 2 / True (the default): also the
        next batch's segments ahead on a third stream, with the shorter FFT
        lengths' correlations beside the longest's."""
        mode = (2 if on else 0) if isinstance(on, bool) else int(on)
        check(lib().pzl_plan_set_overlap(self.h, mode))

    def sum_form(self):
        """(workgroup order, r per thread) of the harmonic sum."""
        o, r = C.c_int(0), C.c_int(0)
        check(lib().pzl_plan_sum_form(self.h, C.byref(o), C.byref(r)))
        return o.value, r.value

    def set_sum_form(self, order, rpt):
        check(lib().pzl_plan_set_sum_form(self.h, int(order), int(rpt)))

    def optimize(self, cands, fdot_only=False):
        """optimize_accelcand of (r, z, w, numharm) candidates with the plan's
        resident spectrum (r in bins of it)."""
        arr = _opt_array(cands, fdot_only)
        check(lib().pzl_optimize(self.h, arr, len(cands)))
        return _opt_result(arr, len(cands))

    def maximize(self, points, fdot_only=False):
        """-noharmpolish: one harmonic per (r, z, w) point with the plan's
        resident spectrum (max_rzw_arr / max_rz_arr)."""
        arr = _opt_array([(r, z, w, 1) for r, z, w in points], fdot_only)
        check(lib().pzl_maximize(self.h, arr, len(points)))
        return _opt_result(arr, len(points))

    def fftlens(self):
        """The FFT lengths of the plan's harmonics (one kernel variant each)."""
        return sorted({h.fftlen for h in self.harms})

    def variant(self, fftlen):
        v = C.c_int(0)
        check(lib().pzl_plan_variant(self.h, int(fftlen), C.byref(v)))
        return v.value

    def set_variant(self, fftlen, v):
        check(lib().pzl_plan_set_variant(self.h, int(fftlen), int(v)))

    def corr_group(self):
        """(group, all): the correlation group of the shorter FFT lengths (all
        False) or of every length (all True); group -1: one launch per length."""
        g, a = C.c_int(0), C.c_int(0)
        check(lib().pzl_plan_corr_group(self.h, C.byref(g), C.byref(a)))
        return g.value, bool(a.value)

    def set_corr_group(self, group, all=False):
        check(lib().pzl_plan_set_corr_group(self.h, int(group), int(bool(all))))

    def set_spectrum_device(self, ptr, numbins):
        check(lib().pzl_set_spectrum_device(self.h, C.c_void_p(ptr), numbins))

    def search(self, rstep_lo=0, rstep_hi=-1, maxcands=1 << 16, stats=True):
        cands = (Cand * maxcands)()
        n = C.c_longlong(0)
        st = Stats() if stats else None
        check(lib().pzl_search(self.h, rstep_lo, rstep_hi, cands, maxcands, C.byref(n),
                               C.byref(st) if st is not None else None))
        out = [(cands[i].power, cands[i].sigma, cands[i].numharm, cands[i].r, cands[i].z, cands[i].w)
               for i in range(min(n.value, maxcands))]
        return out, (st.as_dict() if st is not None else None)

    def set_hit_chunk(self, records):
        """Records per pinned buffer on their way to the host (default 2^20)."""
        check(lib().pzl_plan_set_hit_chunk(self.h, int(records)))

    def set_keep_hits(self, limit):
        """Hits kept by a search for hits() (default 2^22)."""
        check(lib().pzl_plan_set_keep_hits(self.h, int(limit)))

    def hits(self):
        n = C.c_longlong(0)
        check(lib().pzl_get_hits(self.h, None, None, 0, C.byref(n)))
        keys = np.zeros(max(n.value, 1), dtype=np.int64)
        pows = np.zeros(max(n.value, 1), dtype=np.float32)
        check(lib().pzl_get_hits(self.h, keys.ctypes.data_as(C.POINTER(C.c_longlong)), _fptr(pows), n.value,
                                 C.byref(n)))
        return self.decode_keys(keys[: n.value]), pows[: n.value]

    def decode_keys(self, keys):
        g = self.geom
        R, Z, W, S = g.corr_uselen, self.harms[0].numz, self.harms[0].numw, g.numharmstages
        r = keys % R
        k = keys // R
        z = k % Z
        k //= Z
        w = k % W
        k //= W
        stage = k % S
        rstep = k // S
        return np.stack([rstep, stage, w, z, r], axis=1)

    def kernel(self, idx, row):
        h = self.harms[idx]
        out = np.zeros(2 * h.fftlen, dtype=np.float32)
        check(lib().pzl_get_kernel(self.h, idx, row, _fptr(out)))
        return out[0::2] + 1j * out[1::2]

    def plane(self, rstep, idx):
        h = self.harms[idx]
        cap = h.numw * h.numz * (h.max_numrs + 4)
        out = np.zeros(cap, dtype=np.float32)
        numrs = C.c_int(0)
        rlo = C.c_longlong(0)
        check(lib().pzl_compute_plane(self.h, rstep, idx, _fptr(out), cap, C.byref(numrs), C.byref(rlo)))
        return out[: h.numw * h.numz * numrs.value].reshape(h.numw, h.numz, numrs.value), rlo.value

    def sums(self, rstep):
        h = self.harms[0]
        n = h.numw * h.numz * self.geom.corr_uselen
        out = np.zeros(n, dtype=np.float32)
        check(lib().pzl_compute_sums(self.h, rstep, _fptr(out), n))
        return out.reshape(h.numw, h.numz, self.geom.corr_uselen)


def tune_batchsize(cfg, fft, candidates=(4, 8, 16), band=64, device=0):
    """The autotuner's choice of rsteps per launch for this search shape and
    spectrum: (best, {batchsize: ms per rstep})."""
    a = np.ascontiguousarray(fft)
    if a.dtype == np.complex64:
        a = a.view(np.float32)
    cand = (C.c_int * len(candidates))(*candidates)
    ms = (C.c_double * len(candidates))()
    best = C.c_int(0)
    check(lib().pzl_tune_batchsize(C.byref(cfg), int(device), _fptr(a), len(a) // 2, cand, len(candidates),
                                   int(band), C.byref(best), ms))
    return best.value, {int(c): ms[i] for i, c in enumerate(candidates)}


def _opt_array(cands, fdot_only=False):
    arr = (OptCand * max(len(cands), 1))()
    for i, (r, z, w, nh) in enumerate(cands):
        arr[i].r, arr[i].z, arr[i].w, arr[i].numharm = float(r), float(z), float(w), int(nh)
        arr[i].fdot_only = int(bool(fdot_only))
    return arr


def _opt_result(arr, n):
    out = []
    for i in range(n):
        a = arr[i]
        out.append(dict(r=a.ropt, z=a.zopt, w=a.wopt,
                        derivs=[{k: getattr(a.derivs[h], k) for k, _ in RDerivs._fields_} for h in range(a.numharm)],
                        pows=[a.pows[h] for h in range(a.numharm)]))
    return out


def optimize_spectrum(fft, cands, device=0, fdot_only=False):
    """optimize_accelcand (harmonic polishing) of each (r, z, w, numharm), r
    in bins of 'fft': [{r, z, w, derivs: [...], pows: [...]}, ...];
    fdot_only: the candidates of a w = 0 search (simplex in r, z)."""
    a = np.ascontiguousarray(fft)
    if a.dtype == np.complex64:
        a = a.view(np.float32)
    arr = _opt_array(cands, fdot_only)
    check(lib().pzl_optimize_spectrum(int(device), _fptr(a), len(a) // 2, arr, len(cands)))
    return _opt_result(arr, len(cands))


def maximize_spectrum(fft, points, device=0, fdot_only=False):
    """One harmonic per (r, z, w) point, r in bins of 'fft', as
    optimize_accelcand's -noharmpolish branch: max_rzw_arr (maximize_rzw.c:3)
    or, fdot_only, max_rz_arr (maximize_rz.c:25).  [{r, z, w, derivs: [d],
    pows: [raw maximum power]}, ...]."""
    a = np.ascontiguousarray(fft)
    if a.dtype == np.complex64:
        a = a.view(np.float32)
    arr = _opt_array([(r, z, w, 1) for r, z, w in points], fdot_only)
    check(lib().pzl_maximize_spectrum(int(device), _fptr(a), len(a) // 2, arr, len(points)))
    return _opt_result(arr, len(points))


class Server:
    """The request loop (pzl_server_*): searches admitted by HBM footprint and
    run concurrently on one GPU."""

    def __init__(self, device=0, budget=0, max_inflight=2):
        self.h = C.c_void_p()
        check(lib().pzl_server_create(int(device), int(budget), int(max_inflight), C.byref(self.h)))
        self._keep = {}

    def submit(self, cfg, fft, deadline_ms=0.0):
        """Queue a search of a host spectrum; deadline_ms > 0: the request
        fails with PZL_ERR_TIMEOUT once that long has passed since now."""
        a = np.ascontiguousarray(fft)
        if a.dtype == np.complex64:
            a = a.view(np.float32)
        assert a.dtype == np.float32
        rid = C.c_longlong(0)
        if deadline_ms > 0:
            check(lib().pzl_server_submit_deadline(self.h, C.byref(cfg), _fptr(a), 0, len(a) // 2,
                                                   float(deadline_ms), C.byref(rid)))
        else:
            check(lib().pzl_server_submit(self.h, C.byref(cfg), _fptr(a), len(a) // 2, C.byref(rid)))
        self._keep[rid.value] = a  # the spectrum stays valid until the request is done
        return rid.value

    def submit_device(self, cfg, ptr, numbins, deadline_ms=0.0):
        """A spectrum of numbins complex bins in device memory at 'ptr' (kept
        valid by the caller until the request is done)."""
        rid = C.c_longlong(0)
        if deadline_ms > 0:
            check(lib().pzl_server_submit_deadline(self.h, C.byref(cfg), C.c_void_p(ptr), 1, int(numbins),
                                                   float(deadline_ms), C.byref(rid)))
        else:
            check(lib().pzl_server_submit_device(self.h, C.byref(cfg), C.c_void_p(ptr), int(numbins),
                                                 C.byref(rid)))
        return rid.value

    def wait(self, rid, maxcands=1 << 16):
        cands = (Cand * maxcands)()
        n = C.c_longlong(0)
        st = RequestStats()
        try:
            check(lib().pzl_server_wait(self.h, int(rid), cands, maxcands, C.byref(n), C.byref(st)))
        finally:
            self._keep.pop(rid, None)
        out = [(cands[i].power, cands[i].sigma, cands[i].numharm, cands[i].r, cands[i].z, cands[i].w)
               for i in range(min(n.value, maxcands))]
        return out, st.as_dict()

    def poll(self, maxn=1024):
        ids = (C.c_longlong * maxn)()
        n = C.c_int(0)
        check(lib().pzl_server_poll(self.h, ids, maxn, C.byref(n)))
        return [ids[i] for i in range(min(n.value, maxn))]

    def stats(self):
        st = ServerStats()
        check(lib().pzl_server_stats(self.h, C.byref(st)))
        return st.as_dict()

    def close(self):
        if self.h:
            lib().pzl_server_destroy(self.h)
            self.h = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
