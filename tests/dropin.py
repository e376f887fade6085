"""ctypes view of the accel.h drop-in API (include/accel.h), as a caller such
as the reference's ctypes pipeline (bin/accelsearch_pipeline_cu.py:51-233)
binds it: the struct layouts below are written from the field lists of
include/accel.h, include/makeinf.h and include/accelsearch_cmd.h of the
reference."""
import ctypes as C
import os

from prestozl_amd import binding

GLIB = "libglib-2.0.so.0"


class FComplex(C.Structure):
    _fields_ = [("r", C.c_float), ("i", C.c_float)]


class InfoData(C.Structure):
    _fields_ = ([(n, C.c_double) for n in ("ra_s", "dec_s", "N", "dt", "fov", "mjd_f", "dm", "freq", "freqband",
                                          "chan_wid", "wavelen", "waveband", "energy", "energyband")]
                + [("onoff", C.c_double * 80)]
                + [(n, C.c_int) for n in ("num_chan", "mjd_i", "ra_h", "ra_m", "dec_d", "dec_m", "bary",
                                          "numonoff")]
                + [("notes", C.c_char * 500), ("name", C.c_char * 200), ("object", C.c_char * 100),
                   ("instrument", C.c_char * 100), ("observer", C.c_char * 100), ("analyzer", C.c_char * 100),
                   ("telescope", C.c_char * 40), ("band", C.c_char * 40), ("filt", C.c_char * 7)])


def _opt(name, ctype):
    return [(name + "P", C.c_char), (name, ctype), (name + "C", C.c_int)]


class Cmdline(C.Structure):
    _fields_ = (_opt("batchsize", C.c_int) + _opt("ncpus", C.c_int) + _opt("lobin", C.c_int)
                + _opt("numharm", C.c_int) + _opt("zmax", C.c_int) + _opt("wmax", C.c_int)
                + _opt("sigma", C.c_float) + _opt("rlo", C.c_double) + _opt("rhi", C.c_double)
                + _opt("flo", C.c_double) + _opt("fhi", C.c_double)
                + [("inmemP", C.c_char), ("photonP", C.c_char), ("medianP", C.c_char), ("locpowP", C.c_char)]
                + _opt("zaplist", C.c_char_p) + _opt("baryv", C.c_double)
                + [("otheroptP", C.c_char), ("noharmpolishP", C.c_char), ("noharmremoveP", C.c_char),
                   ("argc", C.c_int), ("argv", C.POINTER(C.c_char_p)), ("full_cmd_line", C.c_char_p),
                   ("cudaP", C.c_char), ("cuda", C.c_char)])


class AccelObs(C.Structure):
    _fields_ = ([(n, C.c_longlong) for n in ("N", "numbins", "lobin", "highestbin")]
                + [(n, C.c_int) for n in ("maxkernlen", "corr_uselen", "fftlen", "numharmstages", "numz", "numw",
                                          "numbetween", "numzap", "dat_input", "mmap_file", "inmem",
                                          "norm_type")]
                + [(n, C.c_double) for n in ("dt", "T", "rlo", "rhi", "dr", "zlo", "zhi", "dz", "wlo", "whi",
                                             "dw", "baryv")]
                + [("nph", C.c_float), ("sigma", C.c_float), ("powcut", C.POINTER(C.c_float)),
                   ("ffdotplane", C.POINTER(C.c_float)), ("lobins", C.POINTER(C.c_double)),
                   ("hibins", C.POINTER(C.c_double)), ("numindep", C.POINTER(C.c_longlong)),
                   ("fftfile", C.c_void_p), ("workfile", C.c_void_p), ("fft", C.POINTER(FComplex)),
                   ("rootfilenm", C.c_char_p), ("candnm", C.c_char_p), ("accelnm", C.c_char_p),
                   ("workfilenm", C.c_char_p), ("use_harmonic_polishing", C.c_int)])


class AccelCand(C.Structure):
    _fields_ = [("power", C.c_float), ("sigma", C.c_float), ("numharm", C.c_int), ("r", C.c_double),
                ("z", C.c_double), ("w", C.c_double), ("pows", C.c_void_p), ("hirs", C.c_void_p),
                ("hizs", C.c_void_p), ("hiws", C.c_void_p), ("derivs", C.c_void_p)]


class Kernel(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("z", "w", "fftlen", "numgoodbins", "numbetween", "kern_half_width")] + \
               [("data", C.POINTER(FComplex))]


class SubharmInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("numharm", "harmnum", "zmax", "wmax", "numkern_zdim", "numkern_wdim",
                                       "numkern")] + \
               [("kern", C.POINTER(C.POINTER(Kernel))), ("rinds", C.POINTER(C.c_ushort)),
                ("zinds", C.POINTER(C.c_ushort))]


class FFDotPows(C.Structure):
    _fields_ = [("rlo", C.c_longlong), ("zlo", C.c_int), ("wlo", C.c_int), ("numrs", C.c_int), ("numzs", C.c_int),
                ("numws", C.c_int), ("powers", C.POINTER(C.POINTER(C.POINTER(C.c_float)))),
                ("rinds", C.POINTER(C.c_ushort)), ("zinds", C.POINTER(C.c_ushort))]


class GSList(C.Structure):
    pass


GSList._fields_ = [("data", C.c_void_p), ("next", C.POINTER(GSList))]

_api = None
_glib = None


def api():
    global _api
    if _api is None:
        L = binding.lib()
        L.parseCmdline.restype = C.POINTER(Cmdline)
        L.parseCmdline.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
        L.create_accelobs.argtypes = [C.POINTER(AccelObs), C.POINTER(InfoData), C.POINTER(Cmdline), C.c_int]
        L.free_accelobs.argtypes = [C.POINTER(AccelObs)]
        L.create_subharminfos.restype = C.POINTER(C.POINTER(SubharmInfo))
        L.create_subharminfos.argtypes = [C.POINTER(AccelObs)]
        L.free_subharminfos.argtypes = [C.POINTER(AccelObs), C.POINTER(C.POINTER(SubharmInfo))]
        L.sort_accelcands.restype = C.POINTER(GSList)
        L.sort_accelcands.argtypes = [C.POINTER(GSList)]
        L.eliminate_harmonics.restype = C.POINTER(GSList)
        L.eliminate_harmonics.argtypes = [C.POINTER(GSList), C.POINTER(C.c_int)]
        L.get_fourier_amplitudes.restype = C.POINTER(FComplex)
        L.get_fourier_amplitudes.argtypes = [C.c_longlong, C.c_int, C.POINTER(AccelObs)]
        L.subharm_fderivs_vol.restype = C.POINTER(FFDotPows)
        L.subharm_fderivs_vol.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(SubharmInfo),
                                          C.POINTER(AccelObs)]
        L.free_ffdotpows.argtypes = [C.POINTER(FFDotPows)]
        L.add_ffdotpows.argtypes = [C.POINTER(FFDotPows), C.POINTER(FFDotPows), C.c_int, C.c_int]
        L.search_ffdotpows.restype = C.POINTER(GSList)
        L.search_ffdotpows.argtypes = [C.POINTER(FFDotPows), C.c_int, C.POINTER(AccelObs), C.POINTER(GSList)]
        L.candidate_sigma.restype = C.c_double
        L.candidate_sigma.argtypes = [C.c_double, C.c_int, C.c_double]
        L.power_for_sigma.restype = C.c_double
        L.power_for_sigma.argtypes = [C.c_double, C.c_int, C.c_double]
        L.free_accelcand.argtypes = [C.c_void_p, C.c_void_p]
        _api = L
    return _api


def glib():
    global _glib
    if _glib is None:
        G = C.CDLL(GLIB)
        G.g_slist_append.restype = C.POINTER(GSList)
        G.g_slist_append.argtypes = [C.POINTER(GSList), C.c_void_p]
        G.g_slist_length.restype = C.c_uint
        G.g_slist_length.argtypes = [C.POINTER(GSList)]
        G.g_slist_free.argtypes = [C.POINTER(GSList)]
        G.g_malloc0.restype = C.c_void_p
        G.g_malloc0.argtypes = [C.c_size_t]
        _glib = G
    return _glib


def argv(args):
    arr = (C.c_char_p * (len(args) + 1))(*[a.encode() for a in args], None)
    return len(args), arr


def make_list(cands):
    """A GSList of accelcand (malloc'd, freed by free_accelcand) from
    (power, sigma, numharm, r, z, w) tuples."""
    G = glib()
    lst = C.POINTER(GSList)()
    libc = C.CDLL(None)
    libc.calloc.restype = C.c_void_p
    libc.calloc.argtypes = [C.c_size_t, C.c_size_t]
    for c in cands:
        p = libc.calloc(1, C.sizeof(AccelCand))
        a = AccelCand.from_address(p)
        a.power, a.sigma, a.numharm, a.r, a.z, a.w = c
        lst = G.g_slist_append(lst, p)
    return lst


def read_list(lst):
    out = []
    node = lst
    while node:
        a = AccelCand.from_address(node.contents.data)
        out.append((a.power, a.sigma, a.numharm, a.r, a.z, a.w))
        node = node.contents.next
    return out


def free_list(lst):
    node = lst
    while node:
        api().free_accelcand(node.contents.data, None)
        node = node.contents.next
    glib().g_slist_free(lst)


def setup_files(tmpdir, fft, N, dt, name="obs"):
    from prestozl_amd import synth
    root = os.path.join(str(tmpdir), name)
    synth.write_inf(root, N, dt)
    synth.write_fft(root, fft)
    return root
