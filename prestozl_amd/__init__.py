"""MI355X-native jerk-search core (PrestoZL accelsearch hot path).

The search runs in libaccelsearch_lib.so (HIP kernels for gfx950 behind the
C ABI of include/pzl_accel.h and the accel.h drop-in).  This package holds the
ctypes binding, the build step, synthetic inputs and the serving loop.
"""
