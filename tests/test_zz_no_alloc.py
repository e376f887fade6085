"""pzl_realfft_device_noalloc: the realfft transforms with the library's fixed
device workspace instead of allocated memory.  Its output must equal
pzl_realfft_device_batch's bit for bit (the same kernels on the same data),
on two calls in a row whose second is smaller, and a request beyond the
workspace must fail with an error, leaving the data untouched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    return ctypes.CDLL("libamdhip64.so")


def _run(n, nseries, isign, seed):
    from prestozl_amd import binding as B
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n * nseries).astype(np.float32)
    a, b = B.DeviceArray(x.nbytes), B.DeviceArray(x.nbytes)
    a.upload(x)
    b.upload(x)
    B.realfft_device(a.ptr, n, isign, nseries=nseries)
    B.realfft_device_noalloc(b.ptr, n, isign, nseries=nseries)
    assert _hip().hipDeviceSynchronize() == 0
    ra = a.download((nseries, n), np.float32)
    rb = b.download((nseries, n), np.float32)
    a.free()
    b.free()
    return ra, rb


def test_noalloc_equals_allocating_entry_point_twice():
    # the largest transform the workspace holds, then a smaller batch
    for n, nseries, isign, seed in ((1 << 22, 1, -1, 1), (1 << 18, 4, 1, 2)):
        ra, rb = _run(n, nseries, isign, seed)
        print("n %d x %d, isign %d: identical %s" % (n, nseries, isign, np.array_equal(ra, rb)))
        assert np.array_equal(ra, rb)


def test_noalloc_small_single_workgroup_lengths():
    ra, rb = _run(1 << 12, 3, -1, 3)
    assert np.array_equal(ra, rb)


def test_noalloc_rejects_beyond_workspace():
    from prestozl_amd import binding as B
    n = 1 << 20
    x = np.arange(n * 8, dtype=np.float32)
    d = B.DeviceArray(x.nbytes)
    d.upload(x)
    with pytest.raises(B.PzlError) as e:
        B.realfft_device_noalloc(d.ptr, n, -1, nseries=8)  # 2^23 samples
    print(e.value)
    assert "workspace" in str(e.value)
    assert _hip().hipDeviceSynchronize() == 0
    assert np.array_equal(d.download((n * 8,), np.float32), x)
    d.free()
