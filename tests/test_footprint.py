"""The HBM footprint model that admission uses (the exact replacement of
accelsearch_maxsize, accelsearch_maxsize.c:551-619, which adds a fixed 286 MB
context and a 1.1 factor to estimates of its own buffers).

For each search shape: the model of a created plan equals, byte for byte, what
the plan allocated (counted at every allocation) plus its kernel bank; the
model before the plan exists is not below it (the hit sort reports its scratch
without the device, so the two agree); and the device's free memory drops by that amount
plus at most the runtime allowance that admission adds per search
(PZL_RUNTIME_BYTES_PER_SEARCH, 32 MiB: the HIP runtime's memory behind the
plan's streams; measured up to 27 MiB)."""
import pytest

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def _shapes():
    from prestozl_amd import synth
    fft1, N1, dt1, o1 = synth.config1()
    fft2, N2, dt2, _, _ = synth.config2(N=1 << 21)
    return [
        ("config 1: zmax 20, wmax 0, 4 harmonics", fft1,
         dict(N=N1, dt=dt1, numbins=len(fft1), zmax=20, wmax=0, numharm=4, sigma=2.0, wmaxP=False)),
        ("N=2^21, zmax 200, wmax 300, 16 harmonics, batchsize 8", fft2,
         dict(N=N2, dt=dt2, numbins=len(fft2), zmax=200, wmax=300, numharm=16, sigma=5.0, batchsize=8)),
        ("N=2^21, zmax 200, wmax 150, 16 harmonics, batchsize 3", fft2,
         dict(N=N2, dt=dt2, numbins=len(fft2), zmax=200, wmax=150, numharm=16, sigma=5.0, batchsize=3)),
        ("N=2^21, zmax 50, wmax 0, 8 harmonics, batchsize 16", fft2,
         dict(N=N2, dt=dt2, numbins=len(fft2), zmax=50, wmax=0, numharm=8, sigma=5.0, wmaxP=False, batchsize=16)),
    ]


def test_footprint_model_is_exact():
    from prestozl_amd import binding as B
    # load the library's code objects and runtime state before measuring
    label, fft, kw = _shapes()[0]
    with B.Plan(B.make_config(**kw)) as p:
        p.set_spectrum(fft)
    B.bank_cache_clear()
    rows = []
    for label, fft, kw in _shapes():
        cfg = B.make_config(**kw)
        est = B.geometry(cfg).device_bytes
        free0 = B.device_memory(0)[0]
        p = B.Plan(cfg)
        p.set_spectrum(fft)
        held = p.device_bytes()
        model = p.geometry().device_bytes
        free1 = B.device_memory(0)[0]
        p.close()
        B.bank_cache_clear()
        free2 = B.device_memory(0)[0]
        drop = free0 - free1
        print("%s: model %.2f MiB, held %.2f MiB, before the plan %.2f MiB, free memory drop %.2f MiB, "
              "not returned at close %.2f MiB" % (label, model / MiB, held / MiB, est / MiB, drop / MiB,
                                                  (free0 - free2) / MiB), flush=True)
        rows.append((label, model, held, est, drop))
    for label, model, held, est, drop in rows:
        assert held == model, label
        assert est >= model, label
        # the runtime's memory behind the plan's streams, and the 2 MiB
        # granularity of the free-memory figure
        assert -8 * MiB <= drop - held <= B.RUNTIME_BYTES_PER_SEARCH, label
