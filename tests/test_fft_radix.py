"""The in-register DFTs of the correlation kernels (pzl_fft.h), compiled for
the host and checked against a direct double-precision DFT for every radix
and both signs, plus the compile-time twiddle constants."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "prestozl_amd", "csrc")


def test_radix_kernels(tmp_path):
    exe = str(tmp_path / "fft_radix_check")
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-x", "hip", "--offload-host-only", "--offload-arch=gfx950",
           "-I" + CSRC, os.path.join(HERE, "native", "fft_radix_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        pytest.fail("build failed:\n" + r.stderr[-3000:])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("radix") == 20
