import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def config1_data():
    from prestozl_amd import synth
    fft, N, dt, opts = synth.config1()
    return fft, N, dt, opts


def rel_err(a, b, floor=None):
    """max |a-b| / max(|b|, floor) elementwise, and RMS relative error."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b) if floor is None else np.maximum(np.abs(b), floor)
    r = np.abs(a - b) / den
    return float(r.max()), float(np.sqrt(np.mean(r * r)))
