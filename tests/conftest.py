"""Test configuration: import paths and the ``gpu`` marker.

GPU tests (``-m gpu``) call the HIP library through the product API and
compare with the FP64 oracle; everything else runs on the CPU.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "astra-kernelkit_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (HIP device) and the built library")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def device():
    if not gpu_available():
        pytest.skip("no HIP device")
    import torch
    return torch.device("cuda", 0)


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def assert_parity(got, ref, rel=1e-5, point=1e-4, label=""):
    """The parity bar: relative L2 <= rel and |got - ref| <= point*max|ref|."""
    import numpy as np
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, (label, got.shape, ref.shape)
    e = rel_l2(got, ref)
    scale = float(np.abs(ref).max()) if ref.size else 0.0
    worst = float(np.abs(got - ref).max()) if ref.size else 0.0
    assert e <= rel, f"{label}: relative L2 {e:.3e} > {rel:.1e}"
    assert worst <= point * max(scale, 1e-30), (
        f"{label}: max abs err {worst:.3e} > {point:.0e} * {scale:.3e}")
    return e, worst / max(scale, 1e-30)
