import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nutcracker-1_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

LIB = os.path.join(PKG, "nutcracker_amd", "libnutcracker_hip.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _ensure_built():
    if not os.path.exists(LIB):
        sys.path.insert(0, PKG)
        import build  # nutcracker-1_amd/build.py
        build.build()


def _gpu_visible() -> bool:
    """Ask the HIP core itself (no torch import: the core links /opt/rocm's HIP runtime).
    Where a GPU driver node exists, a failure to reach the device is an error, not a skip."""
    if not os.path.exists("/dev/kfd"):
        return False
    _ensure_built()
    from nutcracker_amd import core
    info = core.device_info()
    print("\nHIP device:", info, flush=True)
    return True


def pytest_collection_modifyitems(config, items):
    if not any("gpu" in it.keywords for it in items):
        return
    if _gpu_visible():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
