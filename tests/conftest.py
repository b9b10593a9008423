import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU host only)")
    config.addinivalue_line("markers", "slow: long CPU case")
    # The oracle's own library is cheap to build; the reference-operator library only
    # builds where the reference checkout exists (tests needing it skip otherwise).
    so = os.path.join(ROOT, "oracle", "_build", "libcpu_ref.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"], check=False,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    if os.path.isdir("/root/reference/src") and not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref_ops_f.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=False,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def rng():
    import numpy as np
    return np.random.default_rng(1234)


@pytest.fixture(scope="session", autouse=True)
def _one_hip_runtime(request):
    """PyTorch-ROCm bundles its own libamdhip64.so.7 (same soname as the system one the
    library links). Initialising torch before the library is first loaded makes both
    bind to that one runtime; the other order leaves torch without a device."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    yield
