import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C-ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.build()
    return o


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from fast_rpa_eval_amd import _lib
    _lib.lib()  # raises if libfrpa.so is missing: no silent fallback
    print(f"[gpu] {torch.cuda.get_device_name(0)}; {_lib.lib().frpa_version().decode()}", flush=True)
    return torch.device("cuda:0")
