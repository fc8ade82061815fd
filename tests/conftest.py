import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    # torch carries its own HIP runtime, which has to reach the device before
    # the library's does (bench.py keeps the same order); on a host without a
    # GPU this only finds none
    import torch
    if torch.cuda.is_available():
        torch.cuda.init()


def load_npz(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def filter_fixtures():
    return sorted(f for f in os.listdir(GOLDEN) if f.startswith("filter_") and f.endswith(".npz"))


@pytest.fixture(scope="session")
def gpu_available():
    # torch (which some GPU tests use for device buffers) carries its own HIP
    # runtime; it has to initialise the device before the library's runtime
    # does, as in bench.py
    import torch
    torch.cuda.init()
    assert torch.cuda.is_available(), "GPU tests need a visible GPU: torch reports none"
    from microtomodenoise_amd import _capi
    n = _capi.lib().inlm_device_count()
    assert n >= 1, "GPU tests need a visible GPU: the HIP library reports none"
    return n
