import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; run with -m gpu")


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    with np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu():
    """Initialise the HIP library on cuda:0.  Raises (never falls back) when the
    library is missing; skips only when the machine has no GPU at all."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU on this machine")
    import estl_amd
    estl_amd.lib()  # raises if libestl_hip.so is missing
    estl_amd.init(0)
    return torch.device("cuda:0")
