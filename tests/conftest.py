import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# every workspace, batch slot and scratch buffer the library allocates starts as NaN bytes, so that a
# kernel reading memory no kernel wrote shows up in the outputs the tests compare
os.environ.setdefault("FASTFFT_POISON", "1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; run with -m gpu")
    config.addinivalue_line("markers", "slow: large-size case (seconds of CPU oracle work)")


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    path = os.path.join(ROOT, "tests", "golden", "golden_small.npz")
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}
