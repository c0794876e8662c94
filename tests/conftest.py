import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "tope-fft_amd", "python"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU oracle work")
    # torch (used by the caller-stream tests) must bring up its HIP runtime
    # before libtopefft does in this process, or it finds no device: import
    # it first whenever GPU tests may run
    if "not gpu" not in (config.getoption("markexpr") or ""):
        try:
            import torch  # noqa: F401
        except ImportError:
            pass


def rel_l2(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="session")
def oracle():
    import oracle_py
    oracle_py.load()
    return oracle_py


@pytest.fixture(scope="session")
def golden():
    path = os.path.join(REPO, "tests", "golden", "golden_fft.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def fw():
    import topefft
    f = topefft.Framework()
    yield f
    f.destroy()
