import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: full-size configuration (seconds to minutes)")


@pytest.fixture(scope="session")
def golden_codebooks():
    return dict(np.load(os.path.join(GOLDEN, "codebooks.npz"), allow_pickle=False))


@pytest.fixture(scope="session")
def golden_tables():
    return dict(np.load(os.path.join(GOLDEN, "tables.npz"), allow_pickle=False))


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def capi():
    from huffman_coding_gap_arrays_amd import _capi

    return _capi.load()


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch.device("cuda:0")
