import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "large: multi-GB GPU case")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hip():
    """The HIP binding on a GPU; raises (fails the test) when the library is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test collected without a GPU")
    import libmarshal_amd as lm
    lm.lib()
    return lm
