import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def dspx():
    return importlib.import_module("dspsr-kanyixia_amd")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx(dspx):
    import torch
    assert torch.cuda.is_available(), "GPU test needs a HIP device"
    torch.cuda.init()
    # one stream for torch's allocations / fills and for the library's kernels
    # (the default stream's handle is 0, which the library would replace with a
    # stream of its own, unordered with torch's work)
    s = torch.cuda.Stream()
    torch.cuda.set_stream(s)
    c = dspx.Context(0, s.cuda_stream)
    yield c
    c.synchronize()
    c.close()
