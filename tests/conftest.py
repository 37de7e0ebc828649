import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real GPU (MI355X) and the built libpsy_gpu.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import ora
    if not os.path.exists(ora.LIB):
        ora.build(with_ref=False)
    return ora


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    from psychopath_cpp_amd import gpu
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    ctx = gpu.Context(0)
    yield ctx
    ctx.close()
