import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "momentumed.jl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libedm.so kernels)")
    config.addinivalue_line("markers", "slow: large-size GPU checks")
    # a solve that stops with unconverged pairs fails the test that ran it, unless the test
    # expects it (pytest.warns)
    config.addinivalue_line("filterwarnings", "error:Lanczos eigsolve stopped")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_ctx():
    from momentumed import device as dev
    if not dev.activate_gpu_device("rocm"):
        pytest.fail("no GPU visible to libedm.so")
    return dev.get_context(0)
