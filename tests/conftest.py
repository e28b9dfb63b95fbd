import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "ai-ran-npu-kernels_amd"
for p in (str(PKG), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import load_oracle
    return load_oracle()


@pytest.fixture(scope="session")
def reflib():
    from oracle_lib import load_ref
    lib = load_ref()
    if lib is None:
        pytest.skip("oracle/_ref/libref.so (reference host sources) is not built on this machine")
    return lib


@pytest.fixture(scope="session")
def lib():
    from puschrx import load_library
    return load_library()


@pytest.fixture()
def gpu_stream():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        yield s
    s.synchronize()
