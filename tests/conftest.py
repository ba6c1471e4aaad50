import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
for p in (str(ROOT), str(ROOT / "oracle"), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _built(path: pathlib.Path) -> bool:
    return path.exists() and path.stat().st_size > 0


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc  # oracle/oracle.py
    if not _built(orc.LIB):
        orc.build(with_ref=True)
    return orc


@pytest.fixture(scope="session")
def ttmlib():
    """The product's ctypes handle (built in-tree if nothing is built yet)."""
    import ttm_amd
    if not _built(ttm_amd.LIB_PATH):
        from ttm_amd import build
        build.build()
    return ttm_amd


@pytest.fixture(scope="session")
def gpu(ttmlib):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    torch.cuda.init()
    return torch.device("cuda:0")


@pytest.fixture(autouse=True)
def _reset_force(request):
    yield
    if "ttmlib" in request.fixturenames or "gpu" in request.fixturenames:
        import ttm_amd
        ttm_amd.force_kernel(None)
