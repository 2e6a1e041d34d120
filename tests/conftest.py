import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as om
    if not om.LIB_ORACLE.exists():
        om.build(ref=False)
    return om.Oracle()


def _ref(dim):
    from oracle import oracle as om
    if not om.REF_LIBS[dim].exists():
        pytest.skip(f"reference library {om.REF_LIBS[dim].name} not built (needs the reference sources)")
    return om.Ref(dim)


@pytest.fixture(scope="session")
def ref1d():
    return _ref("1d")


@pytest.fixture(scope="session")
def ref2d():
    return _ref("2d")


@pytest.fixture(scope="session")
def ref3d():
    return _ref("3d")


@pytest.fixture(scope="session")
def refhlld():
    return _ref("hlld")


@pytest.fixture(scope="session")
def refsp():
    return _ref("spicule")


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from utrip_amd import cans
    cans.lib()  # fails loudly if the HIP library is missing
    return torch.device("cuda:0")
