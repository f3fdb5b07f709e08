import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "contest-sigmod2025_amd"), ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


def _have_gpu():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def ctx():
    import gpujoin

    if not _have_gpu():
        pytest.skip("no GPU in this container")
    c = gpujoin.Context(0)
    yield c
    c.close()
