import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "raw-native_amd"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def ctx():
    """One GPU context for the session (the HIP library must be built and a GPU present).
    torch loads its HIP runtime before the library does, so tests can share device memory."""
    import torch  # noqa: F401
    import raw_ao
    c = raw_ao.Context(0)
    yield c
    c.close()
