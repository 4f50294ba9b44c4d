import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for sub in ("manifolds-rs_amd", "oracle", "tests"):
    p = str(ROOT / sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs on the GPU box")


@pytest.fixture(scope="session")
def ctx():
    import mrs_amd

    c = mrs_amd.Context(0)
    yield c
    c.close()
