import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libarcwarden_hip.so)")
    config.addinivalue_line("markers", "slow: long-running physics test")


@pytest.fixture(scope="session")
def oracle():
    import pic_oracle
    pic_oracle.build()
    return pic_oracle


@pytest.fixture(scope="session")
def gpu():
    """Loads the HIP library; fails loudly (no fallback) if it is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import arcwarden_amd
    arcwarden_amd.lib()
    return arcwarden_amd


@pytest.fixture(autouse=True)
def _name_case(request):
    # print the case name before each test so a GPU fault can be attributed
    if os.environ.get("AW_TEST_ECHO", "1") == "1":
        print(f"\n>>> {request.node.nodeid}", flush=True)
    yield
