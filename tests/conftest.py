import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"

# torch ships its own HIP runtime, and libpyr3hip.so binds to the one the
# process holds when it loads: importing torch first keeps one runtime in
# the test process whichever tests run (the torch-reference tests need
# torch's device).
try:
    import torch  # noqa: F401
except ImportError:
    pass

# Specialised chaos kernels compiled by one test process are reused by the
# next (and by a host of the same arch) from this directory.
os.environ.setdefault("P3_KERNEL_CACHE", str(ROOT / "pyr3_amd" / "kcache"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.load()
    return oracle
