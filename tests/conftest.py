import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

# torch's wheel carries its own HIP runtime under the same soname as the
# system one; whichever loads first serves the whole process. Loading torch
# before libsplat_hip.so keeps one runtime for the tests that use both.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


def pytest_report_header(config):
    return f"splat_amd tests (root {ROOT}, pid {os.getpid()})"
