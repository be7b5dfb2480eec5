"""Shared test setup: import paths, the `gpu` marker, common fixtures.

CPU tests (`-m "not gpu"`) pin the oracle and check the host logic and the C-ABI
library's exports; GPU tests (`-m gpu`) run the HIP path through the C-ABI and
compare it with the oracle.
"""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "river-basin-sim_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def orc():
    """The CPU restatement oracle (built by `make -C oracle`)."""
    from oracle import oracle

    oracle.lib()
    return oracle
