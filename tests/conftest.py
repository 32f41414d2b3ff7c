"""Shared test set-up: the ``gpu`` marker and the oracle library."""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    import oracle

    if not oracle.LIB_PATH.exists():
        oracle.build()
