"""Test configuration: the `gpu` marker, import paths and shared fixtures."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "gkm-dns_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgkm.so on the device)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    if not O.ORACLE_SO.exists():
        O.build(ref=False)
    return O
