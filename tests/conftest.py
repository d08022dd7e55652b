import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
for p in (REPO, REPO / "caelix-renderer_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def cx():
    """The product binding; loads libcaelix_amd.so (fails loudly if it is missing)."""
    import caelix_amd

    caelix_amd.load()
    return caelix_amd


@pytest.fixture(scope="session")
def oracle():
    import oracle as o

    o.load()
    return o
