import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP solve phase)")


def _built(path):
    return os.path.exists(os.path.join(ROOT, path))


@pytest.fixture(scope="session", autouse=True)
def _native_libs():
    """Build what is missing (never rebuild what exists: GPU runs use the prebuilt files)."""
    if not _built("chronoslab_amd/libchronos_setup.so"):
        from chronoslab_amd import build as B

        B.build_setup()
    if not _built("oracle/liboracle.so"):
        import oracle

        oracle.build(reference=False)
    yield


@pytest.fixture(scope="session")
def gpu():
    from chronoslab_amd import amg

    amg.init(0)
    return amg
