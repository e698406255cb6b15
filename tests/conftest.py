import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP solve phase)")


@pytest.fixture(scope="session", autouse=True)
def _native_libs():
    """Rebuild the host set-up library and the oracle when their sources changed since the
    last build (a digest of the sources is stored beside each library).  CHRONOS_PREBUILT=1
    keeps whatever is there (runs that must use shipped prebuilt files)."""
    if os.environ.get("CHRONOS_PREBUILT") != "1":
        from chronoslab_amd import build as B

        if not B.is_current(B.os.path.join(B.HERE, "libchronos_setup.so"), B.setup_sources()):
            B.build_setup()
        import oracle

        if not oracle.is_current():
            oracle.build(reference=False)
    yield


@pytest.fixture(scope="session")
def gpu():
    from chronoslab_amd import amg, build as B

    lib = B.os.path.join(B.HERE, "libchronos_amg.so")
    if os.environ.get("CHRONOS_PREBUILT") != "1" and not B.is_current(lib, B.amg_sources()):
        pytest.fail("libchronos_amg.so is older than its sources: run python -m chronoslab_amd.build")

    amg.init(0)
    return amg
