import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    lib = ROOT / "axiomr_amd" / "libaxiomr_amd.so"
    orc = ROOT / "oracle" / "_build" / "liboracle.so"
    if not lib.exists():
        from axiomr_amd import build as B
        B.build()
    if not orc.exists():
        import oracle
        oracle.build()


@pytest.fixture(scope="session", autouse=True)
def built_libraries():
    _ensure_built()


RASTER_MODES = {"binned": 0, "visibility": 1}


@pytest.fixture(scope="session", params=list(RASTER_MODES))
def ctx(request):
    """One GPU context per raster mode (GPU tests only): every GPU test runs
    in both modes, which must give the same pixels."""
    import axiomr_amd as axr
    c = axr.Context(0)
    c.set_config(raster_mode=RASTER_MODES[request.param])
    yield c
    c.close()
