import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"
DATA = ROOT / "data"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def _built():
    pkg = ROOT / "gip_amd"
    return all((pkg / f).exists() for f in ("libgip_hip.so", "libgip_cavity.so", "libgip_dns.so")) and \
        (ROOT / "oracle" / "_build" / "liboracle.so").exists()


@pytest.fixture(scope="session", autouse=True)
def native_built():
    """Build the native parts once when nothing is built yet."""
    if not _built():
        from gip_amd import build
        build.build_cavity()
        build.build_hip()
        build.build_dns()
        build.build_oracle()
    return True


@pytest.fixture(scope="session")
def cav46_dir(tmp_path_factory):
    """The 46^3 lid-driven cavity (Re=100) written in the InputFiles layout."""
    from gip_amd import cavity
    d = tmp_path_factory.mktemp("cav46")
    cavity.generate(46, gamma=1e-2).write(d, "cav46")
    return d


@pytest.fixture(scope="session")
def c100k_root():
    """Root holding the one-partition C100K case (merged from the shipped four partitions)."""
    marker = DATA / "InputFiles" / "Topos" / "C100K" / "C100K_1p" / "topoColsC100K_1p-0.topo"
    if not marker.exists():
        from gip_amd import build
        if not build.build_data():
            pytest.skip("C100K one-partition data absent (data/ is produced by build() where the reference is present)")
    return DATA
