import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


GOLDEN = os.path.join(ROOT, "tests", "golden")
GOLDEN_DATA = os.path.join(GOLDEN, "data")


@pytest.fixture(scope="session")
def ene():
    import geauxdock_amd as g
    return g.load_enepara()


@pytest.fixture(scope="session")
def golden_job(ene):
    import geauxdock_amd as g
    d = GOLDEN_DATA
    return g.load_complex(os.path.join(d, "1a07C1.sdf"), os.path.join(d, "1a07C1-0.8.ff"),
                          os.path.join(d, "1a07C-1.pdb"), "1a07C1", ene=ene, n_tmp=1,
                          floor_temp=0.044, ceiling_temp=0.036)


@pytest.fixture(scope="session")
def oracle():
    from tests import oracle_ffi
    return oracle_ffi.Oracle()
