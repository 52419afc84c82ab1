import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; parity tests of the HIP engine")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as orc
    if not orc.ORACLE_LIB.exists():
        orc.build(ref=False)
    return orc.oracle()


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    d = ROOT / "tests" / "golden"

    def load(name):
        return dict(np.load(d / name, allow_pickle=False))
    return load
