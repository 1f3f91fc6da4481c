import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def lbs_fixture():
    import json
    return json.loads((ROOT / "tests/golden/lbs_matrix.json").read_text())


@pytest.fixture(scope="session")
def engine():
    import privacyamplification_amd as P
    eng = P.Engine()
    yield eng
    eng.close()
