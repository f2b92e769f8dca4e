import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for path in (ROOT / "openmvg-rs_amd", ROOT / "tests"):
    if str(path) not in sys.path:
        sys.path.insert(0, str(path))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libba_hip.so")
    config.addinivalue_line("markers", "slow: larger CPU parity cases")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (oracle/_build/libba_oracle.so), the checker of every parity test."""
    from oracle_lib import load_oracle

    return load_oracle()
