import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long CPU test")


def has_ref() -> bool:
    return os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libcathy_ref.so"))


@pytest.fixture(scope="session")
def oracle():
    """The C oracle (oracle/liboracle.so), built with gcc if absent."""
    import subprocess
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle"], check=True,
                       stdout=subprocess.DEVNULL)
    import pyoracle
    return pyoracle.Oracle()
