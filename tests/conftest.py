import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libkapsel_hip.so)")
    # the oracle is the CPU library the tests check against; build it if absent
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


@pytest.fixture(scope="session")
def gpu():
    """The product library, loaded for a GPU test; fails loudly if absent."""
    import kapsel3_amd
    kapsel3_amd.lib()
    return kapsel3_amd
