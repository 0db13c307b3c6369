import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture
def c1_small(tmp_path):
    """A 10 A TIP3P droplet (qdyn-format files) for quick CPU/GPU parity."""
    from qsource_amd import gensys
    s = gensys.water_droplet(10.0, 1001)
    paths = gensys.write_system(s, str(tmp_path), "c1s", steps=100, output=10)
    return paths


@pytest.fixture
def c1(tmp_path):
    """Configuration 1: R = 15 A TIP3P droplet, seeded Maxwell velocities."""
    from qsource_amd import gensys
    s = gensys.water_droplet(15.0, 1001)
    paths = gensys.write_system(s, str(tmp_path), "c1", steps=100, output=10)
    return paths
