import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

# The flat FP64 stage runs on the tile kernel when its tiles fill the device
# and as split launches on smaller grids.  The test grids are all small: keep
# them on the tile kernel (the one the 512x512x128 workload runs), and let
# the tests of the split launches ask for them (SCALE_HIP_FLAT_SPLIT=1).
os.environ.setdefault("SCALE_HIP_FLAT_SPLIT", "0")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libscalehip.so")
