import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "stencil-optimalizations_amd")
for p in (ROOT, PKG, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: CPU test that takes tens of seconds")


def pytest_collection_modifyitems(config, items):
    # GPU tests print the case they run first, so a GPU fault names its test
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(pytest.mark.timeout(300))
