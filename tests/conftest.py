import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libkma.so on the device)")


def pytest_collection_modifyitems(config, items):
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(pytest.mark.timeout(90))
