import os
import sys

import pytest

WURZEL = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if WURZEL not in sys.path:
    sys.path.insert(0, WURZEL)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI of libgcsim.so)")
    config.addinivalue_line("markers", "langsam: CPU test that takes more than a few seconds")


def pytest_collection_modifyitems(config, items):
    # GPU tests need the device; fail loudly if the library is missing rather than skipping.
    pass
