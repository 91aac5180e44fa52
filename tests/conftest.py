import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcdjhip.so)")


def pytest_runtest_setup(item):
    if "gpu" in item.keywords:
        print("\n[gpu-test] %s" % item.nodeid, flush=True)
