import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: longer CPU test")


def pytest_runtest_setup(item):
    # Print and flush which test is about to run, so a GPU fault that ends the
    # process still names its case in the captured output.
    if item.get_closest_marker("gpu") is not None:
        sys.stderr.write(f"\n[gpu-test] {item.nodeid}\n")
        sys.stderr.flush()
