"""Test configuration: the gpu marker (tests that need an MI355X) and the import paths of the package and the
oracle."""

import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "cpp-spo-dvr_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (run with -m gpu on the MI355X)")
    config.addinivalue_line("markers", "slow: a long CPU run (minutes)")
