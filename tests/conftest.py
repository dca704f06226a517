import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "oracle", "enzo-e_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line(
        "markers", "slow: long CPU case, runs only with VLCT_RUN_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("VLCT_RUN_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow case; set VLCT_RUN_SLOW=1 to run")
    for item in items:
        if "slow" in item.keywords:
            item.add_marker(skip)
