import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("pp-es-dg-mhd_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "long: an oracle run of several minutes to tens of minutes on the CPU; "
                                       "runs when PPES_LONG=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("PPES_LONG") == "1":
        return
    skip = pytest.mark.skip(reason="long oracle run: set PPES_LONG=1 to include it")
    for item in items:
        if "long" in item.keywords:
            item.add_marker(skip)
