import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "physics-engine_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref (the reference engine compiled from its sources)")


def pytest_collection_modifyitems(config, items):
    from pyoracle import ref_available
    skip_ref = pytest.mark.skip(reason="oracle/_ref not built (reference checkout absent at build time)")
    for it in items:
        if "ref" in it.keywords and not ref_available(32):
            it.add_marker(skip_ref)
