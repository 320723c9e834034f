import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "ref: needs the compiled reference core oracle/_ref")


def pytest_collection_modifyitems(config, items):
    import cpu_checkers
    if not cpu_checkers.ref_available():
        skip = pytest.mark.skip(reason="oracle/_ref not built (reference sources absent)")
        for it in items:
            if "ref" in it.keywords:
                it.add_marker(skip)
