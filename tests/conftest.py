import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libfsgpu.so on the device)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref built from the reference sources")


def pytest_collection_modifyitems(config, items):
    from oracle_bridge import build_ref, ref_available

    have_ref = ref_available() or build_ref()
    skip_ref = pytest.mark.skip(reason="oracle/_ref (reference kernels) not built on this machine")
    for it in items:
        if "ref" in it.keywords and not have_ref:
            it.add_marker(skip_ref)
