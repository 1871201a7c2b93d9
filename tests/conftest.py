import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "spatial-cell_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    lib = os.path.join(ROOT, "spatial-cell_amd", "libspatial_cell.so")
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(orc):
        import importlib.util
        spec = importlib.util.spec_from_file_location("orc_build", os.path.join(ROOT, "oracle", "build.py"))
        m = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(m)
        m.build()
    if not os.path.exists(lib):
        import importlib.util
        spec = importlib.util.spec_from_file_location("sc_build", os.path.join(ROOT, "spatial-cell_amd", "build.py"))
        m = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(m)
        m.build()


_ensure_built()
