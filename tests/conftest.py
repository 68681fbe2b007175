"""Shared pytest setup: the repository root on sys.path and the ``gpu`` marker (tests that need a
gfx950 device; the reference declares the same marker, pyproject.toml [tool.pytest.ini_options])."""

import os
import sys

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built libdriftpt.so")
