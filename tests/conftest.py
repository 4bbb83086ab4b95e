"""Shared test setup.

* ``gpu`` marks tests that need an MI355X (run with ``-m gpu``).
* OMP_NUM_THREADS=1 before any OpenMP runtime loads: the oracle's bit-exact
  comparisons with the reference Fortran assume one thread (the reference's
  own reduction order varies with the thread count).
"""
import os
import sys
from pathlib import Path

os.environ.setdefault("OMP_NUM_THREADS", "1")
ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import pytest  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library, loaded; the test fails loudly if it is not built."""
    from yukawamd_amd import _lib
    return _lib.load()

TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))
