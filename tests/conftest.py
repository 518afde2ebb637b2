"""Test configuration: markers and import paths."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    # Every plan the tests create fills its device output with a non-canonical
    # NaN before each host-buffer launch: an output that no kernel writes
    # keeps that pattern, which no filter can produce, instead of inheriting
    # a correct value from an earlier launch into the same device buffer.
    try:
        from ndfiltrs_amd import _core
    except ImportError:  # library not built: the CPU-only tests still run
        return
    from tests._cases import POISON32, POISON64

    _core.set_output_poison((POISON32, POISON64))
