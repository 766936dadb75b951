"""Shared test setup: the `gpu` marker, import paths and small fixtures.

Tests marked `gpu` need an AMD GPU (they call the HIP library through the C ABI and compare with the CPU
oracle); everything else runs on the CPU.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "taral-lp_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def packing_10000():
    from taral_pdhg import gen
    return gen.packing_mps(10000)


@pytest.fixture(scope="session")
def transport_100x100():
    from taral_pdhg import gen
    return gen.transport_mps(100, 100)


@pytest.fixture(scope="session")
def cfg0():
    """The 2,000 x 5,000 planted LP (10 nonzeros per column, seed 0)."""
    from taral_pdhg import gen
    return gen.planted_lp(2000, 5000, 10, seed=0, name="CFG0")
