"""Shared test setup: the `gpu` marker, import paths, seeded workloads."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def workload1():
    from surface_match_amd import synth
    return synth.workload(1)


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))
