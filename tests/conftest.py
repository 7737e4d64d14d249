"""Shared test setup: import paths for the binding package and the oracle,
and the `gpu` marker (tests that need an MI355X; run with -m gpu)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "hiprt-path-tracer_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X); parity tests through the C-ABI")


@pytest.fixture(scope="session")
def mipt_mod():
    import mipt

    mipt.load_library()
    return mipt


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def scene1(mipt_mod):
    return mipt_mod.generate_scene(1, 512, 512)


@pytest.fixture(scope="session")
def scene2(mipt_mod):
    return mipt_mod.generate_scene(2, 1920, 1080)
