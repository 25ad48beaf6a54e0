"""Shared test setup: import paths, the `gpu` marker, cached synthetic problems."""
import functools
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("sparse-dft_amd", "oracle"):
    path = os.path.join(REPO, sub)
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs through the HIP C ABI)")


@functools.lru_cache(maxsize=None)
def oracle_problem(config, kfrac=(0.0, 0.0, 0.0)):
    """Synthetic problem whose G-sphere comes from the oracle's enumeration."""
    import pyoracle
    from sdft_amd import synthetic
    return synthetic.config_problem(config, kfrac=kfrac, enumerate_fn=pyoracle.enumerate_g)


@pytest.fixture(scope="session")
def si8():
    return oracle_problem(0)


@pytest.fixture(scope="session")
def si8_oracle(si8):
    import pyoracle
    return pyoracle.Operator(si8)


@pytest.fixture(scope="session")
def hip_ctx():
    from sdft_amd import native
    ctx = native.Context(0)
    yield ctx
    ctx.close()
