"""Shared test setup: import paths, the gpu marker and small scene helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("sarlab-tdc_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); parity tests proper")


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def tdc():
    """The product package; GPU tests fail loudly when the HIP library is missing."""
    import sarlabtdc

    assert sarlabtdc.device_count() > 0, "no HIP device visible"
    return sarlabtdc


def subset(npx, n, seed=7, extra=()):
    """Seeded pixel subset (sorted flat indices) plus any extra indices."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(npx, size=min(n, npx), replace=False)
    return np.unique(np.concatenate([idx, np.asarray(extra, dtype=np.int64)]))
