"""Shared test set-up: import paths, the `gpu` marker, small scene builders."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("gpu-sims_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def dfsph_mod():
    import dfsph

    dfsph.lib()
    return dfsph


def dam_break_small(n=16, r=0.01, box=(2.0, 1.0, 1.0), seed=1):
    """C1-like scene: an n^3 jittered block at the min corner with 2r clearance."""
    import dfsph

    x = dfsph.scene_block(n, n, n, r, (2 * r, 2 * r, 2 * r), seed)
    return x, (0.0, 0.0, 0.0), box


def rel_err(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = float(np.abs(b).max())
    if scale == 0.0:
        return float(np.abs(a).max())
    return float(np.abs(a - b).max() / scale)
