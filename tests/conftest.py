import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "voronota-ltr_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
DATA = os.path.join(GOLDEN, "data")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


def load_xyzr(name):
    """Last four numeric columns of each line are x y z r (benches/tessellation.rs:6-24)."""
    rows = []
    with open(os.path.join(DATA, name)) as f:
        for line in f:
            parts = line.split()
            if len(parts) < 4:
                continue
            try:
                rows.append([float(v) for v in parts[-4:]])
            except ValueError:
                continue
    return np.array(rows, dtype=np.float64)


RING17 = np.array(
    [
        [0.0, 0.0, 2.0, 1.0], [0.0, 1.0, 0.0, 0.5], [0.382683, 0.92388, 0.0, 0.5], [0.707107, 0.707107, 0.0, 0.5],
        [0.92388, 0.382683, 0.0, 0.5], [1.0, 0.0, 0.0, 0.5], [0.92388, -0.382683, 0.0, 0.5],
        [0.707107, -0.707107, 0.0, 0.5], [0.382683, -0.92388, 0.0, 0.5], [0.0, -1.0, 0.0, 0.5],
        [-0.382683, -0.92388, 0.0, 0.5], [-0.707107, -0.707107, 0.0, 0.5], [-0.92388, -0.382683, 0.0, 0.5],
        [-1.0, 0.0, 0.0, 0.5], [-0.92388, 0.382683, 0.0, 0.5], [-0.707107, 0.707107, 0.0, 0.5],
        [-0.382683, 0.92388, 0.0, 0.5],
    ]
)

CENTRAL8 = np.array(
    [[x, y, z, 1.5] for x, y, z in [
        [0.0, 0.0, 0.0], [2.5, 0.0, 0.0], [1.25, 2.0, 0.0], [1.25, 0.8, 2.0],
        [0.0, 2.5, 1.0], [2.5, 2.5, 1.0], [1.25, -1.5, 1.0], [3.5, 1.25, 1.0]]]
)


def corners_box(lo, hi):
    """PeriodicBox::from_corners (types.rs:302-308) as 9 doubles a|b|c."""
    return [hi[0] - lo[0], 0.0, 0.0, 0.0, hi[1] - lo[1], 0.0, 0.0, 0.0, hi[2] - lo[2]]


@pytest.fixture(scope="session")
def oracle():
    import pyoracle

    pyoracle._load()
    return pyoracle
