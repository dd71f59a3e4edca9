import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "show-me-dynamic-parallelism_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "data_test_answers.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def data_test_path():
    return os.path.join(GOLDEN, "data_test.gr")


@pytest.fixture(scope="session")
def ctx():
    import smdp
    c = smdp.Context(0)
    yield c
    c.close()


def random_graph(n, m, seed, wlo=1, whi=100, self_loops=True):
    """Small random directed multigraph as (row, col, w) in CSR, duplicates kept."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    if not self_loops:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    w = rng.integers(wlo, whi + 1, len(src))
    order = np.argsort(src, kind="stable")
    src, dst, w = src[order], dst[order], w[order]
    row = np.zeros(n + 1, dtype=np.int32)
    np.add.at(row, src + 1, 1)
    row = np.cumsum(row).astype(np.int32)
    return row, dst.astype(np.int32), w.astype(np.int32)


def symmetrize(row, col):
    n = len(row) - 1
    src = np.repeat(np.arange(n), np.diff(row))
    s2 = np.concatenate([src, col])
    d2 = np.concatenate([col, src])
    order = np.argsort(s2, kind="stable")
    s2, d2 = s2[order], d2[order]
    r = np.zeros(n + 1, dtype=np.int64)
    np.add.at(r, s2 + 1, 1)
    return np.cumsum(r).astype(np.int32), d2.astype(np.int32)
