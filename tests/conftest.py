import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def random_tree(rng, n_leaves):
    """Random rooted binary tree in the phy_topo layout (random joins)."""
    N = n_leaves
    pool = list(range(N))
    children = np.zeros(2 * (N - 1), np.int32)
    nxt = N
    while len(pool) > 1:
        i, j = rng.choice(len(pool), 2, replace=False)
        a, b = pool[i], pool[j]
        for k in sorted((i, j), reverse=True):
            pool.pop(k)
        children[2 * (nxt - N)] = a
        children[2 * (nxt - N) + 1] = b
        pool.append(nxt)
        nxt += 1
    # relabel so the root is not always the last id
    return children, pool[0]


def caterpillar(n_leaves):
    N = n_leaves
    children = np.zeros(2 * (N - 1), np.int32)
    prev = 0
    for k in range(N - 1):
        children[2 * k] = prev
        children[2 * k + 1] = k + 1
        prev = N + k
    return children, 2 * N - 2


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
