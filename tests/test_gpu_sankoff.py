"""GPU Sankoff against the oracle: bit-exact per-site costs and weighted
totals for 2..8 states, symmetric and asymmetric cost matrices, random and
caterpillar trees, ragged site counts; unit costs equal the GPU Fitch pass."""
import os
import sys

import numpy as np
import pytest

from conftest import caterpillar, random_tree
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from phylocaml_amd import device as D

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return D.Context(0)


def _run(ctx, N, ch, root, tcm, leaves, w, stride=None):
    from phylocaml_amd import device as D, sankoff as SK

    S = leaves.shape[1]
    stride = stride or S
    buf = np.zeros((N, stride), np.uint8)
    buf[:, :S] = leaves
    topo = D.Topology(N, ch, root, ctx)
    wd = None if w is None else D.to_device(w, ctx.device)
    sc, tot = SK.downpass(ctx, topo, SK.create(tcm), D.to_device(buf, ctx.device), S, stride, wd)
    ctx.sync()
    return sc.cpu().numpy()[:S], int(tot.cpu().numpy()[0])


@pytest.mark.parametrize("K", [2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("N,S", [(2, 1), (3, 63), (17, 64), (64, 65), (300, 5000)])
def test_sankoff_matches_oracle(ctx, rng, K, N, S):
    for kind in ("random", "caterpillar"):
        _say(f"sankoff K={K} N={N} S={S} {kind}")
        ch, root = random_tree(rng, N) if kind == "random" else caterpillar(N)
        tcm = rng.integers(0, 50, (K, K)).astype(np.int32)
        if kind == "caterpillar":
            tcm = np.minimum(tcm, tcm.T)
        leaves = rng.integers(0, 1 << K, (N, S)).astype(np.uint8)
        w = rng.integers(1, 7, S).astype(np.uint32) if kind == "random" else None
        sc, tot = _run(ctx, N, ch, root, tcm, leaves, w, stride=S + 13)
        osc, otot = O.sankoff(N, ch, root, tcm, leaves, w, threads=min(16, os.cpu_count() or 1))
        assert np.array_equal(sc, osc) and tot == otot


def test_unit_costs_equal_gpu_fitch(ctx, rng):
    from phylocaml_amd import device as D, nonadditive as NA

    N, S = 200, 10007
    ch, root = random_tree(rng, N)
    leaves = rng.integers(1, 16, (N, S)).astype(np.uint8)
    w = rng.integers(1, 4, S).astype(np.uint32)
    sc, tot = _run(ctx, N, ch, root, 1 - np.eye(4, dtype=np.int32), leaves, w)
    topo = D.Topology(N, ch, root, ctx)
    ft = NA.alloc_tree(ctx, topo, S, D.pack_nibbles(leaves), with_final=False)
    NA.downpass(ctx, topo, ft, D.to_device(w, ctx.device))
    ctx.sync()
    assert np.array_equal(sc, ft.site_cost.cpu().numpy()[:S])
    assert tot == int(ft.total.cpu().numpy()[0])


def test_sankoff_argument_errors(ctx, rng):
    from phylocaml_amd import _lib, device as D, sankoff as SK

    ch, root = random_tree(rng, 5)
    topo = D.Topology(5, ch, root, ctx)
    leaves = D.to_device(np.ones((5, 10), np.uint8), ctx.device)
    with pytest.raises(ValueError):
        SK.downpass(ctx, topo, SK.create(np.full((3, 3), 2 ** 29)), leaves, 10, 10)
    t = torch.zeros((81,), dtype=torch.int32, device="cuda:0")
    sc = torch.zeros((10,), dtype=torch.uint32, device="cuda:0")
    tot = torch.zeros((1,), dtype=torch.uint64, device="cuda:0")
    lib = _lib.lib
    assert lib.phy_sankoff_downpass(ctx.handle, topo.handle, 9, D.dptr(t), D.dptr(leaves), 10,
                                    10, None, D.dptr(sc), D.dptr(tot)) == 1
    assert lib.phy_sankoff_downpass(ctx.handle, topo.handle, 4, D.dptr(t), D.dptr(leaves), 5,
                                    10, None, D.dptr(sc), D.dptr(tot)) == 1
    assert lib.phy_sankoff_downpass(ctx.handle, topo.handle, 4, D.dptr(t), D.dptr(leaves), 10,
                                    0, None, D.dptr(sc), D.dptr(tot)) == 0
    ctx.sync()
    assert int(tot.cpu()[0]) == 0
