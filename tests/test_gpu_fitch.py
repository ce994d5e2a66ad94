"""GPU parity of the Fitch kernels against the oracle: bit-exact sets, per-site
costs, per-node costs and totals, through the C ABI."""
import os
import sys
import time

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


def _sets(rng, n, amb=True):
    if amb:
        return rng.integers(1, 16, n).astype(np.uint8)
    return (1 << rng.integers(0, 4, n)).astype(np.uint8)


@pytest.mark.parametrize("S", [1, 15, 16, 17, 1000, 100003])
@pytest.mark.parametrize("weighted", [False, True])
def test_node_ops_match_oracle(ctx, rng, S, weighted):
    from phylocaml_amd import nonadditive as NA

    _say(f"node ops S={S} weighted={weighted}")
    w = rng.integers(1, 7, S).astype(np.uint32) if weighted else None
    wf = None if w is None else w.astype(np.float64)
    a, b, c = (_sets(rng, S) for _ in range(3))
    x, y, z = (NA.of_sets(ctx, s, w) for s in (a, b, c))
    m = NA.median_2(ctx, None, None, x, y)
    om, oc = O.fitch_median2(a, b, wf)
    assert np.array_equal(NA.to_sets(m), om) and m.cost == oc
    assert NA.distance_1(ctx, None, x, y) == oc
    m3 = NA.median_3(ctx, None, None, x, y, z)
    o3, oc3 = O.fitch_median3(a, b, c, wf)
    assert np.array_equal(NA.to_sets(m3), o3) and m3.cost == oc3
    d2 = NA.distance_2(ctx, None, x, y, z)
    assert d2 == O.fitch_median2(O.fitch_median2(a, b)[0], c, wf)[1]
    fp = _sets(rng, S)
    f = NA.of_sets(ctx, fp, w)
    adj, changed = NA.adjust_3(ctx, None, None, x, y, z, f)
    assert np.array_equal(NA.to_sets(adj), O.fitch_adjust3(a, b, c, fp))
    # in-place reuse of the previous value's block (the `t option` argument)
    prev = NA.of_sets(ctx, _sets(rng, S), w)
    m2 = NA.median_2(ctx, None, prev, x, y)
    assert m2.states.data_ptr() == prev.states.data_ptr()
    assert np.array_equal(NA.to_sets(m2), om)


def _tree_case(ctx, N, ch, root, leaves_u8, weights):
    from phylocaml_amd import device as D, nonadditive as NA

    S = leaves_u8.shape[1]
    topo = D.Topology(N, ch, root, ctx)
    ft = NA.alloc_tree(ctx, topo, S, D.pack_nibbles(leaves_u8), with_final=True,
                       with_node_cost=True)
    wd = None if weights is None else D.to_device(weights, ctx.device)
    NA.downpass(ctx, topo, ft, wd)
    NA.uppass(ctx, topo, ft)
    ctx.sync()
    osets, osc, onc, otot = O.fitch_downpass(N, ch, root, leaves_u8, weights, threads=8)
    ofin = O.fitch_uppass(N, ch, root, osets)
    got_sets = ft.sets.cpu().numpy()
    assert np.array_equal(got_sets[N:], D.pack_nibbles(osets[N:])), "prelim sets differ"
    assert np.array_equal(ft.final.cpu().numpy(), D.pack_nibbles(ofin)), "final sets differ"
    assert np.array_equal(ft.site_cost.cpu().numpy()[:S], osc), "site costs differ"
    assert np.array_equal(ft.node_cost.cpu().numpy(), onc), "node costs differ"
    assert int(ft.total.cpu().numpy()[0]) == otot


@pytest.mark.parametrize("N", [2, 3, 17, 128, 1024])
@pytest.mark.parametrize("S", [1, 17, 5000])
def test_tree_passes_match_oracle(ctx, rng, N, S):
    for kind in ("random", "caterpillar"):
        _say(f"tree passes N={N} S={S} {kind}")
        ch, root = random_tree(rng, N) if kind == "random" else caterpillar(N)
        leaves = rng.integers(1, 16, (N, S)).astype(np.uint8)
        w = rng.integers(1, 5, S).astype(np.uint32) if kind == "random" else None
        _tree_case(ctx, N, ch, root, leaves, w)


@pytest.mark.parametrize("nt", ["0", "1"])
@pytest.mark.parametrize("words", ["1", "2", "4"])
@pytest.mark.parametrize("batch", ["4", "8", "16", "32"])
@pytest.mark.parametrize("split", ["0", "1"])
def test_prefetch_depth_and_split_forms(ctx, rng, monkeypatch, batch, split, words, nt):
    # the whole-tree and the split (parts side by side, then the top) forms,
    # at every operand prefetch depth and lane width, with cached or
    # non-temporal row accesses, give the oracle's bits
    monkeypatch.setenv("PHY_FITCH_BATCH", batch)
    monkeypatch.setenv("PHY_FITCH_SPLIT", split)
    monkeypatch.setenv("PHY_FITCH_WORDS", words)
    monkeypatch.setenv("PHY_FITCH_NT", nt)
    N, S = 300, 3001
    _say(f"fitch forms batch={batch} split={split} words={words} nt={nt}")
    ch, root = random_tree(rng, N)
    leaves = rng.integers(1, 16, (N, S)).astype(np.uint8)
    w = rng.integers(1, 5, S).astype(np.uint32)
    if batch == "32" and words == "4":
        # (2 x 32 + depth) sets of 16 B per thread exceed one block's LDS even
        # at 64 threads: the launch is refused with PHY_ERANGE
        from phylocaml_amd._lib import PhyError

        with pytest.raises(PhyError) as e:
            _tree_case(ctx, N, ch, root, leaves, w)
        assert e.value.code == 5
        return
    _tree_case(ctx, N, ch, root, leaves, w)


@pytest.mark.parametrize("N,narrow", [(1000, "1"), (1000, "0"), (8192, "1")])
def test_split_part_counts_as_bytes_or_words(ctx, rng, monkeypatch, N, narrow):
    # the parts hand their per-site counts to the top as bytes when every part
    # has at most 255 internal nodes (1,000 leaves: parts of <= 103), else as
    # 32-bit words (8,192 leaves: parts of up to ~1,500); PHY_FITCH_NARROW=0
    # forces words
    monkeypatch.setenv("PHY_FITCH_NARROW", narrow)
    S = 517
    _say(f"fitch split counts N={N} narrow={narrow}")
    ch, root = random_tree(rng, N)
    leaves = rng.integers(1, 16, (N, S)).astype(np.uint8)
    w = rng.integers(1, 5, S).astype(np.uint32)
    _tree_case(ctx, N, ch, root, leaves, w)


def test_1024_taxa_1e6_sites_bit_exact(ctx):
    """1,024-taxon Yule tree, 10^6 IUPAC sites (the parsimony benchmark
    workload): every preliminary and final set, every site cost and the total
    against the oracle."""
    from phylocaml_amd import device as D, nonadditive as NA, synth

    N, S = 1024, 1_000_000
    _say("1,024 taxa x 10^6 sites: generating")
    tree = synth.yule(N, 1, 0.1)
    lv = synth.fitch_leaves(N, S, 1)
    topo = D.Topology(N, tree.children, tree.root, ctx)
    ft = NA.alloc_tree(ctx, topo, S, lv, with_final=True, with_node_cost=True)
    NA.downpass(ctx, topo, ft)
    NA.uppass(ctx, topo, ft)
    ctx.sync()
    _say("1,024 taxa x 10^6 sites: oracle")
    t0 = time.time()
    leaves = D.unpack_nibbles(lv, S)
    osets, osc, onc, otot = O.fitch_downpass(N, tree.children, tree.root, leaves,
                                             threads=min(16, os.cpu_count() or 1))
    _say(f"oracle downpass {time.time() - t0:.1f}s")
    assert int(ft.total.cpu().numpy()[0]) == otot
    assert np.array_equal(ft.site_cost.cpu().numpy(), osc)
    assert np.array_equal(ft.node_cost.cpu().numpy(), onc)
    assert int(onc.sum()) == otot
    assert np.array_equal(ft.sets.cpu().numpy()[N:], D.pack_nibbles(osets[N:]))
    ofin = O.fitch_uppass(N, tree.children, tree.root, osets)
    assert np.array_equal(ft.final.cpu().numpy(), D.pack_nibbles(ofin))
    _say(f"1,024 taxa x 10^6 sites: total cost {otot}")


def test_empty_and_bad_arguments(ctx):
    from phylocaml_amd import _lib, device as D

    lib = _lib.lib
    buf = torch.zeros(4, dtype=torch.uint64, device="cuda:0")
    cost = torch.ones(1, dtype=torch.uint64, device="cuda:0")
    # zero sites: cost 0, no launch
    assert lib.phy_fitch_median2(ctx.handle, D.dptr(buf), D.dptr(buf), D.dptr(buf), 0, None,
                                 D.dptr(cost)) == 0
    ctx.sync()
    assert int(cost.cpu()[0]) == 0
    assert lib.phy_fitch_median2(ctx.handle, D.dptr(buf), D.dptr(buf), D.dptr(buf), -1, None,
                                 None) == 1
    topo = D.Topology(3, np.array([0, 1, 2, 3], np.int32), 4, ctx)
    # stride smaller than the block
    assert lib.phy_fitch_downpass(ctx.handle, topo.handle, D.dptr(buf), 1, 100, None,
                                  D.dptr(buf), None, D.dptr(cost)) == 1
    # host-only topology cannot run on the device
    host_topo = D.Topology(3, np.array([0, 1, 2, 3], np.int32), 4)
    assert lib.phy_fitch_uppass(ctx.handle, host_topo.handle, D.dptr(buf), D.dptr(buf), 1,
                                1) == 1
