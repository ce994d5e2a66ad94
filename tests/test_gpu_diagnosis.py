"""Whole-tree diagnosis glue: node-by-node diagnosis through Make1D (one
median per node, the reference's call pattern) against the whole-tree
kernels and the oracle -- preliminary and final sets, per-node costs,
component cost -- and the same for the likelihood NodeData."""
import os
import sys

import numpy as np
import pytest

from conftest import random_tree
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


def _leaves(ctx, blocks, w):
    from phylocaml_amd import nonadditive as NA

    N = blocks[0].shape[0]
    out = {}
    for i in range(N):
        d = {}
        for code, sets in enumerate(blocks):
            x = NA.of_sets(ctx, sets[i], w)
            x.codes = [code]
            d[code] = x
        out[i] = d
    return out


@pytest.mark.parametrize("N,S", [(3, 17), (40, 1000), (130, 3001)])
def test_fitch_node_by_node_equals_whole_tree_and_oracle(ctx, rng, N, S):
    from phylocaml_amd import device as D, diagnosis as G, nonadditive as NA

    _say(f"diagnosis N={N} S={S}")
    ch, root = random_tree(rng, N)
    sets = rng.integers(1, 16, (N, S)).astype(np.uint8)
    w = rng.integers(1, 4, S).astype(np.uint32)
    osets, osc, onc, otot = O.fitch_downpass(N, ch, root, sets, w)
    ofin = O.fitch_uppass(N, ch, root, osets)
    mk = G.Make1D(ctx)
    pre = G.diagnose_nodes(mk, None, N, ch, root, _leaves(ctx, [sets], w), final=False)
    fin = G.diagnose_nodes(mk, None, N, ch, root, _leaves(ctx, [sets], w), final=True)
    topo = D.Topology(N, ch, root, ctx)
    wt = G.diagnose_tree(ctx, topo, sets, D.to_device(w, ctx.device))
    for pt in (pre, fin, wt):
        assert pt.get_comp_root(0).component_cost == float(otot)
        edge = (int(ch[2 * (root - N)]), int(ch[2 * (root - N) + 1]))
        assert pt.get_root_data(edge).code == root
    for v in range(N, 2 * N - 1):
        a = pre.get_node_data(v)
        b = wt.get_node_data(v)
        assert a.cost == b.cost == float(onc[v - N])
        assert a.height == b.height and a.topot == b.topot
        assert np.array_equal(NA.to_sets(a.prelim[0]), osets[v])
        assert np.array_equal(NA.to_sets(b.prelim[0]), osets[v])
    for v in range(2 * N - 1):
        assert np.array_equal(NA.to_sets(fin.get_node_data(v).final[0]), ofin[v])
        assert np.array_equal(NA.to_sets(wt.get_node_data(v).final[0]), ofin[v])


def test_two_character_blocks(ctx, rng):
    from phylocaml_amd import device as D, diagnosis as G

    N, S1, S2 = 25, 500, 77
    ch, root = random_tree(rng, N)
    s1 = rng.integers(1, 16, (N, S1)).astype(np.uint8)
    s2 = rng.integers(1, 16, (N, S2)).astype(np.uint8)
    mk = G.Make1D(ctx)
    leaves = _leaves(ctx, [s1], None)
    l2 = _leaves(ctx, [s2], None)
    for i in range(N):
        leaves[i][1] = l2[i][0]
        leaves[i][1].codes = [1]
    pt = G.diagnose_nodes(mk, None, N, ch, root, leaves, final=False)
    t1 = O.fitch_downpass(N, ch, root, s1)[3]
    t2 = O.fitch_downpass(N, ch, root, s2)[3]
    assert pt.get_comp_root(0).component_cost == float(t1 + t2)
    r = pt.get_node_data(root)
    assert mk.get_codes(r) == {0, 1} and mk.cardinal(r) == S1 + S2
    only = mk.filter_codes([1], r)
    assert set(only.prelim) == {1}
    assert mk.compare(r, r) == 0


@pytest.mark.parametrize("form,tol", [(0, 1e-12), (1, 1e-6)])
def test_likelihood_node_by_node_equals_whole_tree(ctx, rng, form, tol):
    from phylocaml_amd import device as D, diagnosis as G, likelihood as LK, mlmodel as MM

    N, P = 30, 700
    ch, root = random_tree(rng, N)
    bl = rng.exponential(0.1, 2 * N - 1) + 1e-3
    bl[root] = 0.0
    pi = rng.dirichlet(np.full(4, 5.0))
    model = MM.create(MM.Spec(4, ("GTR", rng.uniform(0.5, 2.0, 5)), ("DiscreteGamma", 4, 0.6),
                              ("Empirical", pi)))
    tips = (1 << rng.integers(0, 4, (N, P))).astype(np.uint8)
    w = rng.integers(1, 5, P).astype(np.uint32)
    dm = LK.DeviceModel(ctx, model)
    wd = D.to_device(w, ctx.device)
    leaves = {i: {0: LK.leaf(ctx, dm, tips[i], bl[i], wd, form)} for i in range(N)}
    _, a, _ = G.diagnose_lnl_nodes(G.Make1D(ctx, LK), dm, N, ch, root, leaves, bl)
    b = G.lnl_tree(ctx, D.Topology(N, ch, root, ctx), dm, bl, tips, w, form)
    ov = O.lnl(N, ch, root, bl, model.priors, model.q, model.rates, model.probs,
               tips.astype(np.uint32), w, threads=min(16, os.cpu_count() or 1))[0]
    assert abs(a - ov) <= tol * abs(ov) and abs(b - ov) <= tol * abs(ov)
    assert abs(a - b) <= tol * abs(ov)
