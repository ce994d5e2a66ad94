"""Edge cases of the GPU paths against closed forms and the oracle: one
pattern, all-missing tips (lnL 0), zero weights, very short and very long
branches, a perfectly balanced 1,024-taxon tree (deepest stack) and a
2,048-taxon caterpillar, in every pruning form; the same shapes for Fitch
and Sankoff."""
import os
import sys

import numpy as np
import pytest

from conftest import caterpillar, random_tree
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FORMS = [(0, 1e-10), (2, 1e-10), (1, 1e-6), (3, 1e-6)]  # f64, f64 fused, f32, f32 fused


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from phylocaml_amd import device as D

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return D.Context(0)


def balanced(N):
    """Perfectly balanced tree of N = 2^k leaves (stack depth k)."""
    ch, nxt, layer = [], N, list(range(N))
    while len(layer) > 1:
        nl = []
        for i in range(0, len(layer), 2):
            ch += [layer[i], layer[i + 1]]
            nl.append(nxt)
            nxt += 1
        layer = nl
    return np.array(ch, np.int32), layer[0]


def _model(R=4):
    from phylocaml_amd import mlmodel as MM

    return MM.create(MM.Spec(4, ("GTR", [1.2, 3.1, 0.7, 0.9, 2.5]), ("DiscreteGamma", R, 0.6),
                             ("Empirical", [0.3, 0.2, 0.25, 0.25])))


def _lnl(ctx, N, ch, root, bl, model, tips, w, form):
    from phylocaml_amd import device as D, likelihood as LK

    P = tips.shape[1]
    topo = D.Topology(N, ch, root, ctx)
    ev = LK.Evaluator(ctx, topo, LK.DeviceModel(ctx, model), P, form)
    ev.run(D.to_device(D.pack_nibbles(tips), ctx.device), D.words(P),
           D.to_device(np.asarray(w, np.uint32), ctx.device),
           D.to_device(np.asarray(bl, np.float64), ctx.device), site_lnl=True)
    ctx.sync()
    return float(ev.lnl.cpu()[0]), ev.site_lnl.cpu().numpy()


def _oracle(N, ch, root, bl, model, tips, w):
    return O.lnl(N, ch, root, bl, model.priors, model.q, model.rates, model.probs,
                 tips.astype(np.uint32), w, threads=min(16, os.cpu_count() or 1))


@pytest.mark.parametrize("form,tol", FORMS)
def test_one_pattern_and_missing_tips(ctx, rng, form, tol):
    N = 20
    ch, root = random_tree(rng, N)
    bl = rng.exponential(0.1, 2 * N - 1)
    bl[root] = 0.0
    m = _model()
    tips = (1 << rng.integers(0, 4, (N, 1))).astype(np.uint8)
    w = np.array([3], np.uint32)
    v, _ = _lnl(ctx, N, ch, root, bl, m, tips, w, form)
    ov, _ = _oracle(N, ch, root, bl, m, tips, w)
    assert abs(v - ov) <= tol * abs(ov)
    # every tip missing: L = sum_i pi_i = 1 at every pattern, lnL = 0, up to
    # the rounding of the P row sums (FP32: ~6e-8 each, over 2N - 2 products)
    miss = np.full((N, 100), 15, np.uint8)
    v, site = _lnl(ctx, N, ch, root, bl, m, miss, np.ones(100, np.uint32), form)
    atol = 1e-12 if form in (0, 2) else (2 * N - 2) * 1.2e-7
    assert np.all(np.abs(site) <= atol) and abs(v) <= 100 * atol


@pytest.mark.parametrize("form,tol", FORMS)
def test_zero_weights_and_extreme_branches(ctx, rng, form, tol):
    N, P = 33, 500
    ch, root = random_tree(rng, N)
    m = _model()
    tips = (1 << rng.integers(0, 4, (N, P))).astype(np.uint8)
    w = rng.integers(0, 3, P).astype(np.uint32)
    w[:50] = 0
    for scale in (1e-4, 3.0, 20.0):
        bl = rng.uniform(0.5, 1.5, 2 * N - 1) * scale
        bl[root] = 0.0
        v, site = _lnl(ctx, N, ch, root, bl, m, tips, w, form)
        ov, osite = _oracle(N, ch, root, bl, m, tips, w)
        assert abs(v - ov) <= tol * max(abs(ov), 1.0), (scale, v, ov)
        assert np.allclose(site, osite, rtol=10 * tol, atol=10 * tol)
    # t < 1e-10 gives P = I (lib/mlModel.ml:332-341): with every leaf in the
    # same state s at a pattern, L = pi_s exactly
    bl = np.full(2 * N - 1, 1e-12)
    bl[root] = 0.0
    st = rng.integers(0, 4, P)
    same = np.tile((1 << st).astype(np.uint8), (N, 1))
    v, site = _lnl(ctx, N, ch, root, bl, m, same, w, form)
    want = np.log(m.priors[st])
    assert np.allclose(site, want, rtol=tol, atol=tol)
    assert abs(v - float((w * want).sum())) <= tol * abs(v)


@pytest.mark.parametrize("form,tol", FORMS)
def test_deep_balanced_and_long_caterpillar(ctx, rng, form, tol):
    m = _model()
    for name, (N, (ch, root)) in (("balanced 1024", (1024, balanced(1024))),
                                  ("caterpillar 2048", (2048, caterpillar(2048)))):
        _say(f"edge pruning {name} form={form}")
        P = 300
        bl = rng.exponential(0.05, 2 * N - 1)
        bl[root] = 0.0
        tips = (1 << rng.integers(0, 4, (N, P))).astype(np.uint8)
        w = rng.integers(1, 4, P).astype(np.uint32)
        v, _ = _lnl(ctx, N, ch, root, bl, m, tips, w, form)
        ov, _ = _oracle(N, ch, root, bl, m, tips, w)
        assert abs(v - ov) <= tol * abs(ov), (name, v, ov)


def test_fitch_and_sankoff_on_extreme_shapes(ctx, rng):
    from phylocaml_amd import device as D, nonadditive as NA, sankoff as SK

    for name, (N, (ch, root)) in (("balanced 1024", (1024, balanced(1024))),
                                  ("caterpillar 2048", (2048, caterpillar(2048)))):
        _say(f"edge fitch/sankoff {name}")
        S = 777
        leaves = rng.integers(1, 16, (N, S)).astype(np.uint8)
        topo = D.Topology(N, ch, root, ctx)
        ft = NA.alloc_tree(ctx, topo, S, D.pack_nibbles(leaves), with_final=True)
        NA.downpass(ctx, topo, ft)
        NA.uppass(ctx, topo, ft)
        osets, osc, _, otot = O.fitch_downpass(N, ch, root, leaves, threads=8)
        ctx.sync()
        assert int(ft.total.cpu()[0]) == otot
        assert np.array_equal(ft.final.cpu().numpy(),
                              D.pack_nibbles(O.fitch_uppass(N, ch, root, osets)))
        tcm = rng.integers(0, 30, (4, 4)).astype(np.int32)
        sc, tot = SK.downpass(ctx, topo, SK.create(tcm), D.to_device(leaves, ctx.device), S, S)
        osc2, otot2 = O.sankoff(N, ch, root, tcm, leaves, threads=8)
        ctx.sync()
        assert int(tot.cpu()[0]) == otot2
        assert np.array_equal(sc.cpu().numpy()[:S], osc2)
