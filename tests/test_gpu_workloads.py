"""The batched 20-state and learned-SPR paths at their full benchmark sizes.

* 20-state batch: 2,048 trees, taxa ~ U{16..256}, patterns ~ U{1,000..
  50,000} (GTR-20 + Gamma-4, the benchmark workload), all in one persistent
  launch.  A stratified sample of 27 trees -- per taxon bucket the largest-
  pattern tree, the deepest tree and random ones, plus the batch's largest
  tree overall -- against the FP64 oracle within the BF16 tolerance 2e-3;
  every tree's lnL finite and <= 0; a second run bit-identical.
* Learned SPR: 256 taxa x 50,000 patterns.  One round's complete list of
  rooted SPR moves scored exactly in FP64 and FP32-scaled; 16 sampled
  moves against the oracle's lnL of the moved tree (FP64 1e-10, FP32 1e-6
  relative); FP32 within 1e-6 of FP64 on every move; and the MLP, trained on
  another instance's moves, puts an improving move in its top 64 when the
  exhaustive round has one."""
import os
import sys
import time

import numpy as np
import pytest

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


def _stack_depth(N, ch, root):
    """Sethi-Ullman stack depth of the post-order (the tree's 'depth' for the
    kernels: the number of partials waiting at once)."""
    ch = np.asarray(ch).reshape(-1, 2)
    memo = {}

    def need(v):
        if v < N:
            return 0
        if v in memo:
            return memo[v]
        a, b = need(ch[v - N][0]), need(ch[v - N][1])
        memo[v] = max(a, b) if a != b else a + 1
        return memo[v]

    sys.setrecursionlimit(10000)
    return need(root)


def test_aa_batch_2048_trees_one_launch(ctx):
    from phylocaml_amd import aa

    t0 = time.time()
    trees = aa.synth_workload(2048, seed=3)
    _say(f"20-state batch: 2,048 trees generated in {time.time() - t0:.1f}s")
    b = aa.AABatch(ctx, trees, waves=4)
    b.run()
    ctx.sync()
    got = b.lnl.cpu().numpy()
    assert np.all(np.isfinite(got)) and np.all(got <= 0.0)
    b.run()
    ctx.sync()
    assert np.array_equal(b.lnl.cpu().numpy(), got), "second run differs"
    rng = np.random.default_rng(0)
    pick = {int(np.argmax([t.n_patterns * t.n_leaves for t in trees]))}
    for lo, hi in aa.BUCKETS:
        idx = [i for i, t in enumerate(trees) if lo <= t.n_leaves < hi]
        if not idx:
            continue
        pick.add(max(idx, key=lambda i: trees[i].n_patterns))
        pick.add(max(idx, key=lambda i: _stack_depth(trees[i].n_leaves, trees[i].children,
                                                      trees[i].root)))
        pick |= {int(i) for i in rng.choice(idx, 7, replace=False)}
    assert len(pick) >= 24
    worst = 0.0
    for i in sorted(pick):
        t = trees[i]
        ov, _ = O.lnl(t.n_leaves, t.children, t.root, t.blen, t.pi, t.q, t.rates, t.probs,
                      aa.tips_bitmask(t.tips), t.weights, threads=min(16, os.cpu_count() or 1))
        rel = abs(got[i] - ov) / abs(ov)
        worst = max(worst, rel)
        _say(f"tree {i}: N={t.n_leaves} P={t.n_patterns} oracle {ov:.4f} gpu {got[i]:.4f} "
             f"rel {rel:.2e}")
        assert rel <= 2e-3, (i, got[i], ov)
    _say(f"20-state batch: {len(pick)} trees checked, worst relative error {worst:.2e}")


def test_spr_round_256_taxa_50k_patterns(ctx):
    from phylocaml_amd import device as D, spr, spr_search as SS

    N, P = 256, 50_000
    t0 = time.time()
    inst = SS.make_instance(7, N, P, 10)
    c = SS.Context(ctx, inst)
    mv = spr.moves(N, inst.children, inst.root)
    _say(f"SPR round: {len(mv)} moves of a {N}-taxon tree, {P} patterns "
         f"({time.time() - t0:.1f}s to set up)")
    base = c.lnl(inst.children, inst.root, inst.blen)
    t0 = time.time()
    s64 = c.exact(inst.children, inst.root, inst.blen, mv, fp64=True)
    t64 = time.time() - t0
    t0 = time.time()
    s32 = c.exact(inst.children, inst.root, inst.blen, mv, fp64=False)
    t32 = time.time() - t0
    _say(f"SPR round scored: FP64 {t64:.1f}s, FP32 {t32:.1f}s")
    assert np.all(np.isfinite(s64)) and np.all(np.isfinite(s32))
    rel32 = np.abs(s32 - s64) / np.abs(s64)
    _say(f"FP32 vs FP64 over every move: max relative {rel32.max():.2e}")
    assert rel32.max() <= 1e-6
    # the whole round went through the incremental scorer (moves per pruned
    # subtree >> 4); 1,024 of its moves through the per-candidate scorer
    sample = np.arange(0, len(mv), len(mv) // 1024)[:1024]
    p64 = c.exact(inst.children, inst.root, inst.blen, mv[sample], fp64=True, incremental=False)
    rel = np.abs(p64 - s64[sample]) / np.abs(p64)
    _say(f"incremental vs per-candidate FP64 on 1,024 moves: max relative {rel.max():.2e}")
    assert rel.max() <= 1e-10
    # sampled moves against the oracle (the best move, the worst, 14 random)
    rng = np.random.default_rng(1)
    picks = [int(np.argmax(s64)), int(np.argmin(s64))] + \
        [int(i) for i in rng.choice(len(mv), 14, replace=False)]
    tips = D.unpack_nibbles(inst.tips_np, P)
    for i in picks:
        ch, r, bl = spr.apply(N, inst.children, inst.root, inst.blen, *mv[i])
        ov, _ = O.lnl(N, ch, r, bl, inst.model.priors, inst.model.q, inst.model.rates,
                      inst.model.probs, tips.astype(np.uint32), inst.weights,
                      threads=min(16, os.cpu_count() or 1))
        assert abs(s64[i] - ov) <= 1e-10 * abs(ov), (i, s64[i], ov)
        assert abs(s32[i] - ov) <= 1e-6 * abs(ov), (i, s32[i], ov)
    improving = s64 > base + 1e-6 * abs(base)
    _say(f"{int(improving.sum())} improving moves; best dlnL {s64.max() - base:.3f}")
    # the learned proposal, trained on another instance's exact scores
    X, T = SS.training_pairs(ctx, [101], N, P, 10)
    net, norm, losses = SS.train(ctx, X, T)
    pred = net.forward(D.to_device(norm(c.features(inst.children, inst.root, inst.blen, mv))))
    pred = pred.cpu().numpy()
    top = np.argpartition(-pred, 64)[:64]
    _say(f"MLP trained on {len(X)} pairs (loss {losses[-1]:.4f}); improving moves in the "
         f"top 64: {int(improving[top].sum())}")
    if improving.any():
        assert improving[top].any()
