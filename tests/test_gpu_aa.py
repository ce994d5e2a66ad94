"""GPU checks of the batched 20-state BF16-MFMA pruning against the FP64
oracle.

BF16 operands carry an 8-bit significand, so the per-tree lnL is compared
with a relative tolerance of 2e-3 (the measured error is reported by
bench.py); every tree in a batch of uneven trees must meet it, which also
checks the persistent work queue and the per-tree bookkeeping."""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from phylocaml_amd import device as D

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return D.Context(0)


def _oracle(t):
    from phylocaml_amd import aa

    v, _ = O.lnl(t.n_leaves, t.children, t.root, t.blen, t.pi, t.q, t.rates, t.probs,
                 aa.tips_bitmask(t.tips), t.weights, threads=min(16, os.cpu_count() or 1))
    return v


@pytest.mark.parametrize("waves", [1, 2, 4, 8])
def test_batch_matches_oracle(ctx, waves):
    from phylocaml_amd import aa

    print(f"aa batch waves={waves}", flush=True)
    rng = np.random.default_rng(7)
    trees = []
    for k in range(9):
        N = int(rng.integers(2, 70))
        P = int(rng.integers(1, 900))
        trees.append(aa.synth_tree(1000 + k, N, P, n_sim=300))
    # ambiguous/missing tips
    trees[3].tips[:, ::7] = 255
    b = aa.AABatch(ctx, trees, waves=waves)
    b.run()
    ctx.sync()
    got = b.lnl.cpu().numpy()
    for k, t in enumerate(trees):
        ov = _oracle(t)
        rel = abs(got[k] - ov) / abs(ov)
        print(f"tree {k}: N={t.n_leaves} P={t.n_patterns} oracle {ov:.6f} gpu {got[k]:.6f} "
              f"rel {rel:.2e}", flush=True)
        assert rel <= 2e-3, (k, got[k], ov)
    # a second run on the same batch gives the same answer (queue reset)
    b.run()
    ctx.sync()
    assert np.array_equal(b.lnl.cpu().numpy(), got)


@pytest.mark.parametrize("waves", [4, 8])
def test_register_stack_levels_give_the_same_bits(ctx, monkeypatch, waves):
    # the bottom stack levels held in registers carry the same BF16 operands
    # as the LDS rows: identical results with and without them
    from phylocaml_amd import aa

    rng = np.random.default_rng(11)
    trees = [aa.synth_tree(2000 + k, int(rng.integers(40, 140)), int(rng.integers(100, 700)),
                           n_sim=200) for k in range(6)]
    out = {}
    for reg in ("0", "3"):
        monkeypatch.setenv("PHY_AA_REG", reg)
        print(f"aa register levels {reg} waves={waves}", flush=True)
        b = aa.AABatch(ctx, trees, waves=waves)
        b.run()
        ctx.sync()
        out[reg] = b.lnl.cpu().numpy()
    assert np.array_equal(out["0"], out["3"])
    ov = _oracle(trees[0])
    assert abs(out["3"][0] - ov) <= 2e-3 * abs(ov)


def test_tiny_exact_cases(ctx):
    # constant tips and a two-taxon tree: L = sum_i pi_i P_ij(t1 + t2)
    from phylocaml_amd import aa

    t = aa.synth_tree(5, 2, 40, n_sim=40)
    b = aa.AABatch(ctx, [t], waves=1)
    b.run()
    ctx.sync()
    ov = _oracle(t)
    assert abs(b.lnl.cpu().numpy()[0] - ov) <= 2e-3 * abs(ov)


@pytest.mark.parametrize("waves", [1, 4, 8])
def test_large_trees_cross_leaf_id_chunks(ctx, waves):
    # the kernel holds the leaf-id stream 128 ids at a time and advances it by
    # 64: trees of 129-300 leaves cross several such refills
    from phylocaml_amd import aa

    print(f"aa large trees waves={waves}", flush=True)
    trees = [aa.synth_tree(3000 + k, N, P, n_sim=150)
             for k, (N, P) in enumerate([(129, 70), (200, 130), (300, 40), (257, 17)])]
    trees[1].tips[:, ::5] = 255
    b = aa.AABatch(ctx, trees, waves=waves)
    b.run()
    ctx.sync()
    got = b.lnl.cpu().numpy()
    for k, t in enumerate(trees):
        ov = _oracle(t)
        rel = abs(got[k] - ov) / abs(ov)
        print(f"tree {k}: N={t.n_leaves} P={t.n_patterns} rel {rel:.2e}", flush=True)
        assert rel <= 2e-3, (k, got[k], ov)


@pytest.mark.parametrize("R", [1, 2])
@pytest.mark.parametrize("waves", [1, 4, 8])
def test_one_and_two_rate_classes(ctx, R, waves):
    # the R = 1 and R = 2 kernels (the other tests use four Gamma classes)
    from dataclasses import replace

    from phylocaml_amd import aa

    print(f"aa rate classes R={R} waves={waves}", flush=True)
    rates = {1: [1.0], 2: [0.35, 1.65]}[R]
    rng = np.random.default_rng(31 + R)
    trees = []
    for k in range(6):
        t = aa.synth_tree(5000 + 10 * R + k, int(rng.integers(2, 150)), int(rng.integers(1, 600)),
                          n_sim=120)
        trees.append(replace(t, rates=np.array(rates), probs=np.full(R, 1.0 / R)))
    b = aa.AABatch(ctx, trees, waves=waves)
    b.run()
    ctx.sync()
    got = b.lnl.cpu().numpy()
    for k, t in enumerate(trees):
        ov = _oracle(t)
        assert abs(got[k] - ov) <= 2e-3 * abs(ov), (k, got[k], ov)
