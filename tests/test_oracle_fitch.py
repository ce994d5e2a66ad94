"""Pins the Fitch oracle (oracle/oracle.c) against the reference's own
known-answer tests and against brute force on tiny trees."""
import itertools
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, caterpillar, random_tree
from oracle import oracle as O

KATS = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))


def test_dna_codes_match_reference_alphabet():
    # test/alphabetTest.ml:12-18
    assert KATS["dna_codes"] == {"A": 1, "C": 2, "G": 4, "T": 8, "-": 16}
    names = {v: k for k, v in KATS["dna_states"].items()}
    s = "".join(names[c] for c in KATS["sequence_kat"]["codes"])
    assert s == KATS["sequence_kat"]["string"]  # test/sequenceTest.ml:6-9


def test_fitch_pair_kat():
    # test/costMatrixTest.ml:83-108: every pair of Alphabet.dna codes
    rows = np.array(KATS["fitch_pairs"], dtype=np.int64)
    x, y = rows[:, 0].astype(np.uint8), rows[:, 1].astype(np.uint8)
    med, cost = O.fitch_median2(x, y)
    assert np.array_equal(med, rows[:, 2].astype(np.uint8))
    per_site = [O.fitch_median2(x[i:i + 1], y[i:i + 1])[1] for i in range(len(x))]
    assert per_site == list(rows[:, 3].astype(float))
    assert cost == rows[:, 3].sum()


def test_fitch_all_iupac_pairs_setwise():
    # all 4-bit sets, checked with Python sets of state indices
    pairs = list(itertools.product(range(1, 16), repeat=2))
    x = np.array([p[0] for p in pairs], np.uint8)
    y = np.array([p[1] for p in pairs], np.uint8)
    med, cost = O.fitch_median2(x, y)
    exp_cost = 0
    for i, (a, b) in enumerate(pairs):
        sa = {k for k in range(4) if a >> k & 1}
        sb = {k for k in range(4) if b >> k & 1}
        m = sa & sb or sa | sb
        exp_cost += 0 if sa & sb else 1
        assert med[i] == sum(1 << k for k in m)
    assert cost == exp_cost


def test_cost_equals_distance_invariant(rng):
    # test/bitvectorTest.ml:59-69: snd (fitch_median_2 a b) = distance a b
    for n in (1, 17, 1000):
        a = rng.integers(1, 16, n).astype(np.uint8)
        b = rng.integers(1, 16, n).astype(np.uint8)
        assert O.fitch_median2(a, b)[1] == O.fitch_distance(a, b)


def test_weighted_median2(rng):
    a = rng.integers(1, 16, 500).astype(np.uint8)
    b = rng.integers(1, 16, 500).astype(np.uint8)
    w = rng.integers(1, 9, 500).astype(np.float64)
    _, c = O.fitch_median2(a, b, w)
    assert c == w[(a & b) == 0].sum()


def _leaf_sets(rng, N, S, ambiguous):
    if ambiguous:
        return rng.integers(1, 16, (N, S)).astype(np.uint8)
    return (1 << rng.integers(0, 4, (N, S))).astype(np.uint8)


@pytest.mark.parametrize("N", [2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("ambiguous", [False, True])
def test_downpass_total_is_brute_force_minimum(rng, N, ambiguous):
    for trial in range(6):
        ch, root = random_tree(rng, N) if trial % 2 == 0 else caterpillar(N)
        leaves = _leaf_sets(rng, N, 40, ambiguous)
        sets, sc, nc, tot = O.fitch_downpass(N, ch, root, leaves)
        for s in range(leaves.shape[1]):
            best, _ = O.fitch_brute(N, ch, root, leaves[:, s])
            assert sc[s] == best
        assert tot == sc.sum() == nc.sum()


@pytest.mark.parametrize("N", [3, 4, 5, 6, 7])
def test_uppass_internal_sets_are_mpr_sets(rng, N):
    # For binary trees, Fitch's final sets at internal nodes are exactly the
    # states used by some most-parsimonious reconstruction (Swofford &
    # Maddison 1987); brute force enumerates all reconstructions.
    for trial in range(8):
        ch, root = random_tree(rng, N)
        for ambiguous in (False, True):
            leaves = _leaf_sets(rng, N, 30, ambiguous)
            sets, *_ = O.fitch_downpass(N, ch, root, leaves)
            fin = O.fitch_uppass(N, ch, root, sets)
            for s in range(leaves.shape[1]):
                _, mpr = O.fitch_brute(N, ch, root, leaves[:, s])
                assert np.array_equal(fin[N:, s], mpr[N:]), (trial, s)
                # root final = optimal root states
                assert fin[root, s] == mpr[root]


def test_uppass_leaf_rule(rng):
    N = 6
    ch, root = random_tree(rng, N)
    leaves = _leaf_sets(rng, N, 200, True)
    sets, *_ = O.fitch_downpass(N, ch, root, leaves)
    fin = O.fitch_uppass(N, ch, root, sets)
    parent = np.full(2 * N - 1, -1)
    for k in range(N - 1):
        parent[ch[2 * k]] = parent[ch[2 * k + 1]] = N + k
    for leaf in range(N):
        fp = fin[parent[leaf]]
        inter = fp & leaves[leaf]
        assert np.array_equal(fin[leaf], np.where(inter != 0, inter, leaves[leaf]))


def test_median3_brute(rng):
    a, b, c = (rng.integers(1, 16, 300).astype(np.uint8) for _ in range(3))
    med, cost = O.fitch_median3(a, b, c)
    tot = 0
    for i in range(300):
        costs = [sum(0 if (x[i] >> s) & 1 else 1 for x in (a, b, c)) for s in range(4)]
        m = min(costs)
        tot += m
        assert med[i] == sum(1 << s for s in range(4) if costs[s] == m)
    assert cost == tot


def test_adjust3_rules():
    # the Fitch (1971) rules, case by case
    pv, pa, pb, fp = (np.array(v, np.uint8) for v in
                      ([0b0011, 0b0011, 0b0001], [0b0011, 0b0111, 0b0001],
                       [0b1011, 0b0011, 0b0010], [0b0001, 0b0100, 0b1100]))
    out = O.fitch_adjust3(pv, pa, pb, fp)
    # site 0: F_p subset of P_v -> F_p
    # site 1: P_v was an intersection -> P_v | (F_p & (P_a | P_b)) = 0011 | 0100
    # site 2: P_v was a union ({A} | {C}), F_p not a subset -> P_v | F_p
    assert list(out) == [0b0001, 0b0111, 0b1101]
