"""CPU tests of libphylo_hip.so: it loads, exports every symbol the headers
declare, validates topologies, compiles correct traversal programs (checked by
interpreting them on the host exactly as the kernels do), and generates
deterministic synthetic workloads.  No GPU calls."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, caterpillar, random_tree
from oracle import oracle as O
from phylocaml_amd import _lib, device as D, synth


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(phy_\w+)\s*\(", txt))
    return names


def test_every_declared_symbol_is_exported():
    declared = _declared()
    assert len(declared) > 40
    missing = [n for n in declared if not hasattr(_lib.lib, n)]
    assert not missing, missing
    unbound = declared - set(_lib.SIGNATURES)
    assert not unbound, unbound


def test_words():
    assert [_lib.lib.phy_words(n) for n in (0, 1, 15, 16, 17, 1000000)] == [0, 1, 1, 1, 2, 62500]


def test_pack_roundtrip(rng):
    for S in (1, 15, 16, 17, 100, 1001):
        s = rng.integers(0, 16, (3, S)).astype(np.uint8)
        w = D.pack_nibbles(s)
        assert w.shape == (3, (S + 15) // 16)
        assert np.array_equal(D.unpack_nibbles(w, S), s)
    # site 0 -> bits 0..3 of word 0; site 17 -> bits 4..7 of word 1
    s = np.zeros(18, np.uint8)
    s[0], s[17] = 5, 9
    w = D.pack_nibbles(s)
    assert int(w[0]) == 5 and int(w[1]) == 9 << 4


@pytest.mark.parametrize("bad", ["two_parents", "self", "range", "root_child", "cycle"])
def test_topology_rejects_malformed(bad):
    N = 4
    ch, root = caterpillar(N)
    ch = ch.copy()
    if bad == "two_parents":
        ch[1] = ch[0]
    elif bad == "self":
        ch[0] = N
    elif bad == "range":
        ch[0] = 2 * N - 1
    elif bad == "root_child":
        root = 0
    elif bad == "cycle":
        ch[2 * 0] = N + 2  # node N's child is the root
    with pytest.raises(_lib.PhyError):
        D.Topology(N, ch, root)


def _interp_down(N, dprog, dfifo, leaves):
    """Host copy of k_fitch_down's program semantics (acc + stack + fifo)."""
    sets = {}
    acc, stack, f = None, [], 0

    def nxt():
        nonlocal f
        v = leaves[dfifo[f]]
        f += 1
        return v

    for v, A, B, fl in dprog:
        if fl & 16:
            stack.append(acc)
        sa, sb = fl & 3, (fl >> 2) & 3
        x = nxt() if sa == 0 else (acc if sa == 1 else stack.pop())
        y = nxt() if sb == 0 else acc
        t = x & y
        acc = np.where(t != 0, t, x | y).astype(np.uint8)
        sets[v] = acc
    assert not stack and f == N
    return sets


def _interp_up(N, uprog, ufifo, sets, fin=None, parent_of_root=None, expect=None):
    fin = {} if fin is None else fin
    accP = accF = None
    stack, f = [], 0

    def nxt():
        nonlocal f
        v = sets[ufifo[f]]
        f += 1
        return v

    def final_rule(pv, pa, pb, fp):
        sub = (fp & ~pv & 15) == 0
        inter = (pa & pb) != 0
        r = np.where(inter, pv | (fp & (pa | pb)), pv | fp)
        return np.where(sub, fp, r).astype(np.uint8)

    for v, a, b, fl in uprog:
        src = fl & 3
        if src == 0:
            pv = nxt()
            fp = None
        elif src == 3:  # a part's root: the parent's final from memory
            pv, fp = nxt(), fin[parent_of_root]
        elif src == 1:
            pv, fp = accP, accF
        else:
            pv, fp = stack.pop()
        pa, pb = nxt(), nxt()
        fv = pv if src == 0 else final_rule(pv, pa, pb, fp)
        fin[v] = fv
        for c, pc, leaf_flag in ((a, pa, 4), (b, pb, 8)):
            if fl & leaf_flag:
                i = fv & pc
                fin[c] = np.where(i != 0, i, pc).astype(np.uint8)
        if fl & 32:
            stack.append(((pa if fl & 16 else pb), fv))
        accP, accF = (pb if fl & 16 else pa), fv
    assert not stack and f == (2 * N - 1 if expect is None else expect)
    return fin


@pytest.mark.parametrize("N", [2, 3, 5, 16, 57, 300])
def test_programs_reproduce_oracle(rng, N):
    for kind in ("random", "caterpillar", "yule"):
        if kind == "random":
            ch, root = random_tree(rng, N)
        elif kind == "caterpillar":
            ch, root = caterpillar(N)
        else:
            t = synth.yule(N, int(rng.integers(1 << 30)), 0.1)
            ch, root = t.children, t.root
        topo = D.Topology(N, ch, root)
        dprog, dfifo, uprog, ufifo = topo.programs()
        dd, ud, parent = topo.info()
        # every internal node exactly once, every leaf consumed exactly once
        assert sorted(dprog[:, 0]) == list(range(N, 2 * N - 1))
        assert sorted(dfifo) == list(range(N))
        assert sorted(uprog[:, 0]) == list(range(N, 2 * N - 1))
        assert sorted(ufifo) == list(range(2 * N - 1))
        assert dd <= int(np.ceil(np.log2(N))) + 1 and ud <= int(np.ceil(np.log2(N))) + 1
        leaves = rng.integers(1, 16, (N, 37)).astype(np.uint8)
        osets, *_ = O.fitch_downpass(N, ch, root, leaves)
        got = _interp_down(N, dprog, dfifo, leaves)
        for v in range(N, 2 * N - 1):
            assert np.array_equal(got[v], osets[v])
        ofin = O.fitch_uppass(N, ch, root, osets)
        allsets = {v: osets[v] for v in range(2 * N - 1)}
        fin = _interp_up(N, uprog, ufifo, allsets)
        for v in range(2 * N - 1):
            assert np.array_equal(fin[v], ofin[v]), (kind, v)


@pytest.mark.parametrize("N", [64, 127, 128, 300, 1024])
def test_split_programs_reproduce_oracle(rng, N):
    # parts side by side, then the top (down); the top, then the parts (up):
    # every internal node once, and the oracle's sets
    for kind in ("random", "yule", "caterpillar"):
        if kind == "random":
            ch, root = random_tree(rng, N)
        elif kind == "caterpillar":
            ch, root = caterpillar(N)
        else:
            t = synth.yule(N, int(rng.integers(1 << 30)), 0.1)
            ch, root = t.children, t.root
        topo = D.Topology(N, ch, root)
        cuts, progs = topo.split()
        if N < 128 or kind == "caterpillar":
            assert len(cuts) == 0  # too small / no balanced cut: the tree runs whole
            continue
        assert 2 <= len(cuts) <= 16
        P = len(cuts)
        dd, ud, parent = topo.info()
        leaves = rng.integers(1, 16, (N, 23)).astype(np.uint8)
        osets, *_ = O.fitch_downpass(N, ch, root, leaves)
        ofin = O.fitch_uppass(N, ch, root, osets)
        # down: parts, then the top reading the cut nodes' rows
        seen = []
        sets = {v: osets[v] for v in range(N)}
        for k in range(P + 1):
            prog, fifo = progs[("down", k)]
            seen += list(prog[:, 0])
            rows = np.array([sets[v] for v in range(N)] + [np.zeros(23, np.uint8)] * (N - 1))
            for v, s_ in sets.items():
                rows[v] = s_
            acc, stack, f = None, [], 0
            for v, A, B, fl in prog:
                if fl & 16:
                    stack.append(acc)
                sa, sb = fl & 3, (fl >> 2) & 3
                x = rows[fifo[f]] if sa == 0 else (acc if sa == 1 else stack.pop())
                f += sa == 0
                y = rows[fifo[f]] if sb == 0 else acc
                f += sb == 0
                tt = x & y
                acc = np.where(tt != 0, tt, x | y).astype(np.uint8)
                sets[v] = acc
            assert not stack and f == len(fifo)
            if k < P:
                assert prog[-1, 0] == cuts[k]
        assert sorted(seen) == list(range(N, 2 * N - 1))
        for v in range(N, 2 * N - 1):
            assert np.array_equal(sets[v], osets[v]), v
        # up: the top, then every part from its parent's final
        allsets = {v: osets[v] for v in range(2 * N - 1)}
        prog, fifo = progs[("up", 0)]
        fin = _interp_up(N, prog, fifo, allsets, expect=len(fifo))
        seen = list(prog[:, 0])
        for k in range(P):
            prog, fifo = progs[("up", k + 1)]
            assert prog[0, 0] == cuts[k] and (prog[0, 3] & 3) == 3
            fin = _interp_up(N, prog, fifo, allsets, fin=fin, parent_of_root=parent[cuts[k]],
                             expect=len(fifo))
            seen += list(prog[:, 0])
        assert sorted(seen) == list(range(N, 2 * N - 1))
        for v in range(2 * N - 1):
            assert np.array_equal(fin[v], ofin[v]), (kind, v)


def test_yule_is_valid_and_deterministic():
    t1 = synth.yule(1024, 7, 0.05)
    t2 = synth.yule(1024, 7, 0.05)
    assert np.array_equal(t1.children, t2.children) and np.array_equal(t1.blen, t2.blen)
    D.Topology(1024, t1.children, t1.root)  # validates
    assert t1.blen[t1.root] == 0
    assert abs(np.delete(t1.blen, t1.root).mean() - 0.05) < 0.01


def test_fitch_leaves_distribution():
    lv = synth.fitch_leaves(64, 20000, 3)
    s = D.unpack_nibbles(lv, 20000)
    single = np.isin(s, [1, 2, 4, 8]).mean()
    two = np.isin(s, [3, 5, 6, 9, 10, 12]).mean()
    gap = (s == 15).mean()
    assert abs(single - 0.95) < 0.005 and abs(two - 0.04) < 0.004 and abs(gap - 0.01) < 0.002
    assert np.array_equal(lv, synth.fitch_leaves(64, 20000, 3))


def test_simulate_compress_unique():
    t = synth.yule(12, 5, 0.1)
    pi, co = synth.gtr_params(4, 5.0, 9)
    from phylocaml_amd import mlmodel as MM
    Q = MM.gtr_q(pi, co)
    st = synth.simulate(t, pi, Q, np.array([0.5, 1.5]), 3000, 11)
    assert st.max() < 4
    pats, w = synth.compress(st)
    assert w.sum() == 3000 and pats.shape[1] == w.size
    assert len({tuple(c) for c in pats.T}) == w.size
    u = synth.simulate(t, pi, Q, np.array([1.0]), 500, 11, unique=True)
    assert len({tuple(c) for c in u.T}) == 500
    # base composition of the root-state draw follows pi on a star-like scale
    big = synth.simulate(synth.yule(2, 1, 1e-9), pi, Q, np.ones(1), 200000, 3)
    freq = np.bincount(big[0], minlength=4) / 200000
    assert np.allclose(freq, pi, atol=0.005)


def test_pack_dna_matches_numpy(rng):
    st = rng.integers(0, 4, (5, 77)).astype(np.uint8)
    st[1, 3] = 255
    sets = np.where(st < 4, 1 << np.minimum(st, 3), 15).astype(np.uint8)
    assert np.array_equal(synth.pack_dna(st), D.pack_nibbles(sets))


def test_ptopology_store_is_functional():
    from phylocaml_amd.diagnosis import Ptopology, RootData

    t0 = Ptopology()
    t1 = t0.add_node_data(3, "a").add_root_data((1, 2), "r")
    t2 = t1.add_comp_root(0, RootData({}, 5.0, 4.0)).add_comp_root_i(0, 7, ((1, 2), "r"))
    assert t0.node_data == {} and t1.get_node_data(3) == "a"
    assert t2.get_comp_root_i(0, 7) == ((1, 2), "r")
    assert t2.get_comp_root(0).component_cost == 5.0
    t3 = t2.remove_comp_root_i(0, 7).remove_node_data(3).remove_root_data((1, 2))
    assert t3.get_comp_root(0).assignment == {} and 7 in t2.get_comp_root(0).assignment
    assert t3.node_data == {} and t3.root_data == {} and t2.get_node_data(3) == "a"
    assert t3.remove_comp_root(0).comp_root == {}


def test_entry_points_refuse_a_null_context():
    # argument checks that run before any device work: a null context is
    # PHY_EINVAL for the per-node Fitch gather and the whole-tree passes
    import ctypes as C

    from phylocaml_amd._lib import lib, ptr

    idx = np.zeros(4, np.int64)
    assert lib.phy_fitch_select(None, None, 16, None, ptr(idx, C.c_int64), 4, None, None) == 1
    assert lib.phy_fitch_downpass(None, None, None, 1, 16, None, None, None, None) == 1
    assert lib.phy_fitch_uppass(None, None, None, None, 1, 16) == 1
    assert lib.phy_sankoff_downpass(None, None, 4, None, None, 16, 16, None, None, None) == 1
