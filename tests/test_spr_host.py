"""CPU tests of the SPR neighbourhood (enumeration and application) and of
the host-side tree arrays used by the move features."""
import numpy as np
import pytest

from conftest import random_tree
from phylocaml_amd import device as D, spr, synth


def _leafsets(N, ch, root):
    ch = np.asarray(ch).reshape(-1, 2)
    out = {}

    def rec(v):
        if v < N:
            out[v] = frozenset([v])
        else:
            a, b = ch[v - N]
            out[v] = rec(a) | rec(b)
        return out[v]

    rec(root)
    return out


@pytest.mark.parametrize("N", [3, 4, 6, 10, 25])
def test_moves_count_and_application(rng, N):
    for _ in range(3):
        ch, root = random_tree(rng, N)
        bl = rng.exponential(0.1, 2 * N - 1)
        mv = spr.moves(N, ch, root)
        # brute force: every (s, v) with v outside subtree(s), v != p, sibling, new root
        sets = _leafsets(N, ch, root)
        parent = np.full(2 * N - 1, -1)
        chr_ = np.asarray(ch).reshape(-1, 2)
        for k in range(N - 1):
            parent[chr_[k]] = N + k
        expect = 0
        for s in range(2 * N - 1):
            if s == root:
                continue
            p = parent[s]
            x = chr_[p - N][0] if chr_[p - N][1] == s else chr_[p - N][1]
            new_root = x if p == root else root
            for v in range(2 * N - 1):
                if sets[v] <= sets[s] or v in (p, x, new_root):
                    continue
                expect += 1
        assert len(mv) == expect
        for s, v in mv[: 200]:
            c2, r2, b2 = spr.apply(N, ch, root, bl, s, v)
            D.Topology(N, c2, r2)  # valid tree
            ls = _leafsets(N, c2, r2)
            # the pruned subtree is intact and now the sibling of v's clade
            assert sets[s] in ls.values()
            assert (sets[s] | sets[v]) in ls.values() or v == root
            if parent[s] != root:  # tree length is preserved: t_x + t_p merged, t_v split
                assert abs((b2.sum() - b2[r2]) - (bl.sum() - bl[root])) < 1e-9


def test_apply_rejects_invalid():
    t = synth.yule(8, 1, 0.1)
    ch, root = t.children, t.root
    with pytest.raises(Exception):
        spr.apply(8, ch, root, t.blen, root, 0)   # root cannot be pruned
    s = 0
    with pytest.raises(Exception):
        spr.apply(8, ch, root, t.blen, s, s)      # into its own subtree


def test_treeinfo_lca_and_ancestry(rng):
    N = 40
    ch, root = random_tree(rng, N)
    bl = rng.exponential(0.1, 2 * N - 1)
    info = spr.TreeInfo(N, ch, root, bl)
    parent = info.parent

    def anc(v):
        out = [v]
        while parent[v] >= 0:
            v = parent[v]
            out.append(v)
        return out

    a = rng.integers(0, 2 * N - 1, 300)
    b = rng.integers(0, 2 * N - 1, 300)
    l = info.lca(a, b)
    for x, y, z in zip(a, b, l):
        ax, ay = anc(x), set(anc(y))
        assert z == next(u for u in ax if u in ay)
        assert info.is_anc(z, x) and info.is_anc(z, y)
    sets = _leafsets(N, ch, root)
    assert all(info.size[v] == len(sets[v]) for v in range(2 * N - 1))


def _malformed_trees():
    """(N, children, root) triples that are not rooted binary trees: a node
    that is its own child, a detached two-node cycle, a node with two
    parents, and a root that is someone's child."""
    N = 4  # internal nodes 4, 5, 6; a valid tree: 4 = (0, 1), 5 = (4, 2), 6 = (5, 3)
    return [
        ("self child", N, np.array([0, 1, 5, 2, 6, 3], np.int32), 6),  # 5 is its own child
        # 4 = (5, 0) and 5 = (4, 1): a cycle no path from root 6 = (2, 3) reaches
        ("detached cycle", N, np.array([5, 0, 4, 1, 2, 3], np.int32), 6),
        ("two parents", N, np.array([0, 1, 4, 1, 5, 3], np.int32), 6),
        ("root is a child", N, np.array([0, 1, 6, 2, 4, 3], np.int32), 6),
    ]


def test_malformed_trees_are_refused_by_every_move_entry_point():
    # phylo_hip.h promises PHY_EINVAL for a malformed tree: no entry point may
    # loop, exhaust memory or abort on one
    import ctypes as C

    from phylocaml_amd._lib import lib

    P_i32, P_i64, P_f64 = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_double)
    for what, N, ch, root in _malformed_trees():
        M = 2 * N - 1
        chp = ch.ctypes.data_as(P_i32)
        n = C.c_int64(0)
        out_ch = np.zeros(2 * (N - 1), np.int32)
        out_bl = np.zeros(M)
        bl = np.full(M, 0.1)
        r = C.c_int32(0)
        rcs = {
            "spr_moves": lib.phy_spr_moves(N, chp, root, C.byref(n), None, 0),
            "nni_moves": lib.phy_nni_moves(N, chp, root, C.byref(n), None, 0),
            "tbr_moves": lib.phy_tbr_moves(N, chp, root, -1, -1, C.byref(n), None, 0),
            "spr_apply": lib.phy_spr_apply(N, chp, root, bl.ctypes.data_as(P_f64), 0, 2,
                                           out_ch.ctypes.data_as(P_i32), C.byref(r),
                                           out_bl.ctypes.data_as(P_f64)),
            "tbr_apply": lib.phy_tbr_apply(N, chp, root, bl.ctypes.data_as(P_f64), 0, 0, 2,
                                           out_ch.ctypes.data_as(P_i32), C.byref(r),
                                           out_bl.ctypes.data_as(P_f64)),
        }
        assert all(rc == 1 for rc in rcs.values()), (what, rcs)
    # the valid tree of the same shape is accepted
    ok = np.array([0, 1, 4, 2, 5, 3], np.int32)
    n = C.c_int64(0)
    assert lib.phy_spr_moves(4, ok.ctypes.data_as(P_i32), 6, C.byref(n), None, 0) == 0
    assert n.value > 0
