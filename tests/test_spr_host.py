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
