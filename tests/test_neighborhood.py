"""NNI / SPR / TBR neighbourhoods on the host, pinned without the reference
(whose generators never compiled): NNI gives 2(n-3) distinct unrooted
neighbours one split away; SPR gives 2(n-3)(2n-7) distinct unrooted
neighbours (Allen & Steel 2001); TBR gives exactly the trees of an
independent brute-force TBR on the unrooted tree; the Neighborhood.S
deferred and lazy forms enumerate the same deltas."""
import numpy as np
import pytest

from conftest import caterpillar, random_tree


def _rooted_splits(N, ch, root):
    M = 2 * N - 1
    below = [0] * M
    for i in range(N):
        below[i] = 1 << i
    order, st = [], [root]
    while st:
        v = st.pop()
        order.append(v)
        if v >= N:
            st += [int(ch[2 * (v - N)]), int(ch[2 * (v - N) + 1])]
    for v in reversed(order):
        if v >= N:
            below[v] = below[int(ch[2 * (v - N)])] | below[int(ch[2 * (v - N) + 1])]
    full = (1 << N) - 1
    out = set()
    for v in range(M):
        if v == root:
            continue
        s = below[v]
        if s & 1:
            s = full ^ s
        if 1 < bin(s).count("1") < N - 1:
            out.add(s)
    return frozenset(out)


def _unrooted(N, ch, root):
    adj = {}

    def link(a, b):
        adj.setdefault(a, set()).add(b)
        adj.setdefault(b, set()).add(a)

    for k in range(N - 1):
        v = N + k
        if v == root:
            continue
        for c in (ch[2 * k], ch[2 * k + 1]):
            link(v, int(c))
    a, b = int(ch[2 * (root - N)]), int(ch[2 * (root - N) + 1])
    link(a, b)
    return adj


def _splits(adj, N):
    full = (1 << N) - 1
    out = set()
    for a in adj:
        for b in adj[a]:
            if a < b:
                seen, st, mask = {a, b}, [b], 0
                while st:
                    x = st.pop()
                    if x < N:
                        mask |= 1 << x
                    for y in adj[x]:
                        if y not in seen:
                            seen.add(y)
                            st.append(y)
                s = full ^ mask if mask & 1 else mask
                if 1 < bin(s).count("1") < N - 1:
                    out.add(s)
    return frozenset(out)


def _brute_tbr(N, adj):
    """Every tree one TBR away: cut an edge, join any edge (or lone leaf) of
    one side to any edge (or lone leaf) of the other, by subdivision."""
    res = set()
    nxt = [max(adj) + 1]

    def new():
        nxt[0] += 1
        return nxt[0]

    def comp(g, start):
        seen, st = {start}, [start]
        while st:
            x = st.pop()
            for y in g[x]:
                if y not in seen:
                    seen.add(y)
                    st.append(y)
        return seen

    def suppress(g, x):
        if x >= N and len(g[x]) == 2:
            p, q = g.pop(x)
            g[p].discard(x)
            g[q].discard(x)
            g[p].add(q)
            g[q].add(p)

    edges = [(a, b) for a in adj for b in adj[a] if a < b]
    for a, b in edges:
        g = {k: set(v) for k, v in adj.items()}
        g[a].discard(b)
        g[b].discard(a)
        ca, cb = comp(g, a), comp(g, b)
        suppress(g, a)
        suppress(g, b)
        ca = {x for x in ca if x in g}
        cb = {x for x in cb if x in g}

        def attach_points(c):
            es = [(x, y) for x in c for y in g[x] if x < y]
            return es if es else [(next(iter(c)), None)]

        for ea in attach_points(ca):
            for eb in attach_points(cb):
                h = {k: set(v) for k, v in g.items()}
                ends = []
                for e in (ea, eb):
                    x, y = e
                    if y is None:
                        ends.append(x)
                        continue
                    m = new()
                    h[x].discard(y)
                    h[y].discard(x)
                    h[m] = {x, y}
                    h[x].add(m)
                    h[y].add(m)
                    ends.append(m)
                h[ends[0]].add(ends[1])
                h[ends[1]].add(ends[0])
                res.add(_splits(h, N))
    return res


def _reached(N, ch, root, bl, mv):
    from phylocaml_amd import neighborhood as NB

    out = set()
    for m in mv:
        c2, r2, _ = NB.apply(N, ch, root, bl, m)
        out.add(_rooted_splits(N, c2, r2))
    return out


def _trees(rng):
    for N in (5, 6, 7, 8, 9):
        yield (N, *random_tree(rng, N))
    yield (8, *caterpillar(8))


def test_nni_count_and_one_split_away(rng):
    from phylocaml_amd import neighborhood as NB

    for N, ch, root in list(_trees(rng)) + [(40, *random_tree(rng, 40))]:
        bl = np.full(2 * N - 1, 0.1)
        mv = NB.nni_moves(N, ch, root)
        assert len(mv) == 2 * (N - 3)
        base = _rooted_splits(N, ch, root)
        got = _reached(N, ch, root, bl, mv)
        assert len(got) == 2 * (N - 3) and base not in got
        assert all(len(base - s) == 1 and len(s - base) == 1 for s in got)
        # TBR with join distance 1 and no reroot is NNI
        assert _reached(N, ch, root, bl, NB.tbr_moves(N, ch, root, 1, 0)) == got


def test_spr_neighbourhood_size_allen_steel(rng):
    from phylocaml_amd import neighborhood as NB

    for N, ch, root in _trees(rng):
        bl = np.full(2 * N - 1, 0.1)
        mv = NB.Neighborhood("spr").neighborhood((N, ch, root))
        got = _reached(N, ch, root, bl, mv)
        assert len(got) == 2 * (N - 3) * (2 * N - 7)
        assert _rooted_splits(N, ch, root) not in got
        # one direction of each cut edge alone misses some of them
        fwd = _reached(N, ch, root, bl, NB.tbr_moves(N, ch, root, -1, 0))
        assert fwd <= got


def test_tbr_equals_brute_force(rng):
    from phylocaml_amd import neighborhood as NB

    for N, ch, root in _trees(rng):
        bl = rng.uniform(0.01, 0.5, 2 * N - 1)
        mv = NB.tbr_moves(N, ch, root)
        got = _reached(N, ch, root, bl, mv)
        want = _brute_tbr(N, _unrooted(N, ch, root)) - {_rooted_splits(N, ch, root)}
        assert got == want
        spr = _reached(N, ch, root, bl, NB.Neighborhood("spr").neighborhood((N, ch, root)))
        assert spr <= got


def test_apply_keeps_a_valid_tree_and_its_length(rng):
    from phylocaml_amd import neighborhood as NB

    N = 30
    ch, root = random_tree(rng, N)
    bl = rng.uniform(0.01, 0.5, 2 * N - 1)
    bl[root] = 0.0
    parent = np.full(2 * N - 1, -1)
    for k in range(N - 1):
        parent[ch[2 * k]] = parent[ch[2 * k + 1]] = N + k
    mv = NB.tbr_moves(N, ch, root, 4, 3)
    assert len(mv) > 0
    for m in mv[:: max(1, len(mv) // 400)]:
        c2, r2, b2 = NB.apply(N, ch, root, bl, m)
        assert sorted(np.concatenate([c2, [r2]])) == list(range(2 * N - 1))
        assert b2[r2] == 0.0
        if parent[m[0]] != root:
            assert abs(b2.sum() - bl.sum()) < 1e-12
    with pytest.raises(Exception):
        NB.apply(N, ch, root, bl, (root, root, 0))


def test_neighborhood_interface_forms(rng):
    from phylocaml_amd import neighborhood as NB

    N = 12
    ch, root = random_tree(rng, N)
    t = (N, ch, root)
    for kind in ("nni", "spr", "tbr"):
        nb = NB.Neighborhood(kind, max_join=3, max_reroot=2)
        full = [tuple(int(x) for x in m) for m in nb.neighborhood(t)]
        assert list(nb.lazy_neighborhood(t)) == full
        d, seen = nb.deferred_init(t), []
        while d is not None:
            d, delta = nb.deferred_next(d)
            seen.append(delta)
        assert seen == full
        parts = nb.deferred_partition(t, 5)
        assert sum((list(nb.lazy_of_deferred(t, p)) for p in parts), []) == full


def _clades(N, ch, root):
    """Rooted topology key: the leaf set below every internal non-root node."""
    below = {i: 1 << i for i in range(N)}
    order, st = [], [root]
    while st:
        v = st.pop()
        order.append(v)
        if v >= N:
            st += [int(ch[2 * (v - N)]), int(ch[2 * (v - N) + 1])]
    for v in reversed(order):
        if v >= N:
            below[v] = below[int(ch[2 * (v - N)])] | below[int(ch[2 * (v - N) + 1])]
    return frozenset(below[v] for v in range(N, 2 * N - 1) if v != root)


def _walk(N, start, neighbours, key):
    """Breadth-first walk over the trees reachable by repeated moves."""
    seen = {key(N, *start): start}
    frontier = [start]
    while frontier:
        nxt = []
        for ch, root in frontier:
            for ch2, r2 in neighbours(ch, root):
                k = key(N, ch2, r2)
                if k not in seen:
                    seen[k] = (ch2, r2)
                    nxt.append((ch2, r2))
        frontier = nxt
    return seen


def test_tree_counts_match_reference_fixtures():
    # Tree.num_unrooted_trees / num_rooted_trees / num_nodes / num_edges
    # against the values of test/treeTest.ml:221-255
    import json
    import os

    from conftest import GOLDEN
    from phylocaml_amd import neighborhood as NB

    kat = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))["tree_counts"]
    fns = {"unrooted": NB.num_unrooted_trees, "rooted": NB.num_rooted_trees,
           "nodes": NB.num_nodes, "edges": NB.num_edges}
    for name, fn in fns.items():
        for n, want in kat[name]:
            assert fn(n) == want, (name, n)
        with pytest.raises(ValueError):
            fn(-1)
    # the rooted layout of N leaves has 2N - 1 nodes and 2N - 2 branches; its
    # unrooted view suppresses the root (one node, and its two branches are
    # one edge)
    for N in (2, 3, 10, 1024):
        assert (2 * N - 1) - 1 == NB.num_nodes(N) and (2 * N - 2) - 1 == NB.num_edges(N)


@pytest.mark.parametrize("N", [4, 5, 6, 7])
def test_nni_walk_reaches_every_unrooted_topology(N):
    # NNI connects the space of unrooted trees: walking from a caterpillar
    # reaches exactly (2N - 5)!! distinct split sets
    from phylocaml_amd import neighborhood as NB

    bl = np.full(2 * N - 1, 0.1)

    def nbrs(ch, root):
        for m in NB.nni_moves(N, ch, root):
            c2, r2, _ = NB.apply(N, ch, root, bl, m)
            yield c2, r2

    seen = _walk(N, caterpillar(N), nbrs, _rooted_splits)
    assert len(seen) == NB.num_unrooted_trees(N)


@pytest.mark.parametrize("N", [3, 4, 5])
def test_rooted_spr_walk_reaches_every_rooted_topology(N):
    # rooted SPR (every (s, v) of the rooted layout) connects the rooted
    # trees: the walk reaches exactly (2N - 3)!! distinct clade sets
    from phylocaml_amd import neighborhood as NB, spr

    bl = np.full(2 * N - 1, 0.1)

    def nbrs(ch, root):
        for s, v in spr.moves(N, ch, root):
            c2, r2, _ = spr.apply(N, ch, root, bl, int(s), int(v))
            yield c2, r2

    seen = _walk(N, caterpillar(N), nbrs, _clades)
    assert len(seen) == NB.num_rooted_trees(N)
