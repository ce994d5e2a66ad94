"""Whole-tree diagnosis glue: Node.Make1D over the GPU NodeData values, the
Ptopology store, and tree / component costs.

* ``Make1D`` mirrors the functor body of lib/node.ml:59-279: a node value
  (``Node``, the record ``n`` of lib/node.ml:66-73) maps character codes to
  NodeData values (here ``nonadditive.NonAdditive`` device blocks) and
  carries its height, its topology key and its cost, the sum of the
  characters' node-local costs.  Binary medians order their arguments by the
  topology key first (``Ordering.order``, lib/node.ml:156,184); the key here
  is the smallest leaf id below the node.
* ``Ptopology`` mirrors lib/ptopology.ml:7-54: node data by node id, root
  data by edge, and per component the root assignment with its component
  and adjusted costs.  Updates return a new store, as in the reference.
* ``diagnose_nodes`` diagnoses a tree node by node through ``Make1D`` (one
  median launch per internal node, the reference's call pattern, then the
  final sets top-down with readjust_3 and the leaf rule);
  ``diagnose_tree`` fills the same store from the whole-tree kernels (one
  post-order and one pre-order launch).  Both give the same sets and costs.
* ``diagnose_lnl_nodes`` runs the likelihood NodeData (``likelihood``)
  through the same ``Make1D``: median_2 in post-order, the root cost -lnL,
  and optionally a pre-order sweep of readjust_3 that optimises every branch
  length by Newton-Raphson against the value of the rest of the tree (built
  with median_2); ``lnl_tree`` is the whole-tree kernel's lnL.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

from . import nonadditive as NA


@dataclass
class Node:
    """Node.Make1D.n (lib/node.ml:66-73)."""

    prelim: dict
    final: dict | None
    topot: int
    height: int
    cost: float
    code: int


def _sum(f, d: dict) -> float:
    return float(sum(f(v) for v in d.values()))


class Make1D:
    """lib/node.ml:59-279 over a NodeData module ``nd`` (``nonadditive`` or
    ``likelihood``) on device ``ctx``."""

    def __init__(self, ctx, nd=NA):
        self.ctx, self.nd = ctx, nd

    # -- construction and accessors ------------------------------------------
    def of_data(self, id: int, x: dict) -> Node:
        return Node(dict(x), None, int(id), 0, _sum(self.nd.leaf_cost, x), int(id))

    @staticmethod
    def height(_, x: Node) -> int:
        return x.height

    @staticmethod
    def cardinal(x: Node) -> int:
        return int(sum(v.n_sites for v in x.prelim.values()))

    @staticmethod
    def get_codes(x: Node) -> set:
        out = set()
        for v in x.prelim.values():
            out |= set(v.codes)
        return out

    @staticmethod
    def get_preliminary(_, k, x: Node):
        return x.prelim[k]

    @staticmethod
    def get_adjusted(k, x: Node):
        if x.final is None:
            raise KeyError(k)
        return x.final[k]

    @staticmethod
    def get_topot(x: Node) -> int:
        return x.topot

    def compare(self, a: Node, b: Node) -> int:
        """Order of the preliminary values: the codes, then NodeData.compare
        of each code's values in code order."""
        ka, kb = sorted(a.prelim), sorted(b.prelim)
        if ka != kb:
            return (ka > kb) - (ka < kb)
        for k in ka:
            c = self.nd.compare(a.prelim[k], b.prelim[k])
            if c:
                return c
        return 0

    @staticmethod
    def filter_codes(cs, x: Node) -> Node:
        keep = set(cs)

        def f(d):
            return {k: v for k, v in d.items() if set(v.codes) & keep}

        return replace(x, prelim=f(x.prelim), final=None if x.final is None else f(x.final))

    @staticmethod
    def _order(a: Node, b: Node):
        return (a, b) if a.topot <= b.topot else (b, a)

    # -- distances and medians -----------------------------------------------
    def distance_1(self, m, a: Node, b: Node) -> float:
        a, b = self._order(a, b)
        return float(sum(self.nd.distance_1(self.ctx, m, av, b.prelim[k])
                         for k, av in a.prelim.items()))

    def distance_2(self, m, a: Node, b: Node, c: Node) -> float:
        return float(sum(self.nd.distance_2(self.ctx, m, av, b.prelim[k], c.prelim[k])
                         for k, av in a.prelim.items()))

    def median_1(self, m, id: int, o: Node | None, a: Node) -> Node:
        prelim = {k: self.nd.median_1(m, None if o is None else o.prelim[k], v)
                  for k, v in a.prelim.items()}
        return Node(prelim, None, a.topot, a.height + 1, _sum(self.nd.cost, prelim), int(id))

    def median_2(self, m, id: int, o: Node | None, a: Node, b: Node) -> Node:
        assert o is None or o.code == id
        a, b = self._order(a, b)
        prelim = {k: self.nd.median_2(self.ctx, m, None if o is None else o.prelim[k], av,
                                      b.prelim[k])
                  for k, av in a.prelim.items()}
        return Node(prelim, None, min(a.topot, b.topot), max(a.height, b.height) + 1,
                    _sum(self.nd.cost, prelim), int(id))

    def median_3(self, m, id: int, o: Node | None, a: Node, b: Node, c: Node) -> Node:
        assert o is None or o.code == id
        a, b, c = sorted((a, b, c), key=lambda x: x.topot)
        prelim = {k: self.nd.median_3(self.ctx, m, None if o is None else o.prelim[k], av,
                                      b.prelim[k], c.prelim[k])
                  for k, av in a.prelim.items()}
        return Node(prelim, None, min(a.topot, b.topot), max(a.height, b.height) + 1,
                    _sum(self.nd.cost, prelim), int(id))

    def readjust_3(self, m, codes, n: Node, a: Node | None, b: Node | None, c: Node,
                   prelim: bool = False):
        """NodeData.adjust_3 of n from its children a, b and c (lib/node.ml:
        238-256): for Fitch, c is the parent's final sets; for likelihood, the
        value of the rest of the tree at the parent, and a = b = None for a
        leaf.  Returns (node, changed codes)."""
        def val(x: Node, k):
            return x.prelim[k] if prelim or x.final is None else x.final[k]

        out, changed = {}, set()
        for k in n.prelim:
            v, s = self.nd.adjust_3(self.ctx, m, codes, val(n, k),
                                    None if a is None else a.prelim[k],
                                    None if b is None else b.prelim[k], val(c, k))
            out[k] = v
            changed |= s
        return replace(n, prelim=out, final=out), changed

    def uppass_heuristic_leaf(self, m, leaf: Node, parent: Node) -> Node:
        """Final sets of a leaf below a parent with final sets F (the slot of
        lib/node.ml:265, TODO there): per site F & L if that is non-empty,
        else L.  Built from two Fitch medians: median(F, L) is F & L when
        non-empty, else F | L, and its median with L is then F & L (a subset
        of L), resp. L."""
        fin = {}
        for k, lv in leaf.prelim.items():
            f = parent.final[k] if parent.final is not None else parent.prelim[k]
            t = self.nd.median_2(self.ctx, m, None, f, lv)
            fin[k] = replace(lv, states=self.nd.median_2(self.ctx, m, None, t, lv).states)
        return replace(leaf, final=fin)

    def with_branch(self, x: Node, t: float) -> Node:
        """x with every value's branch length set to t (NodeData.S carries no
        branch lengths; the likelihood values hold their own)."""
        return replace(x, prelim={k: self.nd.with_blen(v, t) for k, v in x.prelim.items()},
                       final=None if x.final is None else
                       {k: self.nd.with_blen(v, t) for k, v in x.final.items()})

    @staticmethod
    def root_cost(_, x: Node) -> float:
        return x.cost

    @staticmethod
    def node_cost(_, x: Node) -> float:
        return x.cost


# ---------------------------------------------------------------------------
# Ptopology (lib/ptopology.ml:7-54)
# ---------------------------------------------------------------------------
@dataclass
class RootData:
    """root_data (lib/ptopology.ml:7-11)."""

    assignment: dict = field(default_factory=dict)  # int -> (jxn, root)
    component_cost: float = 0.0
    adjusted_cost: float = 0.0


@dataclass(frozen=True)
class Ptopology:
    node_data: dict = field(default_factory=dict)
    root_data: dict = field(default_factory=dict)
    comp_root: dict = field(default_factory=dict)

    def _with(self, **kw) -> "Ptopology":
        return replace(self, **kw)

    def get_node_data(self, a):
        return self.node_data[a]

    def add_node_data(self, a, d) -> "Ptopology":
        return self._with(node_data={**self.node_data, a: d})

    def remove_node_data(self, a) -> "Ptopology":
        return self._with(node_data={k: v for k, v in self.node_data.items() if k != a})

    def get_root_data(self, e):
        return self.root_data[e]

    def add_root_data(self, e, d) -> "Ptopology":
        return self._with(root_data={**self.root_data, e: d})

    def remove_root_data(self, e) -> "Ptopology":
        return self._with(root_data={k: v for k, v in self.root_data.items() if k != e})

    def get_comp_root(self, h) -> RootData:
        return self.comp_root[h]

    def add_comp_root(self, h, d: RootData) -> "Ptopology":
        return self._with(comp_root={**self.comp_root, h: d})

    def remove_comp_root(self, h) -> "Ptopology":
        return self._with(comp_root={k: v for k, v in self.comp_root.items() if k != h})

    def get_comp_root_i(self, h, i):
        return self.comp_root[h].assignment[i]

    def add_comp_root_i(self, h, i, d) -> "Ptopology":
        y = self.comp_root[h]
        return self.add_comp_root(h, replace(y, assignment={**y.assignment, i: d}))

    def remove_comp_root_i(self, h, i) -> "Ptopology":
        y = self.comp_root[h]
        return self.add_comp_root(
            h, replace(y, assignment={k: v for k, v in y.assignment.items() if k != i}))


# ---------------------------------------------------------------------------
# Tree diagnosis
# ---------------------------------------------------------------------------
def _post_order(N: int, children: np.ndarray, root: int):
    out, stack = [], [(root, False)]
    while stack:
        v, done = stack.pop()
        if v < N:
            continue
        if done:
            out.append(v)
        else:
            k = 2 * (v - N)
            stack += [(v, True), (int(children[k + 1]), False), (int(children[k]), False)]
    return out


def _finish(pt: Ptopology, N: int, children: np.ndarray, root: int, total: float,
            handle: int) -> Ptopology:
    k = 2 * (root - N)
    edge = (int(children[k]), int(children[k + 1]))
    rn = pt.get_node_data(root)
    pt = pt.add_root_data(edge, rn)
    return pt.add_comp_root(handle, RootData({0: (edge, rn)}, total, total))


def diagnose_nodes(mk: Make1D, m, N: int, children, root: int, leaves: dict,
                   handle: int = 0, final: bool = True) -> Ptopology:
    """Diagnosis node by node through Make1D: leaves[id] = {code: NodeData};
    median_2 in post-order, then (final) readjust_3 top-down, which, as in
    the reference, replaces an internal node's preliminary sets by its final
    sets, and the leaf rule for the leaves."""
    children = np.asarray(children)
    pt = Ptopology()
    for i in range(N):
        pt = pt.add_node_data(i, mk.of_data(i, leaves[i]))
    order = _post_order(N, children, root)
    total = 0.0
    for v in order:
        k = 2 * (v - N)
        n = mk.median_2(m, v, None, pt.get_node_data(int(children[k])),
                        pt.get_node_data(int(children[k + 1])))
        total += n.cost
        pt = pt.add_node_data(v, n)
    if not final:
        return _finish(pt, N, children, root, total, handle)
    # pre-order: the root's final sets are its preliminary sets
    rn = pt.get_node_data(root)
    pt = pt.add_node_data(root, replace(rn, final=dict(rn.prelim)))
    for v in reversed(order):
        k = 2 * (v - N)
        if v != root:
            n, _ = mk.readjust_3(m, None, pt.get_node_data(v),
                                 pt.get_node_data(int(children[k])),
                                 pt.get_node_data(int(children[k + 1])),
                                 pt.get_node_data(_parent_of(v, N, children)))
            pt = pt.add_node_data(v, n)
        for c in (int(children[k]), int(children[k + 1])):
            if c < N:
                pt = pt.add_node_data(c, mk.uppass_heuristic_leaf(m, pt.get_node_data(c),
                                                                  pt.get_node_data(v)))
    return _finish(pt, N, children, root, total, handle)


def _parent_of(v: int, N: int, children: np.ndarray) -> int:
    idx = int(np.flatnonzero(children == v)[0])
    return N + idx // 2


def diagnose_tree(ctx, topo, leaf_sets: np.ndarray | None, weights=None, code: int = 0,
                  handle: int = 0, packed=None, n_sites: int | None = None) -> Ptopology:
    """The same store from the whole-tree kernels: leaf_sets [N, S] uint8 sets
    of one character block (code), or packed [N, words] nibble blocks with
    n_sites; weights a device uint32 [S] or None."""
    from . import device as D

    N = topo.n_leaves
    if packed is None:
        S = leaf_sets.shape[1]
        packed = D.pack_nibbles(leaf_sets)
    else:
        S = int(n_sites)
    ft = NA.alloc_tree(ctx, topo, S, packed, with_final=True, with_node_cost=True)
    NA.downpass(ctx, topo, ft, weights)
    NA.uppass(ctx, topo, ft)
    node_cost = ft.node_cost.cpu().numpy().astype(np.float64)
    total = float(int(ft.total.cpu().numpy()[0]))
    ch = topo.children
    pt = Ptopology()
    height = np.zeros(2 * N - 1, np.int64)
    topot = np.arange(2 * N - 1)
    for v in _post_order(N, ch, topo.root):
        a, b = int(ch[2 * (v - N)]), int(ch[2 * (v - N) + 1])
        height[v] = max(height[a], height[b]) + 1
        topot[v] = min(topot[a], topot[b])
    for v in range(2 * N - 1):
        pre = NA.NonAdditive([code], weights, ft.sets[v], S, 0.0 if v < N else node_cost[v - N])
        fin = replace(pre, states=ft.final[v])
        pt = pt.add_node_data(v, Node({code: pre}, {code: fin}, int(topot[v]), int(height[v]),
                                      pre.cost, v))
    return _finish(pt, N, ch, topo.root, total, handle)


def diagnose_lnl_nodes(mk: Make1D, m, N: int, children, root: int, leaves: dict, blen,
                       handle: int = 0, adjust: bool = False, codes=None):
    """Likelihood diagnosis node by node through ``mk`` (NodeData =
    ``likelihood``, m its DeviceModel): leaves[id] = {code: value}; every
    value gets the tree's branch length blen[id].  median_2 in post-order;
    the root's cost is -lnL.  With ``adjust``, one pre-order sweep of
    readjust_3 over every non-root node v: its branch is optimised by Newton
    on the edge between v's value and C_v, the value at v's parent of the
    rest of the tree (at a root child: the sibling moved to the root, at any
    other node: median_2 of the sibling and the parent's C), and each
    internal value is recomputed once its children are done, so the next
    edges see the new lengths.  Every step keeps or raises lnL.
    Returns (store, lnL, branch lengths)."""
    children = np.asarray(children)
    t = np.array(blen, np.float64)
    pt = Ptopology()
    for i in range(N):
        pt = pt.add_node_data(i, mk.with_branch(mk.of_data(i, leaves[i]), t[i]))
    order = _post_order(N, children, root)
    kids = {v: (int(children[2 * (v - N)]), int(children[2 * (v - N) + 1])) for v in order}

    def down(v):
        a, b = kids[v]
        n = mk.median_2(m, v, None, pt.get_node_data(a), pt.get_node_data(b))
        return mk.with_branch(n, t[v])

    for v in order:
        pt = pt.add_node_data(v, down(v))
    if adjust:
        nd = mk.nd

        def up_value(v, sib, c_parent):
            """C_v as a Node: the sibling joined with the parent's C (or the
            sibling alone, moved to the root), located at v's parent."""
            s_node = pt.get_node_data(sib)
            vals = {}
            for k, sv in s_node.prelim.items():
                other = (nd.missing(mk.ctx, m, sv) if c_parent is None
                         else c_parent.prelim[k])
                vals[k] = nd.median_2(mk.ctx, m, None, sv, other, blen=0.0)
            return Node(vals, None, s_node.topot, s_node.height, 0.0, v)

        def visit(v, c_v):
            nonlocal pt
            a, b = kids.get(v, (None, None))
            n = pt.get_node_data(v)
            an = None if a is None else pt.get_node_data(a)
            bn = None if b is None else pt.get_node_data(b)
            n2, _ = mk.readjust_3(m, codes, n, an, bn, c_v, prelim=True)
            t[v] = next(iter(n2.prelim.values())).blen
            pt = pt.add_node_data(v, replace(n2, final=None))
            if a is None:
                return
            # the parent's C seen from below: c_v moved down v's branch
            c_here = replace(c_v, prelim={k: nd.with_blen(x, t[v]) for k, x in c_v.prelim.items()})
            visit(a, up_value(a, b, c_here))
            visit(b, up_value(b, a, c_here))
            pt = pt.add_node_data(v, down(v))

        ra, rb = kids[root]
        visit(ra, up_value(ra, rb, None))
        visit(rb, up_value(rb, ra, None))
        pt = pt.add_node_data(root, down(root))
    rn = pt.get_node_data(root)
    lnl = -mk.root_cost(m, rn)
    return _finish(pt, N, children, root, -lnl, handle), lnl, t


def lnl_tree(ctx, topo, dm, blen, tip_sets: np.ndarray, weights, form: int) -> float:
    """lnL by the whole-tree kernel (likelihood.Evaluator)."""
    from . import device as D
    from . import likelihood as LK

    P = tip_sets.shape[1]
    ev = LK.Evaluator(ctx, topo, dm, P, form)
    ev.run(D.to_device(D.pack_nibbles(tip_sets), ctx.device), D.words(P),
           D.to_device(np.asarray(weights, np.uint32), ctx.device),
           D.to_device(np.asarray(blen, np.float64), ctx.device))
    return float(ev.lnl.cpu()[0])


def main(argv=None):
    """Time node-by-node against whole-tree Fitch diagnosis on one character
    block (default 1,024 taxa x 10^6 IUPAC sites); prints one JSON line."""
    import argparse
    import json
    import time

    from . import device as D
    from . import synth

    ap = argparse.ArgumentParser()
    ap.add_argument("--taxa", type=int, default=1024)
    ap.add_argument("--sites", type=int, default=1_000_000)
    args = ap.parse_args(argv)
    torch = __import__("torch")
    N, S = args.taxa, args.sites
    tree = synth.yule(N, 1, 0.1)
    packed = synth.fitch_leaves(N, S, 1)
    ctx = D.Context(0)
    topo = D.Topology(N, tree.children, tree.root, ctx)
    mk = Make1D(ctx)
    leaves = {i: {0: NA.NonAdditive([0], None, D.to_device(packed[i], ctx.device), S, 0.0)}
              for i in range(N)}
    out = {"workload": f"{N}-taxon Yule tree, {S} IUPAC sites, one character block"}
    dev_packed = D.to_device(packed, ctx.device)
    for name, fn in (("whole_tree", lambda: diagnose_tree(ctx, topo, None, packed=dev_packed,
                                                          n_sites=S)),
                     ("node_by_node", lambda: diagnose_nodes(mk, None, N, tree.children,
                                                             tree.root, leaves))):
        fn()  # warm-up
        torch.cuda.synchronize()
        t0 = time.time()
        pt = fn()
        torch.cuda.synchronize()
        out[name] = {"s": time.time() - t0, "total_cost": pt.get_comp_root(0).component_cost}
    out["speedup"] = out["node_by_node"]["s"] / out["whole_tree"]["s"]
    print(json.dumps(out), flush=True)
    return out


__all__ = ["Node", "Make1D", "RootData", "Ptopology", "diagnose_nodes", "diagnose_tree",
           "diagnose_lnl_nodes", "lnl_tree"]


if __name__ == "__main__":
    main()
