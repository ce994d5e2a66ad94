"""NonAdditive_c on the GPU: the NodeData.S operations of Fitch parsimony
characters (lib/nodeData.ml:3-35, lib/nonAdditive_c.ml) over device nibble
blocks.

Each value mirrors the reference record ``{codes; weights; states; cost}``
(lib/nonAdditive_c.ml:1-5): ``states`` is a device block of packed 4-bit
IUPAC sets (16 sites per uint64), ``weights`` a device uint32 array shared
by the nodes of one character block (None = unit weights), ``cost`` the
node-local weighted cost.  Functions take the model argument ``m`` first,
like the signature, and ignore it (PModel is not used by median_2 in the
reference either, lib/nonAdditive_c.ml:19).

Costs cross back to the host (they are ``float`` in the signature); the
whole-tree functions at the end keep everything on the device.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace

import numpy as np

from . import device as D
from ._lib import lib, ptr


@dataclass
class NonAdditive:
    codes: list = field(default_factory=list)  # character codes (unused by the medians)
    weights: object = None                     # device uint32 [n_sites] or None
    states: object = None                      # device uint64 [words(n_sites)]
    n_sites: int = 0
    cost: float = 0.0


def _cost_buf(ctx):
    return D.empty((1,), dtype=__import__("torch").uint64, device=ctx.device)


def _read_cost(buf) -> float:
    return float(int(buf.cpu().numpy()[0]))


def of_sets(ctx: D.Context, sets: np.ndarray, weights: np.ndarray | None = None,
            codes=None) -> NonAdditive:
    """Upload one leaf: sets[s] = 4-bit IUPAC set of site s."""
    sets = np.asarray(sets, np.uint8)
    w = None if weights is None else D.to_device(np.asarray(weights, np.uint32), ctx.device)
    st = D.to_device(D.pack_nibbles(sets), ctx.device)
    return NonAdditive(codes=list(codes or []), weights=w, states=st, n_sites=sets.size, cost=0.0)


def to_sets(x: NonAdditive) -> np.ndarray:
    return D.unpack_nibbles(x.states.cpu().numpy(), x.n_sites)


def _out_block(prev: NonAdditive | None, like: NonAdditive):
    # the `t option` argument is the node's previous value (lib/node.ml:171,188):
    # reuse its device block in place
    if prev is not None and prev.states is not None and prev.states.numel() == like.states.numel():
        return prev.states
    return __import__("torch").empty_like(like.states)


def median_1(m, prev, x: NonAdditive) -> NonAdditive:
    """lib/nonAdditive_c.ml:18 -- the identity."""
    return x


def median_2(ctx: D.Context, m, prev, x: NonAdditive, y: NonAdditive) -> NonAdditive:
    """lib/nonAdditive_c.ml:19-35: Fitch median, cost weighted by x.weights."""
    out = _out_block(prev, x)
    cb = _cost_buf(ctx)
    ctx.check(lib.phy_fitch_median2(ctx.handle, D.dptr(x.states), D.dptr(y.states), D.dptr(out),
                                    x.n_sites, D.dptr(x.weights), D.dptr(cb)), "fitch_median2")
    return replace(x, states=out, cost=_read_cost(cb))


def median_3(ctx: D.Context, m, prev, a: NonAdditive, b: NonAdditive,
             c: NonAdditive) -> NonAdditive:
    """3-way Fitch median (slot of NodeData.median_3, lib/nodeData.ml:20)."""
    out = _out_block(prev, a)
    cb = _cost_buf(ctx)
    ctx.check(lib.phy_fitch_median3(ctx.handle, D.dptr(a.states), D.dptr(b.states),
                                    D.dptr(c.states), D.dptr(out), a.n_sites, D.dptr(a.weights),
                                    D.dptr(cb)), "fitch_median3")
    return replace(a, states=out, cost=_read_cost(cb))


def distance_1(ctx: D.Context, m, a: NonAdditive, b: NonAdditive) -> float:
    """Weighted bv_distance (lib/bitvector/bv.c:46-55)."""
    cb = _cost_buf(ctx)
    ctx.check(lib.phy_fitch_distance1(ctx.handle, D.dptr(a.states), D.dptr(b.states), a.n_sites,
                                      D.dptr(a.weights), D.dptr(cb)), "fitch_distance1")
    return _read_cost(cb)


def distance_2(ctx: D.Context, m, a: NonAdditive, b: NonAdditive, c: NonAdditive) -> float:
    """Cost of inserting c on the edge (a, b)."""
    cb = _cost_buf(ctx)
    ctx.check(lib.phy_fitch_distance2(ctx.handle, D.dptr(a.states), D.dptr(b.states),
                                      D.dptr(c.states), a.n_sites, D.dptr(a.weights), D.dptr(cb)),
              "fitch_distance2")
    return _read_cost(cb)


def adjust_3(ctx: D.Context, m, codes, n: NonAdditive, a: NonAdditive, b: NonAdditive,
             c: NonAdditive):
    """Final states of node n from its children a, b and its parent's final
    set c (Fitch 1971).  Returns (t, changed codes): the node's codes when any
    site's set changed (IntSet.t of lib/nodeData.ml:24)."""
    out = __import__("torch").empty_like(n.states)
    cb = _cost_buf(ctx)
    ctx.check(lib.phy_fitch_adjust3(ctx.handle, D.dptr(n.states), D.dptr(a.states),
                                    D.dptr(b.states), D.dptr(c.states), D.dptr(out), n.n_sites,
                                    D.dptr(cb)), "fitch_adjust3")
    changed = set(n.codes) if _read_cost(cb) > 0 else set()
    if codes is not None:
        changed &= set(codes)
    return replace(n, states=out), changed


def cost(x: NonAdditive) -> float:
    return x.cost


def root_cost(x: NonAdditive) -> float:
    return x.cost


def leaf_cost(x: NonAdditive) -> float:
    return 0.0


# ---------------------------------------------------------------------------
# Character bookkeeping of NodeData.S (lib/nodeData.ml:9-16; TODO in
# lib/nonAdditive_c.ml:9-16), as Node.Make1D calls it (lib/node.ml:93-145):
# a value holds one character per site, identified by its code (``codes``,
# the site index when not given).
# ---------------------------------------------------------------------------
def _codes(x: NonAdditive) -> list:
    return list(x.codes) if len(x.codes) == x.n_sites else list(range(x.n_sites))


def cardinal(x: NonAdditive) -> int:
    """Number of characters."""
    return x.n_sites


def get_codes(x: NonAdditive) -> set:
    return set(_codes(x))


def mem(codes, x: NonAdditive) -> bool:
    """Whether x holds any of codes (None: any code)."""
    return True if codes is None else bool(set(codes) & get_codes(x))


def recode(f, x: NonAdditive) -> NonAdditive:
    """The same states under codes f(code)."""
    return replace(x, codes=[f(c) for c in _codes(x)])


def _select(ctx: D.Context, x: NonAdditive, keep: list):
    torch = __import__("torch")
    if not keep:
        return None
    n = len(keep)
    idx = np.ascontiguousarray(keep, np.int64)
    out = torch.empty((D.words(n),), dtype=torch.uint64, device=x.states.device)
    w = None if x.weights is None else torch.empty((n,), dtype=torch.uint32,
                                                    device=x.weights.device)
    ctx.check(lib.phy_fitch_select(ctx.handle, D.dptr(x.states), x.n_sites, D.dptr(x.weights),
                                   ptr(idx, C.c_int64), n, D.dptr(out), D.dptr(w)), "fitch_select")
    codes = _codes(x)
    return NonAdditive(codes=[codes[i] for i in keep], weights=w, states=out, n_sites=n, cost=0.0)


def filter_codes(ctx: D.Context, codes, x: NonAdditive):
    """The characters of x whose code is in codes (phy_fitch_select), None
    when there is none.  The node cost is not carried over (median_2
    recomputes it)."""
    cs = set(codes)
    return _select(ctx, x, [i for i, c in enumerate(_codes(x)) if c in cs])


def filter_codes_comp(ctx: D.Context, codes, x: NonAdditive):
    """The characters of x whose code is not in codes (None when none is left)."""
    cs = set(codes)
    return _select(ctx, x, [i for i, c in enumerate(_codes(x)) if c not in cs])


def compare(a: NonAdditive, b: NonAdditive) -> int:
    """A total order: codes, then the state sets site by site, then the cost."""
    torch = __import__("torch")
    ka, kb = _codes(a), _codes(b)
    if ka != kb:
        return -1 if ka < kb else 1
    sa, sb = a.states.view(torch.int64), b.states.view(torch.int64)  # same bits
    ne = (sa != sb).nonzero()
    if ne.numel():
        k = int(ne[0, 0])
        wa = int(sa[k].cpu()) & (2 ** 64 - 1)
        wb = int(sb[k].cpu()) & (2 ** 64 - 1)
        for j in range(16):
            na, nb = (wa >> (4 * j)) & 15, (wb >> (4 * j)) & 15
            if na != nb:
                return -1 if na < nb else 1
    return (a.cost > b.cost) - (a.cost < b.cost)


# ---------------------------------------------------------------------------
# Whole-tree passes (Node.Make1D over a Tree, in one launch each)
# ---------------------------------------------------------------------------
@dataclass
class FitchTree:
    sets: object        # device uint64 [(2N-1), stride]; rows 0..N-1 leaves
    final: object       # device uint64 [(2N-1), stride] or None
    site_cost: object   # device uint32 [n_sites]
    node_cost: object   # device uint64 [N-1] or None
    total: object       # device uint64 [1]
    n_sites: int
    stride: int


def alloc_tree(ctx: D.Context, topo: D.Topology, n_sites: int, leaves=None,
               with_final: bool = True, with_node_cost: bool = False) -> FitchTree:
    torch = __import__("torch")
    N = topo.n_leaves
    W = D.words(n_sites)
    stride = W
    sets = torch.zeros((2 * N - 1, stride), dtype=torch.uint64, device=f"cuda:{ctx.device}")
    if leaves is not None:
        if isinstance(leaves, np.ndarray):
            sets[:N].copy_(torch.from_numpy(np.ascontiguousarray(leaves, np.uint64)))
        else:
            sets[:N].copy_(leaves)
    fin = torch.zeros_like(sets) if with_final else None
    sc = torch.zeros((max(n_sites, 1),), dtype=torch.uint32, device=f"cuda:{ctx.device}")
    nc = (torch.zeros((N - 1,), dtype=torch.uint64, device=f"cuda:{ctx.device}")
          if with_node_cost else None)
    tot = torch.zeros((1,), dtype=torch.uint64, device=f"cuda:{ctx.device}")
    return FitchTree(sets, fin, sc, nc, tot, n_sites, stride)


def downpass(ctx: D.Context, topo: D.Topology, ft: FitchTree, weights=None) -> None:
    ctx.check(lib.phy_fitch_downpass(ctx.handle, topo.handle, D.dptr(ft.sets), ft.stride,
                                     ft.n_sites, D.dptr(weights), D.dptr(ft.site_cost),
                                     D.dptr(ft.node_cost), D.dptr(ft.total)), "fitch_downpass")


def uppass(ctx: D.Context, topo: D.Topology, ft: FitchTree) -> None:
    ctx.check(lib.phy_fitch_uppass(ctx.handle, topo.handle, D.dptr(ft.sets), D.dptr(ft.final),
                                   ft.stride, ft.n_sites), "fitch_uppass")
