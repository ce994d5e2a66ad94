"""Tree neighbourhoods (NNI, SPR, TBR) with exact GPU scoring.

Mirrors the module type ``Neighborhood.S`` (lib/neighborhood.ml:1-20):
``neighborhood``, ``deferred_init`` / ``deferred_next`` /
``deferred_partition`` and ``lazy_neighborhood`` / ``lazy_of_deferred``.
The reference's generators (``generate_tbr``, ``generate_spr``,
``generate_nni``, lib/treeNeighborhoods.ml:27-75) never compiled; as there,
SPR is TBR where one side of the cut keeps its attachment (on the rooted
layout: no reroot, or no move of the attachment), and NNI is SPR at
distance 1.  Moves are defined in include/phylo_hip.h (phy_tbr_moves,
phy_tbr_apply, phy_nni_moves) on the rooted layout of phy_topo_create: a
delta is an SPR pair (s, v) or a TBR triple (s, w, v), v = -1 keeping the
subtree's attachment.

``Scorer`` scores a batch of deltas exactly on the GPU (many candidate trees
per fused pruning launch); ``hill_climb`` is the best-improvement search
over one neighbourhood kind.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import device as D
from ._lib import PhyError, check, lib, ptr


def num_edges(n: int) -> int:
    """Edges of an unrooted binary tree with n leaves: Tree.num_edges
    (lib/tree.ml:552-559): 0 for n <= 1, else 2n - 3."""
    if n < 0:
        raise ValueError("num_edges: negative leaf count")
    return 0 if n <= 1 else 2 * n - 3


def num_nodes(n: int) -> int:
    """Nodes of an unrooted binary tree with n leaves: Tree.num_nodes
    (lib/tree.ml:561-568): n for n <= 1, else 2n - 2."""
    if n < 0:
        raise ValueError("num_nodes: negative leaf count")
    return n if n <= 1 else 2 * n - 2


def num_unrooted_trees(n: int) -> int:
    """Distinct unrooted binary topologies on n labelled leaves:
    Tree.num_unrooted_trees (lib/tree.ml:570-586): 0 for n = 0, 1 for
    n < 3, else (2n - 5)!!."""
    if n < 0:
        raise ValueError("num_unrooted_trees: negative leaf count")
    if n == 0:
        return 0
    out = 1
    for k in range(2 * n - 5, 1, -2):
        out *= k
    return out


def num_rooted_trees(n: int) -> int:
    """Distinct rooted binary topologies on n labelled leaves:
    Tree.num_rooted_trees (lib/tree.ml:588-595): n for n <= 1, else
    num_edges(n) * num_unrooted_trees(n) = (2n - 3)!!."""
    if n < 0:
        raise ValueError("num_rooted_trees: negative leaf count")
    return n if n <= 1 else num_edges(n) * num_unrooted_trees(n)


def nni_moves(n_leaves: int, children, root: int) -> np.ndarray:
    """The 2(N - 3) NNI neighbours as SPR pairs, int32 [n, 2]."""
    ch = np.ascontiguousarray(children, np.int32)
    n = C.c_int64()
    check(lib.phy_nni_moves(n_leaves, ptr(ch, C.c_int32), root, C.byref(n), None, 0))
    out = np.zeros((n.value, 2), np.int32)
    check(lib.phy_nni_moves(n_leaves, ptr(ch, C.c_int32), root, C.byref(n),
                            ptr(out, C.c_int32), n.value))
    return out


def tbr_moves(n_leaves: int, children, root: int, max_join: int = -1,
              max_reroot: int = -1) -> np.ndarray:
    """TBR triples (s, w, v), int32 [n, 3]; max_reroot = 0 gives the SPR
    moves of the rooted side only (``Neighborhood("spr")`` adds the other)."""
    ch = np.ascontiguousarray(children, np.int32)
    n = C.c_int64()
    check(lib.phy_tbr_moves(n_leaves, ptr(ch, C.c_int32), root, int(max_join), int(max_reroot),
                            C.byref(n), None, 0))
    out = np.zeros((n.value, 3), np.int32)
    check(lib.phy_tbr_moves(n_leaves, ptr(ch, C.c_int32), root, int(max_join), int(max_reroot),
                            C.byref(n), ptr(out, C.c_int32), n.value))
    return out


def apply(n_leaves: int, children, root: int, blen, move):
    """(children, root, blen) after an SPR pair or a TBR triple."""
    ch = np.ascontiguousarray(children, np.int32)
    bl = np.ascontiguousarray(blen, np.float64)
    ch_out, bl_out, r = np.zeros_like(ch), np.zeros_like(bl), C.c_int32()
    if len(move) == 2:
        s, v = (int(x) for x in move)
        w = s
    else:
        s, w, v = (int(x) for x in move)
    check(lib.phy_tbr_apply(n_leaves, ptr(ch, C.c_int32), root, ptr(bl, C.c_double), s, w, v,
                            ptr(ch_out, C.c_int32), C.byref(r), ptr(bl_out, C.c_double)),
          None, f"tbr apply {tuple(move)}")
    return ch_out, r.value, bl_out


@dataclass
class Deferred:
    """A position in a neighbourhood's move list (the ``deferred`` type)."""

    moves: np.ndarray
    pos: int
    end: int


class Neighborhood:
    """Neighborhood.S for kind "nni", "spr" or "tbr" over t = (N, children,
    root); max_join / max_reroot bound the TBR distances (< 0: unlimited)."""

    def __init__(self, kind: str = "tbr", max_join: int = -1, max_reroot: int = -1):
        if kind not in ("nni", "spr", "tbr"):
            raise ValueError(kind)
        self.kind, self.max_join, self.max_reroot = kind, int(max_join), int(max_reroot)

    def neighborhood(self, t) -> np.ndarray:
        N, ch, root = t
        if self.kind == "nni":
            return nni_moves(N, ch, root)
        if self.kind == "spr":
            # either side of each cut edge moves, the other keeps its attachment
            return np.concatenate([tbr_moves(N, ch, root, self.max_join, 0),
                                   tbr_moves(N, ch, root, 0, self.max_join)])
        return tbr_moves(N, ch, root, self.max_join, self.max_reroot)

    def deferred_init(self, t) -> Deferred:
        mv = self.neighborhood(t)
        return Deferred(mv, 0, len(mv))

    @staticmethod
    def deferred_next(d: Deferred):
        """(the next state or None when this was the last delta, the delta)."""
        if d.pos >= d.end:
            raise StopIteration
        delta = tuple(int(x) for x in d.moves[d.pos])
        nxt = Deferred(d.moves, d.pos + 1, d.end) if d.pos + 1 < d.end else None
        return nxt, delta

    def deferred_partition(self, t, i: int) -> list:
        """i deferred states over disjoint, contiguous slices that cover the
        neighbourhood (some empty when it has fewer than i moves)."""
        mv = self.neighborhood(t)
        cuts = np.linspace(0, len(mv), int(i) + 1).astype(int)
        return [Deferred(mv, int(a), int(b)) for a, b in zip(cuts[:-1], cuts[1:])]

    def lazy_of_deferred(self, t, d: Deferred | None):
        while d is not None and d.pos < d.end:
            d, delta = self.deferred_next(d)
            yield delta

    def lazy_neighborhood(self, t):
        return self.lazy_of_deferred(t, self.deferred_init(t))


class Scorer:
    """Exact lnL of neighbours.  Two exact scorers: one candidate tree per
    move, many per fused pruning launch (phy_lk_spr_scores for pairs,
    phy_lk_tbr_scores for triples), or the incremental one
    (phy_lk_spr_scores_incr / phy_lk_tbr_scores_incr: per pruned subtree one
    repair of the base tree's partials, one pass over the subtree for its
    rerootings and one over the pruned tree for every regraft).
    ``incremental=None`` picks the incremental scorer for 4-state models
    with 1, 2, 4 or 8 rate classes when the moves average at least four per
    pruned subtree."""

    def __init__(self, ctx: D.Context, dm, n_leaves: int, tips, stride: int, weights,
                 n_patterns: int, form: int = 0, chunk: int = 2048,
                 incremental: bool | None = None, s_batch: int = 16):
        torch = __import__("torch")
        self.ctx, self.dm, self.N = ctx, dm, int(n_leaves)
        self.tips, self.stride, self.weights = tips, int(stride), weights
        self.n_patterns, self.form, self.chunk = int(n_patterns), int(form), int(chunk)
        self.incremental, self.s_batch = incremental, int(s_batch)
        nb = C.c_size_t()
        check(lib.phy_lk_spr_work_size(self.N, dm.handle, self.n_patterns, self.form, self.chunk,
                                       C.byref(nb)))
        self.work_bytes = nb.value
        dev = f"cuda:{ctx.device}"
        self.work = torch.empty((self.work_bytes,), dtype=torch.uint8, device=dev)
        self.iwork = torch.empty((0,), dtype=torch.uint8, device=dev)
        self.base_lnl = torch.zeros((1,), dtype=torch.float64, device=dev)

    def uses_incremental(self, mv: np.ndarray) -> bool:
        if self.incremental is not None:
            return bool(self.incremental)
        if getattr(self.dm, "K", 4) != 4 or getattr(self.dm, "R", 4) not in (1, 2, 4, 8):
            return False
        return len(mv) > 0 and len(mv) >= 4 * len(np.unique(mv[:, 0]))

    def score(self, children, root: int, blen, mv: np.ndarray):
        """Device float64 tensor of lnL for each delta of mv [n, 2] or [n, 3]
        (with the incremental scorer, the base tree's lnL in self.base_lnl)."""
        torch = __import__("torch")
        mv = np.ascontiguousarray(mv, np.int32)
        ch = np.ascontiguousarray(children, np.int32)
        bl = np.ascontiguousarray(blen, np.float64)
        out = torch.empty((max(len(mv), 1),), dtype=torch.float64,
                          device=f"cuda:{self.ctx.device}")
        tbr = mv.ndim == 2 and mv.shape[1] == 3
        if len(mv) and self.uses_incremental(mv):
            try:
                return self._score_incremental(ch, root, bl, mv, tbr, out)
            except PhyError as e:
                # the automatic choice falls back to one candidate tree per
                # move when the incremental form refuses the tree (an LDS
                # stack too deep for it: PHY_ERANGE)
                if self.incremental is not None or e.code != 5:
                    raise
        fn = lib.phy_lk_tbr_scores if tbr else lib.phy_lk_spr_scores
        self.ctx.check(fn(self.ctx.handle, self.dm.handle, self.form, self.n_patterns,
                          D.dptr(self.tips), self.stride, D.dptr(self.weights), self.N,
                          ptr(ch, C.c_int32), int(root), ptr(bl, C.c_double), len(mv),
                          ptr(mv, C.c_int32), self.chunk, D.dptr(out), D.dptr(self.work),
                          self.work_bytes), "neighbour scores")
        return out[: len(mv)]

    def _score_incremental(self, ch, root, bl, mv, tbr, out):
        torch = __import__("torch")
        nb = C.c_size_t()
        if tbr:
            check(lib.phy_lk_tbr_incr_work_size(
                self.N, ptr(ch, C.c_int32), int(root), self.dm.handle, self.n_patterns,
                self.form, len(mv), ptr(mv, C.c_int32), self.s_batch, C.byref(nb)),
                None, "tbr incremental work size")
        else:
            check(lib.phy_lk_spr_incr_work_size(
                self.N, ptr(ch, C.c_int32), int(root), self.dm.handle, self.n_patterns,
                self.form, self.s_batch, C.byref(nb)), None, "spr incremental work size")
        if self.iwork.numel() < nb.value:
            self.iwork = torch.empty((nb.value,), dtype=torch.uint8,
                                     device=f"cuda:{self.ctx.device}")
        fn = lib.phy_lk_tbr_scores_incr if tbr else lib.phy_lk_spr_scores_incr
        self.ctx.check(fn(self.ctx.handle, self.dm.handle, self.form, self.n_patterns,
                          D.dptr(self.tips), self.stride, D.dptr(self.weights), self.N,
                          ptr(ch, C.c_int32), int(root), ptr(bl, C.c_double), len(mv),
                          ptr(mv, C.c_int32), self.s_batch, D.dptr(out),
                          D.dptr(self.base_lnl), D.dptr(self.iwork), self.iwork.numel()),
                       "neighbour scores (incremental)")
        return out[: len(mv)]


def hill_climb(scorer: Scorer, base_lnl: float, N: int, children, root: int, blen,
               nb: Neighborhood, max_rounds: int = 100, log=None) -> dict:
    """Best-improvement search: each round scores the whole neighbourhood
    exactly and moves to its best tree if that raises lnL."""
    ch, r, bl = np.asarray(children, np.int32), int(root), np.asarray(blen, np.float64)
    lnl = float(base_lnl)
    evals, rounds, t0 = 0, 0, time.time()
    for rounds in range(1, max_rounds + 1):
        mv = nb.neighborhood((N, ch, r))
        if len(mv) == 0:
            break
        sc = scorer.score(ch, r, bl, mv).cpu().numpy()
        evals += len(mv)
        k = int(np.argmax(sc))
        if log:
            log(f"{nb.kind} round {rounds}: {len(mv)} neighbours, best {sc[k]:.4f} (from {lnl:.4f})")
        if not sc[k] > lnl:
            break
        ch, r, bl = apply(N, ch, r, bl, mv[k])
        lnl = float(sc[k])
    return {"kind": nb.kind, "rounds": rounds, "exact_evaluations": evals, "final_lnl": lnl,
            "wall_s": time.time() - t0, "children": ch, "root": r, "blen": bl}


def main(argv=None):
    """Best-improvement NNI, SPR and TBR searches from the same perturbed
    start, every neighbour scored exactly on the GPU; prints one JSON line."""
    import argparse
    import json
    import sys

    from . import likelihood as LK
    from .spr_search import make_instance

    ap = argparse.ArgumentParser()
    ap.add_argument("--taxa", type=int, default=128)
    ap.add_argument("--patterns", type=int, default=20_000)
    ap.add_argument("--perturb", type=int, default=10)
    ap.add_argument("--tbr-join", type=int, default=3)
    ap.add_argument("--tbr-reroot", type=int, default=3)
    ap.add_argument("--fp64", action="store_true", help="score in FP64 (default FP32-scaled)")
    args = ap.parse_args(argv)

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    inst = make_instance(11, args.taxa, args.patterns, args.perturb)
    ctx = D.Context(0)
    dm = LK.DeviceModel(ctx, inst.model)
    tips = D.to_device(inst.tips_np)
    wd = D.to_device(inst.weights)
    stride = inst.tips_np.shape[1]
    sc = Scorer(ctx, dm, inst.N, tips, stride, wd, inst.P, form=0 if args.fp64 else 1)
    topo = D.Topology(inst.N, inst.children, inst.root, ctx)
    ev = LK.Evaluator(ctx, topo, dm, inst.P, 0 if args.fp64 else 1)
    ev.run(tips, stride, wd, D.to_device(np.asarray(inst.blen, np.float64)))
    base = float(ev.lnl.cpu()[0])
    log(f"start lnL {base:.4f}")
    out = {"workload": f"{inst.N}-taxon Yule tree, {inst.P} unique GTR+G4 patterns, start = "
                       f"true tree after {args.perturb} random SPR moves; branch lengths as "
                       "the moves set them (not re-optimised)",
           "form": "f64" if args.fp64 else "f32", "start_lnl": base}
    kinds = [("nni", Neighborhood("nni")), ("spr", Neighborhood("spr")),
             ("tbr", Neighborhood("tbr", args.tbr_join, args.tbr_reroot))]
    for name, nb in kinds:
        res = hill_climb(sc, base, inst.N, inst.children, inst.root, inst.blen, nb, log=log)
        for k in ("children", "root", "blen"):
            res.pop(k)
        res["evaluations_per_s"] = res["exact_evaluations"] / max(res["wall_s"], 1e-9)
        if name == "tbr":
            res["max_join"], res["max_reroot"] = args.tbr_join, args.tbr_reroot
        out[name] = res
        log(json.dumps(res))
    print(json.dumps(out), flush=True)
    return out


__all__ = ["nni_moves", "tbr_moves", "apply", "Deferred", "Neighborhood", "Scorer",
           "hill_climb", "num_edges", "num_nodes", "num_unrooted_trees", "num_rooted_trees"]


if __name__ == "__main__":
    main()
