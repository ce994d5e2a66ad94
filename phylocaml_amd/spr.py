"""SPR neighbourhoods, exact scoring on the GPU, and the learned proposal.

The reference's TreeNeighborhoods.generate_spr (lib/treeNeighborhoods.ml:
27-75) never compiled; the move is defined in include/phylo_hip.h
(phy_spr_moves / phy_spr_apply) on the rooted layout of phy_topo_create.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import device as D
from ._lib import check, lib, ptr


def moves(n_leaves: int, children, root: int) -> np.ndarray:
    """All (s, v) SPR moves of the tree, int32 [n, 2]."""
    ch = np.ascontiguousarray(children, np.int32)
    n = C.c_int64()
    check(lib.phy_spr_moves(n_leaves, ptr(ch, C.c_int32), root, C.byref(n), None, 0))
    out = np.zeros((n.value, 2), np.int32)
    check(lib.phy_spr_moves(n_leaves, ptr(ch, C.c_int32), root, C.byref(n), ptr(out, C.c_int32),
                            n.value))
    return out


def apply(n_leaves: int, children, root: int, blen, s: int, v: int):
    """(children, root, blen) of the tree after move (s, v)."""
    ch = np.ascontiguousarray(children, np.int32)
    bl = np.ascontiguousarray(blen, np.float64)
    ch_out = np.zeros_like(ch)
    bl_out = np.zeros_like(bl)
    r = C.c_int32()
    check(lib.phy_spr_apply(n_leaves, ptr(ch, C.c_int32), root, ptr(bl, C.c_double), int(s), int(v),
                            ptr(ch_out, C.c_int32), C.byref(r), ptr(bl_out, C.c_double)),
          None, f"spr apply ({s}, {v})")
    return ch_out, r.value, bl_out


class Scorer:
    """Exact lnL of SPR neighbours: many candidate trees per fused launch."""

    def __init__(self, ctx: D.Context, dm, n_leaves: int, tips, stride: int, weights,
                 n_patterns: int, form: int = 0, chunk: int = 2048):
        torch = __import__("torch")
        self.ctx, self.dm, self.N = ctx, dm, int(n_leaves)
        self.tips, self.stride, self.weights = tips, int(stride), weights
        self.n_patterns, self.form, self.chunk = int(n_patterns), int(form), int(chunk)
        nb = C.c_size_t()
        check(lib.phy_lk_spr_work_size(self.N, dm.handle, self.n_patterns, self.form, self.chunk,
                                       C.byref(nb)))
        self.work_bytes = nb.value
        self.work = torch.empty((self.work_bytes,), dtype=torch.uint8, device=f"cuda:{ctx.device}")

    def score(self, children, root: int, blen, mv: np.ndarray):
        """Device float64 tensor of lnL for each move of mv [n, 2]."""
        torch = __import__("torch")
        mv = np.ascontiguousarray(mv, np.int32)
        ch = np.ascontiguousarray(children, np.int32)
        bl = np.ascontiguousarray(blen, np.float64)
        out = torch.empty((max(len(mv), 1),), dtype=torch.float64, device=f"cuda:{self.ctx.device}")
        self.ctx.check(lib.phy_lk_spr_scores(
            self.ctx.handle, self.dm.handle, self.form, self.n_patterns, D.dptr(self.tips),
            self.stride, D.dptr(self.weights), self.N, ptr(ch, C.c_int32), int(root),
            ptr(bl, C.c_double), len(mv), ptr(mv, C.c_int32), self.chunk, D.dptr(out),
            D.dptr(self.work), self.work_bytes), "spr_scores")
        return out[: len(mv)]


class IncrementalScorer:
    """Exact lnL of SPR neighbours without a tree per move
    (phy_lk_spr_scores_incr): per pruned subtree, a repair of the base
    tree's partials and one pre-order pass score every regraft.  Same
    values as Scorer up to rounding; the workspace grows with the tree's
    longest repair path."""

    def __init__(self, ctx: D.Context, dm, n_leaves: int, tips, stride: int, weights,
                 n_patterns: int, form: int = 0, s_batch: int = 16):
        torch = __import__("torch")
        self.ctx, self.dm, self.N = ctx, dm, int(n_leaves)
        self.tips, self.stride, self.weights = tips, int(stride), weights
        self.n_patterns, self.form, self.s_batch = int(n_patterns), int(form), int(s_batch)
        self.work = torch.empty((0,), dtype=torch.uint8, device=f"cuda:{ctx.device}")
        self.base_lnl = torch.zeros((1,), dtype=torch.float64, device=f"cuda:{ctx.device}")

    def work_size(self, children, root: int) -> int:
        ch = np.ascontiguousarray(children, np.int32)
        nb = C.c_size_t()
        check(lib.phy_lk_spr_incr_work_size(self.N, ptr(ch, C.c_int32), int(root), self.dm.handle,
                                            self.n_patterns, self.form, self.s_batch, C.byref(nb)))
        return nb.value

    def score(self, children, root: int, blen, mv: np.ndarray):
        """Device float64 tensor of lnL for each move of mv [n, 2]; the base
        tree's lnL lands in self.base_lnl."""
        torch = __import__("torch")
        mv = np.ascontiguousarray(mv, np.int32)
        ch = np.ascontiguousarray(children, np.int32)
        bl = np.ascontiguousarray(blen, np.float64)
        need = self.work_size(ch, root)
        if self.work.numel() < need:
            self.work = torch.empty((need,), dtype=torch.uint8, device=f"cuda:{self.ctx.device}")
        out = torch.empty((max(len(mv), 1),), dtype=torch.float64, device=f"cuda:{self.ctx.device}")
        self.ctx.check(lib.phy_lk_spr_scores_incr(
            self.ctx.handle, self.dm.handle, self.form, self.n_patterns, D.dptr(self.tips),
            self.stride, D.dptr(self.weights), self.N, ptr(ch, C.c_int32), int(root),
            ptr(bl, C.c_double), len(mv), ptr(mv, C.c_int32), self.s_batch, D.dptr(out),
            D.dptr(self.base_lnl), D.dptr(self.work), self.work.numel()), "spr_scores_incr")
        return out[: len(mv)]


class SprPlan:
    """Incremental scoring of a fixed move list on a fixed topology,
    prepared once (phy_spr_plan_create for SPR pairs [n, 2],
    phy_tbr_plan_create for TBR triples [n, 3]) and run on device branch
    lengths (phy_spr_plan_run): a run issues kernels only, so ``capture``
    records it in a HIP graph that rescoring replays after the lengths
    (written into the same device tensor) change.  Capture needs ``ctx`` on
    a side stream."""

    def __init__(self, ctx: D.Context, dm, n_leaves: int, tips, stride: int, weights,
                 n_patterns: int, children, root: int, mv: np.ndarray, form: int = 0,
                 s_batch: int = 16):
        torch = __import__("torch")
        self.ctx, self.dm, self.N = ctx, dm, int(n_leaves)
        self.tips, self.stride, self.weights = tips, int(stride), weights
        mv = np.ascontiguousarray(mv, np.int32)
        tbr = mv.ndim == 2 and mv.shape[1] == 3
        self.mv = mv if tbr else mv.reshape(-1, 2)
        ch = np.ascontiguousarray(children, np.int32)
        h = C.c_void_p()
        create = lib.phy_tbr_plan_create if tbr else lib.phy_spr_plan_create
        ctx.check(create(ctx.handle, dm.handle, int(form), int(n_patterns), self.N,
                         ptr(ch, C.c_int32), int(root), len(self.mv), ptr(self.mv, C.c_int32),
                         int(s_batch), C.byref(h)),
                  "tbr_plan_create" if tbr else "spr_plan_create")
        self.handle = h
        nb = C.c_size_t()
        check(lib.phy_spr_plan_work_size(h, C.byref(nb)))
        dev = f"cuda:{ctx.device}"
        self.work = torch.empty((max(nb.value, 8),), dtype=torch.uint8, device=dev)
        self.lnl = torch.zeros((max(len(self.mv), 1),), dtype=torch.float64, device=dev)
        self.base_lnl = torch.zeros((1,), dtype=torch.float64, device=dev)

    def run(self, blen_dev, dm=None) -> None:
        """Enqueue the scores of every move at the device lengths blen_dev
        into self.lnl (and the base tree's into self.base_lnl)."""
        dm = dm or self.dm
        self.ctx.check(lib.phy_spr_plan_run(
            self.ctx.handle, self.handle, dm.handle, D.dptr(self.tips), self.stride,
            D.dptr(self.weights), D.dptr(blen_dev), D.dptr(self.lnl), D.dptr(self.base_lnl),
            D.dptr(self.work), self.work.numel()), "spr_plan_run")

    def capture(self, blen_dev):
        torch = __import__("torch")
        if not self.ctx.stream:
            raise ValueError("graph capture needs the context on a side stream")
        s = torch.cuda.ExternalStream(self.ctx.stream, device=f"cuda:{self.ctx.device}")
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self.run(blen_dev)
        return g

    def close(self):
        if getattr(self, "handle", None):
            lib.phy_spr_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Move features
# ---------------------------------------------------------------------------
N_FEATURES = 32

FEATURE_NAMES = [
    "log1p size(s)", "log1p size(v)", "log1p (N - size(s))", "size(s)/N", "size(v)/N",
    "t_s", "t_p", "t_x", "t_v", "t_u", "t_x + t_p", "log t_s", "log t_v",
    "depth(s)/D", "depth(v)/D", "path(s,v)/D", "log1p path(s,v)", "patristic(s,v)",
    "v ancestor of s", "v inside subtree(x)", "v adjacent to the prune point",
    "Fitch insert at v (%)", "Fitch insert at origin (%)", "Fitch change (%)",
    "Fitch change / (origin + 1)", "signed log Fitch change", "Fitch at v / (origin + 1)",
    "s child of root", "p is root", "(depth(v) - depth(s))/D", "t_s size(s)/N",
    "t_v/2 + t_s",
]


class TreeInfo:
    """Host arrays of a rooted tree: parent, depth, leaf counts, root distances,
    Euler intervals and binary-lifting ancestors."""

    def __init__(self, N, children, root, blen):
        self.N, self.root = int(N), int(root)
        M = 2 * N - 1
        ch = np.asarray(children, np.int64).reshape(-1, 2)
        self.parent = np.full(M, -1, np.int64)
        self.parent[ch[:, 0]] = np.arange(N, M)
        self.parent[ch[:, 1]] = np.arange(N, M)
        self.blen = np.asarray(blen, np.float64)
        order = [self.root]
        for v in order:
            if v >= N:
                order.extend(ch[v - N])
        self.order = np.array(order)
        self.depth = np.zeros(M, np.int64)
        self.rdist = np.zeros(M)
        for v in order[1:]:
            self.depth[v] = self.depth[self.parent[v]] + 1
            self.rdist[v] = self.rdist[self.parent[v]] + self.blen[v]
        self.size = np.zeros(M, np.int64)
        self.size[:N] = 1
        for v in order[::-1]:
            if v >= N:
                self.size[v] = self.size[ch[v - N]].sum()
        # Euler tour intervals
        self.tin = np.zeros(M, np.int64)
        self.tout = np.zeros(M, np.int64)
        t = 0
        stack = [(self.root, 0)]
        while stack:
            v, st = stack.pop()
            if st == 0:
                self.tin[v] = t
                t += 1
                stack.append((v, 1))
                if v >= N:
                    stack.append((ch[v - N][1], 0))
                    stack.append((ch[v - N][0], 0))
            else:
                self.tout[v] = t
        LOG = max(1, int(self.depth.max()).bit_length())
        up = np.where(self.parent < 0, np.arange(M), self.parent)
        self.up = [up]
        for _ in range(LOG):
            self.up.append(self.up[-1][self.up[-1]])
        self.children = ch

    def is_anc(self, a, b):
        """a is an ancestor of (or equal to) b."""
        return (self.tin[a] <= self.tin[b]) & (self.tout[b] <= self.tout[a])

    def lca(self, a, b):
        a, b = np.asarray(a).copy(), np.asarray(b).copy()
        swap = self.depth[a] < self.depth[b]
        a[swap], b[swap] = b[swap], a[swap]
        diff = self.depth[a] - self.depth[b]
        for k in range(len(self.up)):
            sel = (diff >> k) & 1 == 1
            a[sel] = self.up[k][a[sel]]
        for k in range(len(self.up) - 1, -1, -1):
            ua, ub = self.up[k][a], self.up[k][b]
            ne = ua != ub
            a[ne], b[ne] = ua[ne], ub[ne]
        return np.where(a == b, a, self.up[0][a])


def fitch_sets(ctx, topo, tips_dev, n_sites, weights_dev):
    """Down- and up-pass Fitch sets of the tree (device), for the features."""
    from . import nonadditive as NA

    ft = NA.alloc_tree(ctx, topo, n_sites, None, with_final=True)
    ft.sets[: topo.n_leaves].copy_(tips_dev[:, : ft.stride])
    NA.downpass(ctx, topo, ft, weights_dev)
    NA.uppass(ctx, topo, ft)
    return ft


def features(ctx, info: TreeInfo, ft, n_sites, weights_dev, mv: np.ndarray,
             total_weight: float):
    """float32 [n, 32] features of the moves mv [n, 2] (see FEATURE_NAMES)."""
    torch = __import__("torch")
    N = info.N
    s, v = mv[:, 0].astype(np.int64), mv[:, 1].astype(np.int64)
    p = info.parent[s]
    chp = info.children[p - N]
    x = np.where(chp[:, 0] == s, chp[:, 1], chp[:, 0])
    g = info.parent[p]
    u = info.parent[v]
    Dmax = max(1, int(info.depth.max()))
    bl = info.blen
    tu = np.where(u >= 0, bl[np.maximum(u, 0)], 0.0)
    l = info.lca(s, v)
    path = info.depth[s] + info.depth[v] - 2 * info.depth[l]
    pat = info.rdist[s] + info.rdist[v] - 2 * info.rdist[l]
    # Fitch insertion costs on the device: at v (edge (u, v)) and at the origin
    # (edge (g, x), or x alone when p is the root)
    uq = np.where(u >= 0, u, v)
    q_new = np.stack([s, uq, v], 1).astype(np.int32)
    us = np.unique(s)
    ps = info.parent[us]
    chs = info.children[ps - N]
    xs = np.where(chs[:, 0] == us, chs[:, 1], chs[:, 0])
    gs = info.parent[ps]
    q_old = np.stack([us, np.where(gs >= 0, gs, xs), xs], 1).astype(np.int32)
    q = np.concatenate([q_new, q_old])
    dq = D_to(q, ctx)
    costs = torch.empty((len(q),), dtype=torch.uint64, device=f"cuda:{ctx.device}")
    ctx.check(lib.phy_fitch_insertion_costs(ctx.handle, D.dptr(ft.sets), D.dptr(ft.final), ft.stride,
                                            n_sites, D.dptr(weights_dev), len(q), D.dptr(dq),
                                            D.dptr(costs)), "insertion_costs")
    c = costs.cpu().numpy().astype(np.float64)
    c_new = c[: len(mv)]
    c_old_u = dict(zip(us.tolist(), c[len(mv):].tolist()))
    c_old = np.array([c_old_u[k] for k in s.tolist()])
    W = max(total_weight, 1.0)
    dc = c_new - c_old
    F = np.zeros((len(mv), N_FEATURES), np.float32)
    cols = [
        np.log1p(info.size[s]), np.log1p(info.size[v]), np.log1p(N - info.size[s]),
        info.size[s] / N, info.size[v] / N,
        bl[s], bl[p], bl[x], bl[v], tu, bl[x] + bl[p], np.log(bl[s] + 1e-6), np.log(bl[v] + 1e-6),
        info.depth[s] / Dmax, info.depth[v] / Dmax, path / Dmax, np.log1p(path), pat,
        info.is_anc(v, s).astype(np.float64), info.is_anc(x, v).astype(np.float64),
        (u == g).astype(np.float64),
        100 * c_new / W, 100 * c_old / W, 100 * dc / W, dc / (c_old + 1.0),
        np.sign(dc) * np.log1p(np.abs(dc)), c_new / (c_old + 1.0),
        (p == info.root).astype(np.float64), (g < 0).astype(np.float64),
        (info.depth[v] - info.depth[s]) / Dmax, bl[s] * info.size[s] / N, bl[v] / 2 + bl[s],
    ]
    for k, col in enumerate(cols):
        F[:, k] = col
    return F


def D_to(a, ctx):
    return D.to_device(np.ascontiguousarray(a), ctx.device)
