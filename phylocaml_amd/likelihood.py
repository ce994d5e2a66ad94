"""Likelihood_c on the GPU: Felsenstein pruning behind the NodeData.S
signature (lib/nodeData.ml:3-35; every function is TODO in the reference,
lib/likelihood_c.ml:1-33).

A value holds a node's conditional likelihoods on the device -- partials
L[r][i][p] (FP32 or FP64, pattern-contiguous) with an int32 power-of-two
exponent per (rate class, pattern) -- plus the length of the branch from its
node to wherever the next operation joins it.  NodeData.S has no
branch-length argument, so the value carries it (``blen``).  NodeData costs
are minimised, so ``cost`` / ``root_cost`` are -lnL.

The NodeData operations take the arguments ``Make1D`` passes to every
NodeData module (``ctx``, the model ``m`` = a ``DeviceModel``, the previous
value ``prev`` of the same node or None):
  median_1 / median_2 / median_3, distance_1 / distance_2, adjust_3 (Newton
  on the edge between a node and the rest of the tree), cost / root_cost /
  leaf_cost, and the character bookkeeping.
``Evaluator`` runs a whole tree in one pass (lnL and optional per-branch
first/second derivatives for simultaneous Newton branch-length updates).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace

import numpy as np

from . import device as D
from ._lib import PHY_F32, PHY_F64, PHY_FUSED, check, lib
from .mlmodel import Model


class DeviceModel:
    """phy_model: pi, eigen-decomposition, rates and probabilities on the device."""

    def __init__(self, ctx: D.Context, model: Model):
        self.ctx = ctx
        self.model = model
        K = model.q.shape[0]
        R = model.rates.size
        h = C.c_void_p()
        arr = [np.ascontiguousarray(a, np.float64) for a in
               (model.priors, model.q, model.rates, model.probs)]
        self._keep = arr
        ctx.check(lib.phy_model_create(ctx.handle, K, R, *(a.ctypes.data_as(C.POINTER(C.c_double))
                                                           for a in arr), C.byref(h)),
                  "model_create")
        self.handle = h
        self.K, self.R = K, R

    def close(self):
        if getattr(self, "handle", None):
            lib.phy_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Likelihood:
    partial: object      # device [R*K, pitch] float32/float64
    exps: object         # device int32 [R, pitch]: value = partial * 2**exps
    blen: float          # branch from this node to the next join
    n_patterns: int
    weights: object = None  # device uint32 [n_patterns] (pattern weights)
    form: int = PHY_F64
    lnl: float | None = None   # lnL with this value as the root (set by root_cost / median_3)
    codes: list = field(default_factory=list)  # character code of each pattern
    dm: object = None    # the DeviceModel that made it (cost evaluates through it)

    @property
    def n_sites(self) -> int:
        return self.n_patterns

    @property
    def pitch(self) -> int:
        return int(self.partial.shape[1])


def _dtype(form):
    torch = __import__("torch")
    return torch.float32 if form & PHY_F32 else torch.float64


def _dev(ctx):
    return f"cuda:{ctx.device}"


def leaf(ctx, dm: DeviceModel, tip_sets: np.ndarray, blen: float, weights=None,
         form: int = PHY_F64) -> Likelihood:
    """median_1 of a leaf: partial = indicator of the tip's state set."""
    torch = __import__("torch")
    n = int(tip_sets.size)
    pitch = ((n + 63) // 64) * 64
    tip = D.to_device(D.pack_nibbles(np.asarray(tip_sets, np.uint8)), ctx.device)
    part = torch.empty((dm.R * dm.K, pitch), dtype=_dtype(form), device=_dev(ctx))
    ex = torch.zeros((dm.R, pitch), dtype=torch.int32, device=_dev(ctx))
    ctx.check(lib.phy_lk_leaf(ctx.handle, dm.handle, form, n, D.dptr(tip), D.dptr(part),
                              D.dptr(ex), pitch), "lk_leaf")
    w = weights if weights is None or not isinstance(weights, np.ndarray) else \
        D.to_device(np.asarray(weights, np.uint32), ctx.device)
    return Likelihood(part, ex, float(blen), n, w, form, dm=dm)


def missing(ctx, dm: DeviceModel, like: Likelihood, blen: float = 0.0) -> Likelihood:
    """A value that carries no information (every state possible at every
    pattern, L = 1): joining it changes nothing, so median_2(x, missing)
    moves x along its branch."""
    torch = __import__("torch")
    part = torch.ones_like(like.partial)
    ex = torch.zeros_like(like.exps)
    return replace(like, partial=part, exps=ex, blen=float(blen), lnl=None)


def _out(prev, like: Likelihood):
    torch = __import__("torch")
    if prev is not None and prev.partial.shape == like.partial.shape and \
            prev.partial.dtype == like.partial.dtype:
        return prev.partial, prev.exps
    return torch.empty_like(like.partial), torch.empty_like(like.exps)


def _check_pair(x: Likelihood, y: Likelihood, what: str):
    if x.n_patterns != y.n_patterns or x.pitch != y.pitch or x.form != y.form:
        raise ValueError(f"{what}: values of different pattern sets or forms")


def with_blen(x: Likelihood, t: float) -> Likelihood:
    """x with branch length t (the same device buffers)."""
    return replace(x, blen=float(t))


def median_1(m, prev, x: Likelihood) -> Likelihood:
    """median_1 (lib/likelihood_c.ml:14): a node with one child is its child."""
    return x


def median_2(ctx, m: DeviceModel, prev, x: Likelihood, y: Likelihood,
             blen: float | None = None) -> Likelihood:
    """median_2: L[r] = (P(r t_x) L_x[r]) (*) (P(r t_y) L_y[r]) for every
    rate class r.  The result's branch length is ``blen``, else that of
    ``prev`` (the node's previous value), else 0."""
    _check_pair(x, y, "median_2")
    out, ex = _out(prev, x)
    ctx.check(lib.phy_lk_median2(ctx.handle, m.handle, x.form, x.n_patterns, D.dptr(x.partial),
                                 D.dptr(x.exps), x.blen, D.dptr(y.partial), D.dptr(y.exps),
                                 y.blen, D.dptr(out), D.dptr(ex), x.pitch), "lk_median2")
    t = blen if blen is not None else (prev.blen if prev is not None else 0.0)
    return replace(x, partial=out, exps=ex, blen=float(t), lnl=None, dm=m)


def _lnl_buf(ctx):
    torch = __import__("torch")
    return torch.zeros((3,), dtype=torch.float64, device=_dev(ctx))


def median_3(ctx, m: DeviceModel, prev, x: Likelihood, y: Likelihood, z: Likelihood,
             blen: float | None = None) -> Likelihood:
    """median_3 (the unrooted root of Node.Make1D.median_3, lib/node.ml:200-216):
    the join of three neighbours, with its lnL (so ``cost`` is -lnL of the
    tree seen from this node)."""
    _check_pair(x, y, "median_3")
    _check_pair(x, z, "median_3")
    out, ex = _out(prev, x)
    buf = _lnl_buf(ctx)
    w = x.weights
    ctx.check(lib.phy_lk_median3(ctx.handle, m.handle, x.form, x.n_patterns, D.dptr(x.partial),
                                 D.dptr(x.exps), x.blen, D.dptr(y.partial), D.dptr(y.exps),
                                 y.blen, D.dptr(z.partial), D.dptr(z.exps), z.blen, D.dptr(out),
                                 D.dptr(ex), x.pitch, D.dptr(w), D.dptr(buf)), "lk_median3")
    t = blen if blen is not None else (prev.blen if prev is not None else 0.0)
    return replace(x, partial=out, exps=ex, blen=float(t), lnl=float(buf.cpu()[0]), dm=m)


def edge(ctx, m: DeviceModel, x: Likelihood, y: Likelihood, t: float) -> tuple:
    """(lnL, dlnL/dt, d2lnL/dt2) of x and y joined by one edge of length t."""
    _check_pair(x, y, "edge")
    buf = _lnl_buf(ctx)
    ctx.check(lib.phy_lk_edge(ctx.handle, m.handle, x.form, x.n_patterns, D.dptr(x.partial),
                              D.dptr(x.exps), D.dptr(y.partial), D.dptr(y.exps), x.pitch,
                              float(t), D.dptr(x.weights), D.dptr(buf)), "lk_edge")
    v = buf.cpu().numpy()
    return float(v[0]), float(v[1]), float(v[2])


def distance_1(ctx, m: DeviceModel, a: Likelihood, b: Likelihood) -> float:
    """distance_1 (lib/likelihood_c.ml:23): -lnL of a and b joined across
    their two branches (one edge of length t_a + t_b)."""
    return -edge(ctx, m, a, b, a.blen + b.blen)[0]


def distance_2(ctx, m: DeviceModel, a: Likelihood, b: Likelihood, c: Likelihood) -> float:
    """distance_2 (lib/likelihood_c.ml:24): -lnL of a, b and c joined at one
    point (c inserted on the edge between a and b, each at its own branch
    length from the insertion point)."""
    return -median_3(ctx, m, None, a, b, c).lnl


def adjust_3(ctx, m: DeviceModel, codes, n: Likelihood, a: Likelihood, b: Likelihood,
             c: Likelihood, max_iter: int = 20, tol: float = 1e-8, t_lo: float = 1e-8,
             t_hi: float = 10.0):
    """adjust_3 (lib/likelihood_c.ml:18) as Node.Make1D.readjust_3 calls it
    (lib/node.ml:239-256): n's own value is recomputed from its children a
    and b, then the length of the edge between n and c -- the value of the
    rest of the tree at n's parent -- is optimised by Newton-Raphson (a leaf
    has a = b = None and keeps its value).  Returns (n', changed codes): n's codes when the length moved (limited to
    ``codes`` when given; with ``codes`` sharing nothing with n, n is
    returned unchanged)."""
    if codes is not None and not (set(_codes(n)) & set(codes)):
        return n, set()
    d = n if a is None else median_2(ctx, m, None, a, b, blen=n.blen)
    t, lnl, ev = C.c_double(), C.c_double(), C.c_int32()
    ctx.check(lib.phy_lk_adjust3(ctx.handle, m.handle, d.form, d.n_patterns, D.dptr(d.partial),
                                 D.dptr(d.exps), D.dptr(c.partial), D.dptr(c.exps), d.pitch,
                                 D.dptr(d.weights), float(n.blen), t_lo, t_hi, max_iter, tol,
                                 C.byref(t), C.byref(lnl), C.byref(ev)), "lk_adjust3")
    out = replace(d, blen=float(t.value), lnl=None)
    changed = set(_codes(n)) if t.value != n.blen else set()
    if codes is not None:
        changed &= set(codes)
    return out, changed


def root_lnl(ctx, dm: DeviceModel, x: Likelihood) -> float:
    out = _lnl_buf(ctx)
    ctx.check(lib.phy_lk_root(ctx.handle, dm.handle, x.form, x.n_patterns, D.dptr(x.partial),
                              D.dptr(x.exps), x.pitch, D.dptr(x.weights),
                              D.dptr(out)), "lk_root")
    x.lnl = float(out.cpu().numpy()[0])  # kept for cost
    return x.lnl


def root_cost(ctx, dm: DeviceModel, x: Likelihood) -> float:
    """-lnL of the tree rooted at x (NodeData costs are minimised)."""
    return -root_lnl(ctx, dm, x)


def cost(x: Likelihood) -> float:
    """NodeData.cost: -lnL with x as the root (the patterns' likelihood
    under x's subtree, or the whole tree for a median_3 value); evaluated
    on first use through the model that made x."""
    if x.lnl is None:
        if x.dm is None:
            raise ValueError("cost: the value has no model; evaluate root_cost(ctx, dm, x)")
        return -root_lnl(x.dm.ctx, x.dm, x)
    return -x.lnl


def leaf_cost(x: Likelihood) -> float:
    return 0.0


# Character bookkeeping of NodeData.S (lib/nodeData.ml:9-16; TODO in
# lib/likelihood_c.ml:5-12), as Node.Make1D calls it (lib/node.ml:93-145):
# one character per pattern, identified by its code (``codes``, the pattern
# index when not given).
def _codes(x: Likelihood) -> list:
    return list(x.codes) if len(x.codes) == x.n_patterns else list(range(x.n_patterns))


def cardinal(x: Likelihood) -> int:
    return x.n_patterns


def get_codes(x: Likelihood) -> set:
    return set(_codes(x))


def mem(codes, x: Likelihood) -> bool:
    return True if codes is None else bool(set(codes) & get_codes(x))


def recode(f, x: Likelihood) -> Likelihood:
    return replace(x, codes=[f(c) for c in _codes(x)])


def _select(x: Likelihood, keep: list):
    torch = __import__("torch")
    if not keep:
        return None
    dev = x.partial.device
    idx = torch.as_tensor(keep, dtype=torch.int64, device=dev)
    n = len(keep)
    pitch = ((n + 63) // 64) * 64
    part = torch.ones((x.partial.shape[0], pitch), dtype=x.partial.dtype, device=dev)
    part[:, :n] = x.partial.index_select(1, idx)
    ex = torch.zeros((x.exps.shape[0], pitch), dtype=x.exps.dtype, device=dev)
    ex[:, :n] = x.exps.index_select(1, idx)
    w = None if x.weights is None else \
        x.weights.view(torch.int32).index_select(0, idx).view(torch.uint32)
    codes = _codes(x)
    return replace(x, partial=part, exps=ex, n_patterns=n, weights=w, lnl=None,
                   codes=[codes[i] for i in keep])


def filter_codes(codes, x: Likelihood):
    """The patterns of x whose code is in codes (None when there is none):
    their partials, exponents and weights; padding columns hold ones."""
    cs = set(codes)
    return _select(x, [i for i, c in enumerate(_codes(x)) if c in cs])


def filter_codes_comp(codes, x: Likelihood):
    cs = set(codes)
    return _select(x, [i for i, c in enumerate(_codes(x)) if c not in cs])


def compare(a: Likelihood, b: Likelihood) -> int:
    """A total order: codes, then branch length, then the scaled partials
    pattern by pattern (exponents first, then the values)."""
    ka, kb = _codes(a), _codes(b)
    if ka != kb:
        return -1 if ka < kb else 1
    if a.blen != b.blen:
        return -1 if a.blen < b.blen else 1
    n = a.n_patterns
    ea, eb = a.exps[:, :n], b.exps[:, :n]
    pa, pb = a.partial[:, :n], b.partial[:, :n]
    diff = (ea != eb).any(0) | (pa != pb).any(0)
    nz = diff.nonzero()
    if not nz.numel():
        return 0
    p = int(nz[0, 0])
    col_ea, col_eb = ea[:, p].cpu().tolist(), eb[:, p].cpu().tolist()
    if col_ea != col_eb:
        return -1 if col_ea < col_eb else 1
    col_a, col_b = pa[:, p].cpu().tolist(), pb[:, p].cpu().tolist()
    return -1 if col_a < col_b else 1


class Evaluator:
    """Whole-tree pruning on one compiled topology (phy_lk_tree).

    form: PHY_F64 or PHY_F32, optionally | PHY_FUSED (lnL only)."""

    def __init__(self, ctx, topo: D.Topology, dm: DeviceModel, n_patterns: int, form: int,
                 derivs: bool = False):
        torch = __import__("torch")
        self.ctx, self.topo, self.dm = ctx, topo, dm
        self.n_patterns, self.form, self.derivs = int(n_patterns), int(form), bool(derivs)
        nb = C.c_size_t()
        check(lib.phy_lk_tree_work_size(topo.handle, dm.handle, self.n_patterns, self.form,
                                        1 if derivs else 0, C.byref(nb)), ctx.handle, "work_size")
        self.work_bytes = nb.value
        dev = f"cuda:{ctx.device}"
        self.work = torch.empty((max(self.work_bytes, 8),), dtype=torch.uint8, device=dev)
        M = topo.n_nodes
        self.lnl = torch.zeros((1,), dtype=torch.float64, device=dev)
        self.site_lnl = torch.zeros((self.n_patterns,), dtype=torch.float64, device=dev)
        self.d1 = torch.zeros((M,), dtype=torch.float64, device=dev) if derivs else None
        self.d2 = torch.zeros((M,), dtype=torch.float64, device=dev) if derivs else None

    def run(self, tips, stride: int, weights, blen, site_lnl: bool = False) -> None:
        """Enqueue one evaluation; tips/weights/blen are device tensors."""
        self.ctx.check(lib.phy_lk_tree(
            self.ctx.handle, self.topo.handle, self.dm.handle, self.form, self.n_patterns,
            D.dptr(tips), stride, D.dptr(weights), D.dptr(blen), D.dptr(self.lnl),
            D.dptr(self.site_lnl) if site_lnl else None, D.dptr(self.d1), D.dptr(self.d2),
            D.dptr(self.work), self.work_bytes), "lk_tree")

    def partial(self, v: int):
        """(partial [R*K, pitch], exps [R, pitch]) views of internal node v after
        run(): row r*K + i holds L[r][i][p], and the value is
        L[r][i][p] * 2**exps[r][p] (one exponent per pattern and rate class)."""
        torch = __import__("torch")
        pp, pe, pitch = C.c_void_p(), C.c_void_p(), C.c_int64()
        check(lib.phy_lk_tree_partials(self.topo.handle, self.dm.handle, self.n_patterns,
                                       self.form, 1 if self.derivs else 0, D.dptr(self.work), v,
                                       C.byref(pp), C.byref(pe), C.byref(pitch)), None, "partials")
        esz = 4 if self.form & PHY_F32 else 8
        base = int(self.work.data_ptr())
        off_p, off_e = pp.value - base, pe.value - base
        R, K, P = self.dm.R, self.dm.K, pitch.value
        nbytes = R * K * P * esz
        part = self.work[off_p: off_p + nbytes].view(_dtype(self.form)).view(R * K, P)
        ex = self.work[off_e: off_e + 4 * R * P].view(torch.int32).view(R, P)
        return part, ex


def newton_step(blen: np.ndarray, d1: np.ndarray, d2: np.ndarray, root: int,
                lo: float = 1e-8, hi: float = 10.0) -> np.ndarray:
    """One Newton-Raphson branch-length update t <- t - d1/d2 where d2 < 0
    (a maximum), else a gradient step; clamped to [lo, hi]."""
    t = blen.copy()
    mask = np.arange(t.size) != root
    step = np.where(d2 < 0, -d1 / np.where(d2 < 0, d2, -1.0), 0.1 * np.sign(d1) * t)
    t[mask] = np.clip(t[mask] + step[mask], lo, hi)
    return t


__all__ = ["DeviceModel", "Likelihood", "Evaluator", "leaf", "missing", "with_blen", "median_1", "median_2",
           "median_3", "edge", "distance_1", "distance_2", "adjust_3", "root_lnl",
           "root_cost", "cost", "leaf_cost", "cardinal", "get_codes", "mem", "recode", "filter_codes",
           "filter_codes_comp", "compare", "newton_step", "PHY_F32", "PHY_F64", "PHY_FUSED"]
