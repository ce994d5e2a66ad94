"""Batched 20-state (amino-acid) pruning: many trees in one persistent launch
on BF16 MFMA (phy_aa_batch_* in include/phylo_hip.h).

``AATree`` holds one tree's inputs (topology, branch lengths, GTR-20 model,
tips as state indices 0..19 with >= 20 = missing, weights).  ``AABatch``
uploads a list of them and evaluates every tree's lnL per ``run()``.
``synth_workload`` draws the benchmark workload: taxa ~ U{16..256},
patterns ~ U{1000..50000}, exchangeabilities ~ Dirichlet(1^190), state
frequencies ~ Dirichlet(2^20), Gamma-4 with alpha ~ U(0.3, 2), branch lengths
~ Exp(0.1).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import device as D, mlmodel as MM, synth
from ._lib import check, lib, ptr

K = 20


@dataclass
class AATree:
    n_leaves: int
    children: np.ndarray  # int32 [2(N-1)]
    root: int
    blen: np.ndarray      # float64 [2N-1]
    pi: np.ndarray        # [20]
    q: np.ndarray         # [20, 20]
    rates: np.ndarray     # [R]
    probs: np.ndarray     # [R]
    tips: np.ndarray      # uint8 [N, P]
    weights: np.ndarray   # uint32 [P]

    @property
    def n_patterns(self) -> int:
        return int(self.tips.shape[1])


class AABatch:
    def __init__(self, ctx: D.Context, trees: list[AATree], waves: int = 4):
        torch = __import__("torch")
        self.ctx, self.trees, self.waves = ctx, trees, int(waves)
        T = len(trees)
        R = trees[0].rates.size
        cat = np.concatenate
        nl = np.array([t.n_leaves for t in trees], np.int32)
        npat = np.array([t.n_patterns for t in trees], np.int64)
        ch = cat([np.asarray(t.children, np.int32) for t in trees])
        roots = np.array([t.root for t in trees], np.int32)
        bl = cat([np.asarray(t.blen, np.float64) for t in trees])
        pi = cat([np.asarray(t.pi, np.float64) for t in trees])
        q = cat([np.asarray(t.q, np.float64).ravel() for t in trees])
        rates = cat([np.asarray(t.rates, np.float64) for t in trees])
        probs = cat([np.asarray(t.probs, np.float64) for t in trees])
        tips = cat([np.ascontiguousarray(t.tips, np.uint8).ravel() for t in trees])
        w = cat([np.asarray(t.weights, np.uint32) for t in trees])
        h = C.c_void_p()
        ctx.check(lib.phy_aa_batch_create(
            ctx.handle, T, ptr(nl, C.c_int32), ptr(npat, C.c_int64), ptr(ch, C.c_int32),
            ptr(roots, C.c_int32), ptr(bl, C.c_double), ptr(pi, C.c_double), ptr(q, C.c_double),
            R, ptr(rates, C.c_double), ptr(probs, C.c_double), ptr(tips, C.c_uint8),
            ptr(w, C.c_uint32), self.waves, C.byref(h)), "aa_batch_create")
        self.handle = h
        pn, ni, dep = C.c_int64(), C.c_int64(), C.c_int32()
        check(lib.phy_aa_batch_info(h, C.byref(pn), C.byref(ni), C.byref(dep)))
        self.pattern_nodes, self.n_items, self.stack_depth = pn.value, ni.value, dep.value
        self.lnl = torch.zeros((T,), dtype=torch.float64, device=f"cuda:{ctx.device}")

    def run(self) -> None:
        self.ctx.check(lib.phy_aa_batch_run(self.ctx.handle, self.handle, D.dptr(self.lnl)),
                       "aa_batch_run")

    def close(self):
        if getattr(self, "handle", None):
            lib.phy_aa_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synth_tree(seed: int, n_leaves: int, n_patterns: int, n_sim: int | None = None,
               mean_bl: float = 0.1) -> AATree:
    """One tree of the batched workload.  n_sim sites are simulated under the
    tree's GTR-20+G4 model; the remaining patterns resample those columns
    (n_sim = None simulates all of them)."""
    rng = np.random.default_rng(seed)
    tree = synth.yule(n_leaves, seed, mean_bl)
    ex = rng.dirichlet(np.ones(K * (K - 1) // 2))
    pi = rng.dirichlet(np.full(K, 2.0))
    alpha = rng.uniform(0.3, 2.0)
    spec = MM.Spec(K, ("GTR", (ex / ex[-1])[:-1]), ("DiscreteGamma", 4, alpha),
                   ("Empirical", pi), MM.PHY_GAMMA_MEAN)
    m = MM.create(spec)
    ns = n_patterns if n_sim is None else min(n_sim, n_patterns)
    sim = synth.simulate(tree, m.priors, m.q, m.rates, ns, seed)
    st = sim
    if ns < n_patterns:
        st = np.empty((n_leaves, n_patterns), np.uint8)
        st[:, :ns] = sim
        cols = rng.integers(0, ns, n_patterns - ns)
        np.take(sim, cols, axis=1, out=st[:, ns:])
    w = synth.poisson_weights(n_patterns, 3.0, seed)
    return AATree(n_leaves, tree.children, tree.root, tree.blen, m.priors, m.q, m.rates, m.probs,
                  np.ascontiguousarray(st), w)


def synth_workload(n_trees: int = 2048, seed: int = 3, n_sim: int = 1000,
                   taxa=(16, 256), patterns=(1000, 50000)) -> list[AATree]:
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)
    shapes = [(int(rng.integers(taxa[0], taxa[1] + 1)), int(rng.integers(patterns[0], patterns[1] + 1)))
              for _ in range(n_trees)]
    with ThreadPoolExecutor(max_workers=16) as ex:
        return list(ex.map(lambda a: synth_tree(seed * 100003 + a[0], a[1][0], a[1][1], n_sim=n_sim),
                           enumerate(shapes)))


def tips_bitmask(tips: np.ndarray) -> np.ndarray:
    """State indices -> uint32 state bitmasks (>= 20 -> all 20 states)."""
    t = np.asarray(tips, np.uint32)
    return np.where(t < K, np.left_shift(1, np.minimum(t, 31)), (1 << K) - 1).astype(np.uint32)


# ---------------------------------------------------------------------------
# Autotuner: waves per workgroup (tile = 16 * waves patterns) per size bucket
# ---------------------------------------------------------------------------
BUCKETS = ((2, 64), (64, 128), (128, 1 << 30))  # taxa ranges [lo, hi)
WAVE_CHOICES = (1, 2, 4, 8)


def bucket_of(n_leaves: int) -> int:
    for k, (lo, hi) in enumerate(BUCKETS):
        if lo <= n_leaves < hi:
            return k
    return len(BUCKETS) - 1


def _time_batch(torch, b: AABatch, reps: int) -> float:
    b.run()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        b.run()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def autotune(ctx: D.Context, trees: list[AATree], sample: int = 48, reps: int = 2,
             seed: int = 0) -> dict:
    """For each size bucket, time every waves choice on a sample of that
    bucket's trees; returns {bucket: {"waves": w, "ms": {w: ms}, "range": ...}}."""
    torch = __import__("torch")
    rng = np.random.default_rng(seed)
    out = {}
    for k, (lo, hi) in enumerate(BUCKETS):
        idx = [i for i, t in enumerate(trees) if lo <= t.n_leaves < hi]
        if not idx:
            continue
        pick = [trees[i] for i in rng.choice(idx, min(sample, len(idx)), replace=False)]
        ms = {}
        for w in WAVE_CHOICES:
            try:
                b = AABatch(ctx, pick, waves=w)
            except Exception:
                continue  # this tile does not fit (LDS)
            ms[w] = _time_batch(torch, b, reps)
            b.close()
        best = min(ms, key=ms.get)
        out[k] = {"taxa": [lo, hi], "waves": best, "tile_patterns": 16 * best,
                  "ms": {str(w): v for w, v in ms.items()}, "sample_trees": len(pick)}
    return out


class TunedBatch:
    """The whole batch in one persistent launch when every size bucket's
    autotuned waves agree; otherwise the trees split by bucket, one
    persistent launch per bucket with its own waves.  lnl in the original
    tree order."""

    def __init__(self, ctx: D.Context, trees: list[AATree], choice: dict):
        torch = __import__("torch")
        self.ctx = ctx
        self.parts = []
        waves = {k: choice.get(k, {"waves": 4})["waves"] for k in range(len(BUCKETS))}
        present = {bucket_of(t.n_leaves) for t in trees}
        if len({waves[k] for k in present}) == 1:
            w = waves[next(iter(present))]
            self.parts.append((np.arange(len(trees)), AABatch(ctx, trees, waves=w)))
        else:
            for k in range(len(BUCKETS)):
                idx = [i for i, t in enumerate(trees) if bucket_of(t.n_leaves) == k]
                if not idx:
                    continue
                self.parts.append((np.array(idx),
                                   AABatch(ctx, [trees[i] for i in idx], waves=waves[k])))
        self.launches = len(self.parts)
        self.lnl = torch.zeros((len(trees),), dtype=torch.float64, device=f"cuda:{ctx.device}")
        self.pattern_nodes = sum(b.pattern_nodes for _, b in self.parts)
        self._idx = [torch.from_numpy(i).to(self.lnl.device) for i, _ in self.parts]

    def run(self) -> None:
        if self.launches == 1:
            self.parts[0][1].run()
            self.lnl.copy_(self.parts[0][1].lnl)
            return
        for (idx, b), di in zip(self.parts, self._idx):
            b.run()
            self.lnl.index_copy_(0, di, b.lnl)
