"""Learned SPR proposal: training data from exact pruning, MLP training on the
GPU, and hill-climbing search guided by the MLP versus exhaustive exact
scoring.

Each round of either search enumerates every SPR move of the current tree
(phy_spr_moves).  The exhaustive search scores every move exactly and applies
the best improving one.  The guided search scores every move with the MLP,
re-scores only the top-k exactly (FP64) and applies the best improving one
among them.  If none of the k improves, it falls back to one exhaustive round,
which either finds the improving move or proves the tree a local optimum.
Both stop at a tree with no improving move.

Exact scores keep the branch lengths of phy_spr_apply (not re-optimised).
Targets are transformed ΔlnL: sign(d) * log1p(|d|).

python -m phylocaml_amd.spr_search [--quick] prints one JSON summary.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import device as D, likelihood as LK, mlmodel as MM, spr, synth
from .mlp import MLP

F64_FUSED, F32_FUSED = 2, 3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


@dataclass
class Instance:
    N: int
    P: int
    model: MM.Model
    children: np.ndarray
    root: int
    blen: np.ndarray
    tips_np: np.ndarray   # nibble blocks [N, W]
    weights: np.ndarray   # uint32 [P]


def make_instance(seed: int, N: int = 256, P: int = 50_000, n_perturb: int = 10,
                  mean_bl: float = 0.05) -> Instance:
    """A pruning-benchmark-style instance: Yule tree, GTR+G4 (alpha 0.5), P unique
    simulated patterns with weights Poisson(3)+1; the start tree is the true
    tree after n_perturb random SPR moves."""
    rng = np.random.default_rng(seed)
    tree = synth.yule(N, seed, mean_bl)
    pi, co = synth.gtr_params(4, 5.0, seed)
    model = MM.create(MM.Spec(4, ("GTR", co), ("DiscreteGamma", 4, 0.5), ("Empirical", pi)))
    st = synth.simulate(tree, model.priors, model.q, model.rates, P, seed, unique=True)
    w = synth.poisson_weights(P, 3.0, seed)
    ch, r, bl = tree.children.copy(), tree.root, tree.blen.copy()
    for _ in range(n_perturb):
        mv = spr.moves(N, ch, r)
        s, v = mv[rng.integers(len(mv))]
        ch, r, bl = spr.apply(N, ch, r, bl, s, v)
    return Instance(N, P, model, ch, r, bl, synth.pack_dna(st), w)


class Context:
    """Device state of one instance: model, tips, exact scorers."""

    def __init__(self, ctx: D.Context, inst: Instance, chunk: int = 2048):
        self.ctx, self.inst = ctx, inst
        self.dm = LK.DeviceModel(ctx, inst.model)
        self.tips = D.to_device(inst.tips_np)
        self.wd = D.to_device(inst.weights)
        self.stride = inst.tips_np.shape[1]
        self.total_w = float(inst.weights.sum())
        self.exact64 = spr.Scorer(ctx, self.dm, inst.N, self.tips, self.stride, self.wd, inst.P,
                                  form=0, chunk=chunk)
        self.exact32 = spr.Scorer(ctx, self.dm, inst.N, self.tips, self.stride, self.wd, inst.P,
                                  form=1, chunk=chunk)
        self.inc64 = spr.IncrementalScorer(ctx, self.dm, inst.N, self.tips, self.stride, self.wd,
                                           inst.P, form=0)
        self.inc32 = spr.IncrementalScorer(ctx, self.dm, inst.N, self.tips, self.stride, self.wd,
                                           inst.P, form=1)
        self.n_exact = 0

    def lnl(self, ch, r, bl, form=F64_FUSED) -> float:
        topo = D.Topology(self.inst.N, ch, r, self.ctx)
        ev = LK.Evaluator(self.ctx, topo, self.dm, self.inst.P, form)
        ev.run(self.tips, self.stride, self.wd, D.to_device(np.asarray(bl, np.float64)))
        self.n_exact += 1
        return float(ev.lnl.cpu()[0])

    def exact(self, ch, r, bl, mv, fp64: bool = True, incremental: bool | None = None) -> np.ndarray:
        """Exact lnL of the moves mv [n, 2].  The incremental scorer costs
        about two passes per distinct pruned subtree, the per-candidate
        scorer one pass per move: by default the incremental one runs when
        the moves average at least 4 per pruned subtree."""
        mv = np.asarray(mv, np.int32).reshape(-1, 2)
        if incremental is None:
            incremental = len(mv) >= 4 * len(np.unique(mv[:, 0]))
        if incremental:
            sc = self.inc64 if fp64 else self.inc32
        else:
            sc = self.exact64 if fp64 else self.exact32
        self.n_exact += len(mv)
        return sc.score(ch, r, bl, mv).cpu().numpy()

    def features(self, ch, r, bl, mv) -> np.ndarray:
        info = spr.TreeInfo(self.inst.N, ch, r, bl)
        topo = D.Topology(self.inst.N, ch, r, self.ctx)
        ft = spr.fitch_sets(self.ctx, topo, self.tips, self.inst.P, self.wd)
        return spr.features(self.ctx, info, ft, self.inst.P, self.wd, mv, self.total_w)


def target(delta: np.ndarray) -> np.ndarray:
    return np.sign(delta) * np.log1p(np.abs(delta))


def training_pairs(ctx: D.Context, seeds, N, P, n_perturb=10):
    """(features, transformed exact ΔlnL) of every SPR move of each instance's
    start tree, scored exactly (FP32-scaled, rel. error ~1e-9) in batches."""
    X, T = [], []
    for sd in seeds:
        t0 = time.time()
        inst = make_instance(sd, N, P, n_perturb)
        c = Context(ctx, inst)
        mv = spr.moves(N, inst.children, inst.root)
        base = c.lnl(inst.children, inst.root, inst.blen, F32_FUSED)
        sc = c.exact(inst.children, inst.root, inst.blen, mv, fp64=False)
        X.append(c.features(inst.children, inst.root, inst.blen, mv))
        T.append(target(sc - base).astype(np.float32))
        log(f"  instance {sd}: {len(mv)} moves in {time.time() - t0:.1f}s "
            f"(best dlnL {np.max(sc - base):.2f})")
    return np.concatenate(X), np.concatenate(T)


class Normaliser:
    def __init__(self, X):
        self.mu = X.mean(0)
        self.sd = X.std(0) + 1e-6

    def __call__(self, X):
        return ((X - self.mu) / self.sd).astype(np.float32)


def train(ctx: D.Context, X, T, epochs=4, batch=1024, lr=1e-3, seed=0):
    torch = __import__("torch")
    norm = Normaliser(X)
    Xd = D.to_device(norm(X))
    tmu, tsd = float(T.mean()), float(T.std() + 1e-6)
    Td = D.to_device(((T - tmu) / tsd).astype(np.float32))
    net = MLP(ctx, X.shape[1], 128, seed)
    n = len(X)
    rng = np.random.default_rng(seed)
    losses = []
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n)).to(Xd.device)
        tot, cnt = 0.0, 0
        for k in range(0, n - batch + 1, batch):
            idx = perm[k:k + batch]
            want = (k // batch) % 50 == 0
            l = net.step(Xd[idx], Td[idx], lr=lr * (0.5 ** ep), want_loss=want)
            if want:
                tot += l
                cnt += 1
        losses.append(tot / max(cnt, 1))
        log(f"  epoch {ep}: loss {losses[-1]:.4f}")
    return net, norm, losses


def search(c: Context, net, norm, guided: bool, k: int = 64, eps: float = 1e-6,
           max_rounds: int = 200, fp64: bool = True):
    inst = c.inst
    ch, r, bl = inst.children.copy(), inst.root, inst.blen.copy()
    c.n_exact = 0
    cur = c.lnl(ch, r, bl)
    t0 = time.time()
    rounds, fallbacks, recall_hits, recall_n = 0, 0, 0, 0
    while rounds < max_rounds:
        mv = spr.moves(inst.N, ch, r)
        rounds += 1
        log(f"  {'guided' if guided else 'exhaustive'} round {rounds}: lnL {cur:.3f}, "
            f"{c.n_exact} exact evaluations, {time.time() - t0:.1f}s")
        if guided:
            X = norm(c.features(ch, r, bl, mv))
            pred = net.forward(D.to_device(X)).cpu().numpy()
            top = np.argpartition(-pred, min(k, len(mv) - 1))[:k]
            sc = c.exact(ch, r, bl, mv[top], fp64=fp64)
            j = int(np.argmax(sc))
            if sc[j] > cur + eps * abs(cur):
                ch, r, bl = spr.apply(inst.N, ch, r, bl, *mv[top[j]])
                cur = float(sc[j])
                continue
            fallbacks += 1
        sc = c.exact(ch, r, bl, mv, fp64=fp64)
        j = int(np.argmax(sc))
        if not guided and net is not None:
            X = norm(c.features(ch, r, bl, mv))
            pred = net.forward(D.to_device(X)).cpu().numpy()
            top = np.argpartition(-pred, min(k, len(mv) - 1))[:k]
            recall_n += 1
            recall_hits += int(j in set(top.tolist()))
        if sc[j] <= cur + eps * abs(cur):
            break
        ch, r, bl = spr.apply(inst.N, ch, r, bl, *mv[j])
        cur = float(sc[j])
    return {"final_lnl": cur, "rounds": rounds, "exact_evaluations": c.n_exact,
            "wall_s": time.time() - t0, "fallback_rounds": fallbacks,
            "best_move_recall_top_k": (recall_hits / recall_n) if recall_n else None,
            "moves_per_round": int(len(spr.moves(inst.N, ch, r)))}


def compare(ctx: D.Context, N: int, P: int, pert: int, train_instances: int = 4,
            top_k: int = 64, fp64: bool = True) -> dict:
    """Train the MLP on `train_instances` instances' exact move scores, then
    run the exhaustive and the guided search from the same held-out start
    (seed 7) to convergence; returns their summaries and ratios."""
    log(f"training data: {train_instances} instances, N={N}, P={P}")
    t0 = time.time()
    X, T = training_pairs(ctx, range(100, 100 + train_instances), N, P, pert)
    t_data = time.time() - t0
    t0 = time.time()
    net, norm, losses = train(ctx, X, T)
    t_train = time.time() - t0
    inst = make_instance(7, N, P, pert)
    c = Context(ctx, inst)
    log("exhaustive search")
    ex = search(c, net, norm, guided=False, k=top_k, fp64=fp64)
    log(json.dumps(ex))
    log("guided search")
    gd = search(c, net, norm, guided=True, k=top_k, fp64=fp64)
    log(json.dumps(gd))
    out = {
        "workload": f"{N}-taxon Yule tree, {P} unique GTR+G4 patterns, start = true tree "
                    f"after {pert} random SPR moves; every rooted SPR move each round",
        "training_pairs": int(len(X)), "training_data_s": t_data, "training_s": t_train,
        "train_loss_per_epoch": losses, "top_k": top_k, "exact_scores": "FP64" if fp64 else "FP32",
        "exhaustive": ex, "guided": gd,
        "exact_evaluation_ratio": ex["exact_evaluations"] / max(gd["exact_evaluations"], 1),
        "wall_time_ratio": ex["wall_s"] / max(gd["wall_s"], 1e-9),
        "final_lnl_gap": gd["final_lnl"] - ex["final_lnl"],
    }
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="small instances (tests)")
    ap.add_argument("--train-instances", type=int, default=4)
    ap.add_argument("--top-k", type=int, default=64)
    ap.add_argument("--fp32-exhaustive", action="store_true",
                    help="score the exhaustive baseline in FP32 (faster)")
    args = ap.parse_args(argv)
    N, P, pert = (64, 5_000, 6) if args.quick else (256, 50_000, 10)
    out = compare(D.Context(0), N, P, pert, args.train_instances, args.top_k,
                  fp64=not args.fp32_exhaustive)
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
