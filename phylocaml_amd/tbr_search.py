"""Learned TBR proposals: the SPR proposal of ``spr_search`` extended to the
TBR neighbourhood (``neighborhood.Neighborhood("tbr", join, reroot)``).

A TBR delta (s, w, v) reroots the subtree of s on the edge above w and
regrafts it on the edge above v (v = -1: where it was).  Its features are the
32 SPR features of the regraft (s, v), with v = -1 read as the original
attachment, plus 7 of the reroot: reroot depth, the size and branch length
at w, and the keep / reroot flags.  The MLP, its training on exact scores
and the guided search (the top k re-scored exactly in FP64, an exhaustive
round when none of them improves) are those of ``spr_search``.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import device as D
from . import neighborhood as NB
from . import spr
from .spr_search import Context, F32_FUSED, make_instance, target, train

N_TBR_EXTRA = 7
TBR_EXTRA_NAMES = ["reroot depth / D", "log1p size(w)", "size(w) / size(s)", "t_w",
                   "log t_w", "keeps its attachment", "rerooted"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def features(c: Context, ch, r, bl, mv: np.ndarray) -> np.ndarray:
    """float32 [n, 32 + 7] features of TBR triples mv [n, 3]."""
    N = c.inst.N
    info = spr.TreeInfo(N, ch, r, bl)
    topo = D.Topology(N, ch, r, c.ctx)
    ft = spr.fitch_sets(c.ctx, topo, c.tips, c.inst.P, c.wd)
    s, w, v = (mv[:, k].astype(np.int64) for k in range(3))
    p = info.parent[s]
    chp = info.children[p - N]
    x = np.where(chp[:, 0] == s, chp[:, 1], chp[:, 0])
    keep = v < 0
    # a kept attachment is read as the regraft onto the sibling's edge (which
    # rebuilds the same unrooted placement)
    v_eff = np.where(keep, x, v)
    base = spr.features(c.ctx, info, ft, c.inst.P, c.wd,
                        np.stack([s, v_eff], 1).astype(np.int32), c.total_w)
    Dmax = max(1, int(info.depth.max()))
    d = (info.depth[w] - info.depth[s]) / Dmax
    extra = np.stack([d, np.log1p(info.size[w]), info.size[w] / info.size[s], info.blen[w],
                      np.log(info.blen[w] + 1e-6), keep.astype(np.float64),
                      (w != s).astype(np.float64)], 1).astype(np.float32)
    return np.concatenate([base, extra], 1)


def training_pairs(ctx, seeds, N, P, nb: NB.Neighborhood, n_perturb=10):
    X, T = [], []
    for sd in seeds:
        t0 = time.time()
        inst = make_instance(sd, N, P, n_perturb)
        c = Context(ctx, inst)
        sc32 = NB.Scorer(ctx, c.dm, N, c.tips, c.stride, c.wd, P, form=1)
        mv = nb.neighborhood((N, inst.children, inst.root))
        base = c.lnl(inst.children, inst.root, inst.blen, F32_FUSED)
        sc = sc32.score(inst.children, inst.root, inst.blen, mv).cpu().numpy()
        X.append(features(c, inst.children, inst.root, inst.blen, mv))
        T.append(target(sc - base).astype(np.float32))
        log(f"  instance {sd}: {len(mv)} moves in {time.time() - t0:.1f}s "
            f"(best dlnL {np.max(sc - base):.2f})")
    return np.concatenate(X), np.concatenate(T)


def search(c: Context, nb: NB.Neighborhood, net, norm, guided: bool, k: int = 64,
           eps: float = 1e-6, max_rounds: int = 300, fp64: bool = True) -> dict:
    inst = c.inst
    N = inst.N
    ex = NB.Scorer(c.ctx, c.dm, N, c.tips, c.stride, c.wd, inst.P, form=0 if fp64 else 1)
    ch, r, bl = inst.children.copy(), inst.root, inst.blen.copy()
    c.n_exact = 0
    n_exact = 0
    cur = c.lnl(ch, r, bl)
    t0 = time.time()
    rounds, fallbacks, hits, tries = 0, 0, 0, 0
    while rounds < max_rounds:
        mv = nb.neighborhood((N, ch, r))
        rounds += 1
        if guided:
            pred = net.forward(D.to_device(norm(features(c, ch, r, bl, mv)))).cpu().numpy()
            top = np.argpartition(-pred, min(k, len(mv) - 1))[:k]
            sc = ex.score(ch, r, bl, mv[top]).cpu().numpy()
            n_exact += len(top)
            j = int(np.argmax(sc))
            if sc[j] > cur + eps * abs(cur):
                ch, r, bl = NB.apply(N, ch, r, bl, mv[top[j]])
                cur = float(sc[j])
                continue
            fallbacks += 1
        sc = ex.score(ch, r, bl, mv).cpu().numpy()
        n_exact += len(mv)
        j = int(np.argmax(sc))
        if not guided and net is not None:
            pred = net.forward(D.to_device(norm(features(c, ch, r, bl, mv)))).cpu().numpy()
            top = np.argpartition(-pred, min(k, len(mv) - 1))[:k]
            tries += 1
            hits += int(j in set(top.tolist()))
        if sc[j] <= cur + eps * abs(cur):
            break
        ch, r, bl = NB.apply(N, ch, r, bl, mv[j])
        cur = float(sc[j])
        log(f"  {'guided' if guided else 'exhaustive'} round {rounds}: lnL {cur:.3f}, "
            f"{n_exact} exact evaluations, {time.time() - t0:.1f}s")
    return {"final_lnl": cur, "rounds": rounds, "exact_evaluations": n_exact,
            "wall_s": time.time() - t0, "fallback_rounds": fallbacks,
            "best_move_recall_top_k": (hits / tries) if tries else None}


def main(argv=None):
    """Train on the TBR neighbourhoods of --train-instances instances, then
    run exhaustive and guided TBR searches from each held-out start
    (--eval-seeds); prints one JSON line with per-instance results and the
    totals."""
    ap = argparse.ArgumentParser()
    ap.add_argument("--taxa", type=int, default=128)
    ap.add_argument("--patterns", type=int, default=20_000)
    ap.add_argument("--train-instances", type=int, default=8)
    ap.add_argument("--eval-seeds", default="7,8,9,10")
    ap.add_argument("--join", type=int, default=3)
    ap.add_argument("--reroot", type=int, default=3)
    ap.add_argument("--top-k", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=4)
    args = ap.parse_args(argv)
    N, P = args.taxa, args.patterns
    nb = NB.Neighborhood("tbr", args.join, args.reroot)
    ctx = D.Context(0)
    log(f"training data: {args.train_instances} instances, N={N}, P={P}")
    t0 = time.time()
    X, T = training_pairs(ctx, range(200, 200 + args.train_instances), N, P, nb)
    t_data = time.time() - t0
    t0 = time.time()
    net, norm, losses = train(ctx, X, T, epochs=args.epochs)
    t_train = time.time() - t0
    runs = []
    for sd in (int(x) for x in args.eval_seeds.split(",")):
        inst = make_instance(sd, N, P, 10)
        c = Context(ctx, inst)
        start = c.lnl(inst.children, inst.root, inst.blen)
        log(f"instance {sd}: start lnL {start:.3f}; exhaustive search")
        exs = search(c, nb, net, norm, guided=False, k=args.top_k)
        log(json.dumps(exs))
        log("guided search")
        gd = search(c, nb, net, norm, guided=True, k=args.top_k)
        log(json.dumps(gd))
        runs.append({"seed": sd, "start_lnl": start, "exhaustive": exs, "guided": gd,
                     "final_lnl_gap": gd["final_lnl"] - exs["final_lnl"],
                     "gain_ratio": (gd["final_lnl"] - start) / max(exs["final_lnl"] - start, 1e-9)})
    ev_ex = sum(r["exhaustive"]["exact_evaluations"] for r in runs)
    ev_gd = sum(r["guided"]["exact_evaluations"] for r in runs)
    t_ex = sum(r["exhaustive"]["wall_s"] for r in runs)
    t_gd = sum(r["guided"]["wall_s"] for r in runs)
    out = {"workload": f"{N}-taxon Yule trees, {P} unique GTR+G4 patterns, start = true tree "
                       f"after 10 random SPR moves; TBR neighbourhood, join <= {args.join}, "
                       f"reroot <= {args.reroot}; FP64 exact scores",
           "training_pairs": int(len(X)), "training_data_s": t_data, "training_s": t_train,
           "features": 32 + N_TBR_EXTRA, "train_loss_per_epoch": losses, "top_k": args.top_k,
           "instances": runs,
           "exact_evaluation_ratio": ev_ex / max(ev_gd, 1),
           "wall_time_ratio": t_ex / max(t_gd, 1e-9),
           "mean_final_lnl_gap": float(np.mean([r["final_lnl_gap"] for r in runs])),
           "mean_gain_ratio": float(np.mean([r["gain_ratio"] for r in runs]))}
    print(json.dumps(out), flush=True)
    return out


__all__ = ["features", "training_pairs", "search", "main", "TBR_EXTRA_NAMES"]


if __name__ == "__main__":
    main()
