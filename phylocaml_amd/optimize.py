"""Branch-length and model-parameter optimisation on the GPU likelihood.

The reference declares the hooks but leaves them unimplemented:
``num_parameters`` / ``update`` (lib/mlModel.ml:193-196) and the
``gen_subst_opt_func`` / ``gen_rates_opt_func`` / ``gen_prior_opt_func``
generators (lib/mlModel.ml:821-829), each returning the current parameter
vector and a function from a vector to a new model.  The generators live in
``mlmodel``; this module drives them:

* ``optimize_branches``: simultaneous Newton-Raphson on every branch from one
  derivative pass (``phy_lk_tree`` with d1/d2).  A branch with d2 < 0 takes
  the Newton step -d1/d2; otherwise it moves half its length along the
  gradient.  Lengths are clamped to [BL_MIN, BL_MAX].  The two root branches
  form one unrooted edge: they report that edge's derivatives, take one step
  on the edge length and keep their ratio.  When a step lowers lnL it is
  halved (at most ``max_halvings`` times) before it is accepted.
* ``brent_max``: Brent's method (R. P. Brent, Algorithms for Minimization
  without Derivatives, 1973, ch. 5: golden-section steps with parabolic
  interpolation) for a maximum on an interval.
* ``optimize_model``: Brent along each coordinate of the (log-scale)
  parameter vectors in turn, rates, then substitution, then priors.
* ``optimize``: alternates the two until a round gains less than ``tol``.

The objective is any object with ``lnl(model, blen) -> float`` and
``derivs(model, blen) -> (lnl, d1, d2)``; ``DeviceObjective`` evaluates it
with the whole-tree kernels in FP64 (fused form for lnL, stored form with the
derivative pass for d1/d2).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import mlmodel as MM

BL_MIN = 1e-8
BL_MAX = 10.0
GOLD = 0.3819660112501051  # (3 - sqrt 5) / 2


class DeviceObjective:
    """lnL and branch derivatives of one topology and alignment on the GPU."""

    def __init__(self, ctx, topo, tips, stride: int, weights, n_patterns: int):
        from . import device as D
        from . import likelihood as LK

        self.ctx, self.topo = ctx, topo
        self.tips, self.stride, self.weights = tips, int(stride), weights
        self.n_patterns = int(n_patterns)
        self._D, self._LK = D, LK
        self._ev = {}  # (K, R, derivs) -> Evaluator
        self.n_lnl = 0
        self.n_derivs = 0

    def _eval(self, model, derivs: bool):
        LK = self._LK
        dm = LK.DeviceModel(self.ctx, model)
        key = (dm.K, dm.R, derivs)
        ev = self._ev.get(key)
        if ev is None:
            form = LK.PHY_F64 if derivs else LK.PHY_F64 | LK.PHY_FUSED
            ev = LK.Evaluator(self.ctx, self.topo, dm, self.n_patterns, form, derivs=derivs)
            self._ev[key] = ev
        ev.dm = dm  # same K and R: same workspace
        return ev

    def _blen(self, blen):
        return self._D.to_device(np.ascontiguousarray(blen, np.float64), self.ctx.device)

    def lnl(self, model, blen) -> float:
        ev = self._eval(model, False)
        ev.run(self.tips, self.stride, self.weights, self._blen(blen))
        self.n_lnl += 1
        return float(ev.lnl.cpu()[0])

    def derivs(self, model, blen):
        ev = self._eval(model, True)
        ev.run(self.tips, self.stride, self.weights, self._blen(blen))
        self.n_derivs += 1
        return float(ev.lnl.cpu()[0]), ev.d1.cpu().numpy().copy(), ev.d2.cpu().numpy().copy()


class DeviceNewton:
    """The Newton iteration of ``optimize_branches`` with its step control on
    the device (phy_newton_propose, phy_lk_tree with derivatives at the
    proposal, phy_newton_accept): an iteration issues kernels only, so
    ``capture(k)`` records k iterations in one HIP graph.  A proposal that
    lowers lnL is halved towards the current lengths on the next iteration,
    as ``optimize_branches`` does; there is no stopping test, so iterations
    past convergence leave the lengths in place.  Capture needs ``ctx`` on
    a side stream (not the default stream)."""

    def __init__(self, ctx, topo, dm, tips, stride: int, weights, n_patterns: int, blen):
        from . import device as D
        from . import likelihood as LK
        from ._lib import lib

        self.ctx, self.topo, self._D, self._lib = ctx, topo, D, lib
        self.ev = LK.Evaluator(ctx, topo, dm, n_patterns, LK.PHY_F64, derivs=True)
        self.inputs = (tips, int(stride), weights)
        t = np.asarray(blen, np.float64).copy()
        t[topo.root] = 0.0
        self.t = D.to_device(t, ctx.device)
        self.tn = D.to_device(t, ctx.device)
        self.d1 = D.to_device(np.zeros_like(t), ctx.device)
        self.d2 = D.to_device(np.zeros_like(t), ctx.device)
        self.state = D.to_device(np.array([-np.inf, 1.0, 0.0, 0.0]), ctx.device)
        # derivatives at the start: accepted as a proposal over lnL = -inf,
        # then the counters reset
        self.ev.run(*self.inputs, self.t)
        self._accept()
        ctx.check(lib.phy_newton_init(ctx.handle, D.dptr(self.state), D.dptr(self.ev.lnl)),
                  "newton_init")

    def _accept(self):
        D = self._D
        self.ctx.check(self._lib.phy_newton_accept(
            self.ctx.handle, self.topo.handle, D.dptr(self.state), D.dptr(self.ev.lnl),
            D.dptr(self.t), D.dptr(self.tn), D.dptr(self.d1), D.dptr(self.ev.d1), D.dptr(self.d2),
            D.dptr(self.ev.d2)), "newton_accept")

    def step(self) -> None:
        """Enqueue one iteration (propose, derivative pass, accept)."""
        D = self._D
        self.ctx.check(self._lib.phy_newton_propose(
            self.ctx.handle, self.topo.handle, D.dptr(self.state), D.dptr(self.t), D.dptr(self.d1),
            D.dptr(self.d2), D.dptr(self.tn)), "newton_propose")
        self.ev.run(*self.inputs, self.tn)
        self._accept()

    def run(self, k: int) -> None:
        for _ in range(int(k)):
            self.step()

    def capture(self, k: int):
        """A torch.cuda.CUDAGraph of k iterations on the context's stream;
        each replay advances the state by k iterations."""
        torch = __import__("torch")
        if not self.ctx.stream:
            raise ValueError("graph capture needs the context on a side stream")
        s = torch.cuda.ExternalStream(self.ctx.stream, device=f"cuda:{self.ctx.device}")
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self.run(k)
        return g

    def result(self):
        """(lengths, lnL, accepted steps, halvings) after a synchronisation."""
        self.ctx.sync()
        st = self.state.cpu().numpy()
        return self.t.cpu().numpy(), float(st[0]), int(st[2]), int(st[3])


def _root_children(children: np.ndarray, n_leaves: int, root: int):
    k = 2 * (root - n_leaves)
    return int(children[k]), int(children[k + 1])


def newton_update(blen: np.ndarray, d1: np.ndarray, d2: np.ndarray, children: np.ndarray,
                  n_leaves: int, root: int) -> np.ndarray:
    """One simultaneous Newton-Raphson step on every branch (module doc)."""
    t = np.asarray(blen, np.float64).copy()
    a, b = _root_children(children, n_leaves, root)

    def step(x, g, h):
        if h < 0.0:
            return x - g / h
        return x + (0.5 * x if g > 0 else -0.5 * x)

    for i in range(t.size):
        if i == root or i == a or i == b:
            continue
        t[i] = min(max(step(t[i], d1[i], d2[i]), BL_MIN), BL_MAX)
    T = t[a] + t[b]
    Tn = min(max(step(T, d1[a], d2[a]), 2 * BL_MIN), 2 * BL_MAX)
    fa = t[a] / T if T > 0 else 0.5
    t[a], t[b] = max(Tn * fa, BL_MIN), max(Tn * (1 - fa), BL_MIN)
    t[root] = 0.0
    return t


@dataclass
class BranchResult:
    blen: np.ndarray
    lnl: float
    iterations: int
    derivative_passes: int
    max_abs_d1: float  # over branches not held at a bound


def optimize_branches(obj, model, blen, children, n_leaves: int, root: int,
                      max_iter: int = 30, tol: float = 1e-8, max_halvings: int = 10) -> BranchResult:
    """Newton-Raphson on all branch lengths (module doc).  Stops when a step
    gains less than tol * |lnL| or after max_iter derivative passes."""
    t = np.asarray(blen, np.float64).copy()
    t[root] = 0.0
    lnl, d1, d2 = obj.derivs(model, t)
    passes = 1
    it = 0
    for it in range(1, max_iter + 1):
        tn = newton_update(t, d1, d2, children, n_leaves, root)
        ln_new, d1n, d2n = obj.derivs(model, tn)
        passes += 1
        h = 0
        while ln_new < lnl and h < max_halvings:
            tn = 0.5 * (t + tn)
            ln_new, d1n, d2n = obj.derivs(model, tn)
            passes += 1
            h += 1
        if ln_new < lnl:  # no improving step along the Newton direction
            break
        gain = ln_new - lnl
        t, lnl, d1, d2 = tn, ln_new, d1n, d2n
        if gain <= tol * abs(lnl):
            break
    # branches held at a bound by a gradient pointing outwards are not free
    free = np.ones(t.size, bool)
    free[root] = False
    free &= ~((t <= BL_MIN * 1.0000001) & (d1 < 0)) & ~((t >= BL_MAX) & (d1 > 0))
    return BranchResult(t, lnl, it, passes, float(np.max(np.abs(d1[free]), initial=0.0)))


def brent_max(f, lo: float, hi: float, x0: float | None = None, tol: float = 1e-5,
              max_iter: int = 100):
    """Maximum of f on [lo, hi] by Brent's method (module doc); returns
    (x, f(x), evaluations)."""
    a, b = lo, hi
    x = w = v = (x0 if x0 is not None and lo < x0 < hi else a + GOLD * (b - a))
    fx = fw = fv = -f(x)
    n = 1
    d = e = 0.0
    for _ in range(max_iter):
        m = 0.5 * (a + b)
        tol1 = tol * abs(x) + 1e-10
        tol2 = 2.0 * tol1
        if abs(x - m) <= tol2 - 0.5 * (b - a):
            break
        use_golden = True
        if abs(e) > tol1:
            r = (x - w) * (fx - fv)
            q = (x - v) * (fx - fw)
            p = (x - v) * q - (x - w) * r
            q = 2.0 * (q - r)
            if q > 0:
                p = -p
            q = abs(q)
            if abs(p) < abs(0.5 * q * e) and q * (a - x) < p < q * (b - x):
                e, d = d, p / q
                u = x + d
                if u - a < tol2 or b - u < tol2:
                    d = tol1 if m > x else -tol1
                use_golden = False
        if use_golden:
            e = (a - x) if x >= m else (b - x)
            d = GOLD * e
        u = x + d if abs(d) >= tol1 else x + (tol1 if d > 0 else -tol1)
        fu = -f(u)
        n += 1
        if fu <= fx:
            if u >= x:
                a = x
            else:
                b = x
            v, w, x = w, x, u
            fv, fw, fx = fw, fx, fu
        else:
            if u < x:
                a = u
            else:
                b = u
            if fu <= fw or w == x:
                v, w = w, u
                fv, fw = fw, fu
            elif fu <= fv or v == x or v == w:
                v, fv = u, fu
    return x, -fx, n


# log-scale search window around each parameter's current value
LOG_WINDOW = 4.0


def optimize_model(obj, model, blen, tol: float = 1e-5, which=("rates", "subst", "priors")):
    """Brent along each coordinate of the chosen parameter vectors, in turn;
    returns (model, lnL, evaluations)."""
    gens = {"rates": MM.gen_rates_opt_func, "subst": MM.gen_subst_opt_func,
            "priors": MM.gen_prior_opt_func}
    lnl = obj.lnl(model, blen)
    evals = 1
    for name in which:
        vec, make = gens[name](model)
        vec = np.array(vec, np.float64)
        for i in range(vec.size):
            def f(xi, i=i):
                v = vec.copy()
                v[i] = xi
                return obj.lnl(make(v), blen)

            lo, hi = vec[i] - LOG_WINDOW, vec[i] + LOG_WINDOW
            if name == "rates":  # log alpha: keep the Gamma shape in [1e-3, 1e3]
                lo, hi = max(lo, math.log(1e-3)), min(hi, math.log(1e3))
            x, fx, n = brent_max(f, lo, hi, x0=vec[i], tol=tol)
            evals += n
            if fx > lnl:
                vec[i], lnl = x, fx
                model = make(vec)
    return model, lnl, evals


@dataclass
class OptimizeResult:
    model: object
    blen: np.ndarray
    lnl: float
    rounds: int
    lnl_evaluations: int
    derivative_passes: int


def optimize(obj, model, blen, children, n_leaves: int, root: int, rounds: int = 5,
             tol: float = 1e-6, which=("rates", "subst", "priors")) -> OptimizeResult:
    """Alternate branch-length Newton and model-parameter Brent until a round
    gains less than tol * |lnL|."""
    t = np.asarray(blen, np.float64).copy()
    lnl = -math.inf
    evals = passes = 0
    r = 0
    for r in range(1, rounds + 1):
        br = optimize_branches(obj, model, t, children, n_leaves, root)
        t = br.blen
        passes += br.derivative_passes
        model, ln_new, n = optimize_model(obj, model, t, which=which)
        evals += n
        gain = ln_new - lnl
        lnl = ln_new
        if gain <= tol * abs(lnl):
            break
    return OptimizeResult(model, t, lnl, r, evals, passes)


def main(argv=None):
    """Optimise branch lengths and GTR+Gamma parameters of a simulated
    alignment from a perturbed start on the GPU; prints one JSON line."""
    import argparse
    import json
    import sys
    import time

    from . import device as D
    from . import synth

    ap = argparse.ArgumentParser()
    ap.add_argument("--taxa", type=int, default=512)
    ap.add_argument("--patterns", type=int, default=200_000)
    ap.add_argument("--rounds", type=int, default=10)
    ap.add_argument("--seed", type=int, default=2)
    args = ap.parse_args(argv)

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    N, P = args.taxa, args.patterns
    tree = synth.yule(N, args.seed, 0.05)
    pi, co = synth.gtr_params(4, 5.0, args.seed)
    truth = MM.create(MM.Spec(4, ("GTR", co), ("DiscreteGamma", 4, 0.5), ("Empirical", pi)))
    st = synth.simulate(tree, truth.priors, truth.q, truth.rates, P, args.seed, unique=True)
    w = synth.poisson_weights(P, 3.0, args.seed)
    tips = np.where(st < 4, 1 << np.minimum(st, 3), 15).astype(np.uint8)
    ctx = D.Context(0)
    topo = D.Topology(N, tree.children, tree.root, ctx)
    obj = DeviceObjective(ctx, topo, D.to_device(D.pack_nibbles(tips), ctx.device), D.words(P),
                          D.to_device(w, ctx.device), P)
    # start: GTR with equal exchangeabilities and frequencies, alpha 1, every
    # branch length 0.1
    start = MM.create(MM.Spec(4, ("GTR", np.ones(5)), ("DiscreteGamma", 4, 1.0),
                              ("Empirical", np.full(4, 0.25))))
    t0 = np.full(2 * N - 1, 0.1)
    t0[tree.root] = 0.0
    l_start = obj.lnl(start, t0)
    l_true = obj.lnl(truth, tree.blen)
    log(f"start lnL {l_start:.4f}, generating model at true lengths {l_true:.4f}")
    torch = __import__("torch")
    torch.cuda.synchronize()
    w0 = time.time()
    br = optimize_branches(obj, start, t0, tree.children, N, tree.root)
    torch.cuda.synchronize()
    t_br = time.time() - w0
    log(f"branch lengths: lnL {br.lnl:.4f} after {br.iterations} Newton steps "
        f"({br.derivative_passes} derivative passes) in {t_br:.2f}s")
    w0 = time.time()
    res = optimize(obj, start, t0, tree.children, N, tree.root, rounds=args.rounds)
    torch.cuda.synchronize()
    t_all = time.time() - w0
    log(f"full: lnL {res.lnl:.4f} after {res.rounds} rounds in {t_all:.2f}s")
    out = {
        "workload": f"{N}-taxon Yule tree, {P} unique GTR+G4 patterns (weights Poisson(3)+1); "
                    "start: equal GTR rates and frequencies, alpha 1, all branches 0.1",
        "lnl_start": l_start, "lnl_generating_model": l_true,
        "branches_only": {"lnl": br.lnl, "newton_steps": br.iterations,
                          "derivative_passes": br.derivative_passes, "wall_s": t_br,
                          "ms_per_derivative_pass": 1e3 * t_br / br.derivative_passes,
                          "max_abs_d1_free": br.max_abs_d1},
        "full": {"lnl": res.lnl, "rounds": res.rounds, "lnl_evaluations": res.lnl_evaluations,
                 "derivative_passes": res.derivative_passes, "wall_s": t_all,
                 "alpha": res.model.spec.site_variation[2],
                 "alpha_generating": 0.5, "priors": res.model.priors.tolist(),
                 "priors_generating": truth.priors.tolist(),
                 "exchangeabilities": list(np.asarray(res.model.spec.substitution[1])),
                 "exchangeabilities_generating": list(np.asarray(co))},
    }
    print(json.dumps(out), flush=True)
    return out



__all__ = ["DeviceObjective", "newton_update", "optimize_branches", "brent_max",
           "optimize_model", "optimize", "BranchResult", "OptimizeResult", "BL_MIN", "BL_MAX"]


if __name__ == "__main__":
    main()
