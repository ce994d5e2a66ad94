"""GPU parameter optimisation: the same optimiser driven by the GPU objective
(FP64 whole-tree kernels) and by the FP64 oracle follows the same path and
ends at the same branch lengths, parameters and lnL."""
import sys

import numpy as np
import pytest

from test_optimize import OracleObjective, _instance

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from phylocaml_amd import device as D

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return D.Context(0)


def _gpu_objective(ctx, tree, tips, w, N):
    from phylocaml_amd import device as D
    from phylocaml_amd.optimize import DeviceObjective

    topo = D.Topology(N, tree.children, tree.root, ctx)
    P = tips.shape[1]
    sets = tips.astype(np.uint8)
    tp = D.to_device(D.pack_nibbles(sets), ctx.device)
    return DeviceObjective(ctx, topo, tp, D.words(P), D.to_device(w, ctx.device), P)


@pytest.mark.parametrize("N,P,seed", [(10, 400, 3), (40, 1500, 7)])
def test_branch_newton_gpu_matches_oracle(ctx, N, P, seed):
    from phylocaml_amd.optimize import optimize_branches

    _say(f"branch Newton N={N} P={P}")
    tree, model, tips, w = _instance(N, P, seed)
    rng = np.random.default_rng(seed)
    start = tree.blen * rng.uniform(0.3, 3.0, tree.blen.size)
    og = optimize_branches(_gpu_objective(ctx, tree, tips, w, N), model, start, tree.children,
                           N, tree.root, tol=1e-10)
    oo = optimize_branches(OracleObjective(N, tree.children, tree.root, tips, w), model, start,
                           tree.children, N, tree.root, tol=1e-10)
    assert og.iterations == oo.iterations and og.derivative_passes == oo.derivative_passes
    assert abs(og.lnl - oo.lnl) <= 1e-10 * abs(oo.lnl)
    assert np.allclose(og.blen, oo.blen, rtol=1e-7, atol=1e-12)


def test_full_optimisation_gpu_matches_oracle(ctx):
    from phylocaml_amd.optimize import optimize

    N, P = 12, 600
    _say(f"model + branch optimisation N={N} P={P}")
    tree, model, tips, w = _instance(N, P, 11)
    g = optimize(_gpu_objective(ctx, tree, tips, w, N), model, tree.blen, tree.children, N,
                 tree.root, rounds=3)
    o = optimize(OracleObjective(N, tree.children, tree.root, tips, w), model, tree.blen,
                 tree.children, N, tree.root, rounds=3)
    assert abs(g.lnl - o.lnl) <= 1e-9 * abs(o.lnl)
    assert abs(g.model.spec.site_variation[2] - o.model.spec.site_variation[2]) <= 1e-4
    assert np.allclose(g.model.priors, o.model.priors, atol=1e-5)
    assert np.allclose(g.model.q, o.model.q, rtol=1e-4, atol=1e-6)
    assert np.allclose(g.blen, o.blen, rtol=1e-4, atol=1e-9)
