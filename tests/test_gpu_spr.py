"""GPU checks of the learned-SPR building blocks: batched exact scoring of SPR
neighbours against the FP64 oracle, and the MLP kernels (forward, gradients,
training) against the FP64 numpy oracle.

BF16 MFMA operands: forward outputs within 2e-2 relative to the output
scale, gradients within 5e-2 of the gradient norm."""

import numpy as np
import pytest

from oracle import mlp_oracle as MO
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from phylocaml_amd import device as D

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return D.Context(0)


def _instance(N, P, seed):
    from phylocaml_amd import mlmodel as MM, synth

    tree = synth.yule(N, seed, 0.08)
    pi, co = synth.gtr_params(4, 5.0, seed)
    model = MM.create(MM.Spec(4, ("GTR", co), ("DiscreteGamma", 4, 0.5), ("Empirical", pi)))
    st = synth.simulate(tree, model.priors, model.q, model.rates, P, seed)
    w = synth.poisson_weights(P, 3.0, seed)
    return tree, model, st, w


@pytest.mark.parametrize("form,tol", [(0, 1e-10), (1, 1e-6)])
def test_spr_scores_match_oracle(ctx, form, tol):
    from phylocaml_amd import device as D, likelihood as LK, spr, synth

    N, P = 12, 257
    tree, model, st, w = _instance(N, P, 5)
    mv = spr.moves(N, tree.children, tree.root)
    dm = LK.DeviceModel(ctx, model)
    tips = D.to_device(synth.pack_dna(st))
    sc = spr.Scorer(ctx, dm, N, tips, D.words(P), D.to_device(w), P, form=form, chunk=64)
    got = sc.score(tree.children, tree.root, tree.blen, mv).cpu().numpy()
    tips_m = (1 << st).astype(np.uint32)
    for k in range(len(mv)):
        ch, r, bl = spr.apply(N, tree.children, tree.root, tree.blen, *mv[k])
        ov, _ = O.lnl(N, ch, r, bl, model.priors, model.q, model.rates, model.probs, tips_m, w)
        assert abs(got[k] - ov) <= tol * abs(ov), (k, mv[k], got[k], ov)


def _mlp(ctx, n_in=32, H=128, seed=1):
    from phylocaml_amd import mlp

    return mlp.MLP(ctx, n_in, H, seed)


def test_mlp_forward_and_gradients(ctx, rng):
    from phylocaml_amd import device as D

    n_in, H, B = 32, 128, 300
    net = _mlp(ctx, n_in, H)
    X = rng.normal(0, 1, (B, n_in)).astype(np.float32)
    t = rng.normal(0, 1, B).astype(np.float32)
    p = net.get_params()
    y = net.forward(D.to_device(X)).cpu().numpy()
    yo, _ = MO.forward(p.astype(np.float64), X.astype(np.float64), n_in, H)
    assert np.max(np.abs(y - yo)) <= 2e-2 * (np.max(np.abs(yo)) + 1e-3)
    g, loss = net.gradients(D.to_device(X), D.to_device(t))
    lo, go = MO.loss_and_grad(p.astype(np.float64), X, t, n_in, H)
    assert abs(loss - lo) <= 2e-2 * abs(lo)
    assert np.linalg.norm(g - go) <= 5e-2 * np.linalg.norm(go)
    # each parameter block separately
    sizes = [n_in * H, H, H * H, H, H, 1]
    o = 0
    for s in sizes:
        a, b = g[o:o + s], go[o:o + s]
        assert np.linalg.norm(a - b) <= 6e-2 * (np.linalg.norm(b) + 1e-6), s
        o += s


def test_mlp_training_reduces_loss(ctx, rng):
    from phylocaml_amd import device as D

    n_in, B = 32, 4096
    net = _mlp(ctx, n_in, 128, seed=3)
    X = rng.normal(0, 1, (B, n_in)).astype(np.float32)
    wt = rng.normal(0, 1, n_in)
    t = (np.tanh(X @ wt / 4) + 0.1 * X[:, 0] ** 2).astype(np.float32)
    Xd, td = D.to_device(X), D.to_device(t)
    first = None
    for it in range(300):
        loss = net.step(Xd[(it % 4) * 1024:(it % 4 + 1) * 1024], td[(it % 4) * 1024:(it % 4 + 1) * 1024],
                        lr=2e-3)
        if first is None:
            first = loss
    _, final = net.gradients(Xd, td)
    assert final < 0.2 * first, (first, final)
