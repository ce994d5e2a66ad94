"""End-to-end learned SPR search on a small instance: the MLP trained on
exact-ΔlnL pairs of other instances ranks moves so that hill climbing with a
top-k exact re-score needs far fewer exact evaluations than exhaustive
scoring, and both stop at a tree with no improving SPR move."""
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_guided_search_needs_fewer_exact_evaluations():
    from phylocaml_amd import device as D, spr_search as S

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    ctx = D.Context(0)
    N, P = 32, 2000
    X, T = S.training_pairs(ctx, [11, 12, 13, 14], N, P, n_perturb=5)
    assert X.shape[1] == 32 and len(X) == len(T)
    net, norm, losses = S.train(ctx, X, T, epochs=4, batch=256, lr=2e-3)
    assert losses[-1] < losses[0]
    inst = S.make_instance(21, N, P, 5)
    c = S.Context(ctx, inst)
    ex = S.search(c, net, norm, guided=False, k=16)
    gd = S.search(c, net, norm, guided=True, k=16)
    print("exhaustive", ex, "guided", gd, flush=True)
    assert ex["exact_evaluations"] >= 2 * gd["exact_evaluations"]
    start = c.lnl(inst.children, inst.root, inst.blen)
    gain_ex = ex["final_lnl"] - start
    # the guided search recovers nearly all of the exhaustive search's gain
    assert gd["final_lnl"] - start >= 0.9 * gain_ex
