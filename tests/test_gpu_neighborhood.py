"""Exact GPU scores of NNI, SPR and TBR neighbours against the FP64 oracle
on the trees the moves produce (FP64 within 1e-10, FP32-scaled within 1e-6
relative), and a best-improvement search that ends at a local optimum."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

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


@pytest.fixture(scope="module")
def inst():
    from phylocaml_amd.spr_search import make_instance

    return make_instance(5, 14, 301, 3)


def _scorer(ctx, inst, form):
    from phylocaml_amd import device as D, likelihood as LK, neighborhood as NB

    dm = LK.DeviceModel(ctx, inst.model)
    return NB.Scorer(ctx, dm, inst.N, D.to_device(inst.tips_np), inst.tips_np.shape[1],
                     D.to_device(inst.weights), inst.P, form=form, chunk=64)


def _oracle_lnl(inst, ch, r, bl):
    from phylocaml_amd import device as D

    st = D.unpack_nibbles(inst.tips_np, inst.P).astype(np.uint32)
    m = inst.model
    return O.lnl(inst.N, ch, r, bl, m.priors, m.q, m.rates, m.probs, st, inst.weights)[0]


@pytest.mark.parametrize("kind", ["nni", "spr", "tbr"])
@pytest.mark.parametrize("form,tol", [(0, 1e-10), (1, 1e-6)])
def test_neighbour_scores_match_oracle(ctx, inst, kind, form, tol):
    from phylocaml_amd import neighborhood as NB

    _say(f"neighbour scores {kind} form={form}")
    nb = NB.Neighborhood(kind, max_join=4, max_reroot=3)
    mv = nb.neighborhood((inst.N, inst.children, inst.root))
    assert len(mv) > 0
    got = _scorer(ctx, inst, form).score(inst.children, inst.root, inst.blen, mv).cpu().numpy()
    idx = np.unique(np.linspace(0, len(mv) - 1, 60).astype(int))
    for i in idx:
        ch, r, bl = NB.apply(inst.N, inst.children, inst.root, inst.blen, mv[i])
        ov = _oracle_lnl(inst, ch, r, bl)
        assert abs(got[i] - ov) <= tol * abs(ov), (kind, mv[i], got[i], ov)


def test_hill_climb_ends_at_local_optimum(ctx, inst):
    from phylocaml_amd import neighborhood as NB

    sc = _scorer(ctx, inst, 0)
    base = _oracle_lnl(inst, inst.children, inst.root, inst.blen)
    nb = NB.Neighborhood("tbr", max_join=3, max_reroot=2)
    res = NB.hill_climb(sc, base, inst.N, inst.children, inst.root, inst.blen, nb)
    assert res["final_lnl"] >= base
    assert abs(_oracle_lnl(inst, res["children"], res["root"], res["blen"]) - res["final_lnl"]) \
        <= 1e-10 * abs(base)
    mv = nb.neighborhood((inst.N, res["children"], res["root"]))
    last = sc.score(res["children"], res["root"], res["blen"], mv).cpu().numpy()
    assert last.max() <= res["final_lnl"]
