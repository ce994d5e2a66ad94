"""Incremental SPR scoring (phy_lk_spr_scores_incr) on the GPU.

Every regraft of a pruned subtree is scored from one repair of the base
tree's partials and one pre-order pass over the pruned tree.  Checked
against the FP64 oracle evaluating each candidate tree (phy_spr_apply) in
full -- FP64 within 1e-10 relative, FP32 within 1e-6 -- and against the
per-candidate scorer (phy_lk_spr_scores) on every move of larger trees; the
base tree's lnL against the oracle; rate-class counts 1, 2, 4 and 8, zero
rate classes, long repair paths (caterpillars), the 3-taxon tree, moves in
any order with repeats, batches that do not divide the subtrees, and
invalid moves."""
import numpy as np
import pytest

from conftest import caterpillar, random_tree
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

F64, F32 = 0, 1


@pytest.fixture(scope="module")
def ctx():
    from phylocaml_amd import device as D

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return D.Context(0)


def _say(msg):
    print(msg, flush=True)


def _model(seed, rates="gamma4", ref=False):
    from phylocaml_amd import mlmodel as MM, synth

    pi, co = synth.gtr_params(4, 5.0, seed)
    sv = {"gamma4": ("DiscreteGamma", 4, 0.5), "gamma2": ("DiscreteGamma", 2, 0.7),
          "gamma8": ("DiscreteGamma", 8, 0.4), "constant": ("Constant",),
          "custom0": ("DiscreteCustom", [(0.0, 0.2), (0.5, 0.3), (1.2, 0.3), (2.5, 0.2)])}[rates]
    mode = MM.PHY_GAMMA_REF if ref else MM.PHY_GAMMA_MEAN
    return MM.create(MM.Spec(4, ("GTR", co), sv, ("Empirical", pi), mode))


def _data(tree, model, P, seed):
    from phylocaml_amd import synth

    sim_rates = np.where(model.rates > 0, model.rates, 1.0)
    st = synth.simulate(tree, model.priors, model.q, sim_rates, P, seed)
    w = synth.poisson_weights(P, 3.0, seed)
    return st, w


def _scorers(ctx, model, st, w, N, form, s_batch=16, chunk=256):
    from phylocaml_amd import device as D, likelihood as LK, spr, synth

    P = st.shape[1]
    dm = LK.DeviceModel(ctx, model)
    tips = D.to_device(synth.pack_dna(st))
    wd = D.to_device(w)
    inc = spr.IncrementalScorer(ctx, dm, N, tips, D.words(P), wd, P, form=form, s_batch=s_batch)
    full = spr.Scorer(ctx, dm, N, tips, D.words(P), wd, P, form=form, chunk=chunk)
    return inc, full, dm


def _oracle_all(N, ch, root, bl, model, st, w, mv):
    from phylocaml_amd import spr

    tips_m = (1 << st).astype(np.uint32)
    out = np.empty(len(mv))
    for k in range(len(mv)):
        c2, r2, b2 = spr.apply(N, ch, root, bl, *mv[k])
        out[k], _ = O.lnl(N, c2, r2, b2, model.priors, model.q, model.rates, model.probs, tips_m, w)
    return out


def _tree(kind, N, seed):
    from phylocaml_amd import synth

    rng = np.random.default_rng(seed)
    if kind == "yule":
        t = synth.yule(N, seed, 0.08)
        return t.children, t.root, t.blen
    ch, root = caterpillar(N) if kind == "caterpillar" else random_tree(rng, N)
    bl = rng.exponential(0.08, 2 * N - 1)
    bl[root] = 0.0
    return ch, root, bl


@pytest.mark.parametrize("form,tol", [(F64, 1e-10), (F32, 1e-6)])
def test_incremental_matches_oracle(ctx, form, tol):
    from phylocaml_amd import spr, synth

    N, P = 12, 257
    tree = synth.yule(N, 5, 0.08)
    model = _model(5)
    st, w = _data(tree, model, P, 5)
    mv = spr.moves(N, tree.children, tree.root)
    inc, _, _ = _scorers(ctx, model, st, w, N, form, s_batch=5)
    _say(f"incremental 12 x 257 form {form}: {len(mv)} moves")
    got = inc.score(tree.children, tree.root, tree.blen, mv).cpu().numpy()
    ref = _oracle_all(N, tree.children, tree.root, tree.blen, model, st, w, mv)
    err = np.abs(got - ref) / np.abs(ref)
    assert err.max() <= tol, (int(err.argmax()), mv[err.argmax()], got[err.argmax()], ref[err.argmax()])
    base, _ = O.lnl(N, tree.children, tree.root, tree.blen, model.priors, model.q, model.rates,
                    model.probs, (1 << st).astype(np.uint32), w)
    assert abs(float(inc.base_lnl.item()) - base) <= tol * abs(base)


@pytest.mark.parametrize("rates,ref,kind", [("constant", False, "random"), ("gamma2", False, "yule"),
                                            ("gamma8", False, "random"), ("gamma4", True, "yule"),
                                            ("custom0", True, "caterpillar")])
def test_incremental_rate_classes_and_shapes(ctx, rates, ref, kind):
    """R = 1, 2, 4, 8, the reference's own Gamma categories (rate 0 first)
    and an invariant class, on random, Yule and caterpillar trees."""
    from phylocaml_amd import spr, synth

    N, P = 18, 700
    ch, root, bl = _tree(kind, N, 11)
    model = _model(11, rates, ref)
    st, w = _data(synth.Tree(N, ch, root, bl), model, P, 11)
    mv = spr.moves(N, ch, root)
    ref_v = _oracle_all(N, ch, root, bl, model, st, w, mv)
    for form, tol in ((F64, 1e-10), (F32, 1e-6)):
        inc, _, _ = _scorers(ctx, model, st, w, N, form, s_batch=4)
        _say(f"incremental {rates} {kind} form {form}: {len(mv)} moves")
        got = inc.score(ch, root, bl, mv).cpu().numpy()
        err = np.abs(got - ref_v) / np.abs(ref_v)
        assert np.all(np.isfinite(got))
        assert err.max() <= tol, (rates, kind, form, mv[err.argmax()], got[err.argmax()],
                                  ref_v[err.argmax()])


@pytest.mark.parametrize("kind", ["random", "yule", "caterpillar"])
def test_incremental_few_targets_per_subtree(ctx, kind):
    """Short move lists (the NNI moves, and two random regrafts per pruned
    subtree): each pruned tree's walk then enters only the nodes above its
    targets and reads the partials of the subtrees it skips."""
    from phylocaml_amd import neighborhood as NB, spr, synth

    N, P = 20, 500
    ch, root, bl = _tree(kind, N, 17)
    model = _model(17)
    st, w = _data(synth.Tree(N, ch, root, bl), model, P, 17)
    full = spr.moves(N, ch, root)
    rng = np.random.default_rng(5)
    pick = []
    for s in np.unique(full[:, 0]):
        rows = np.flatnonzero(full[:, 0] == s)
        pick += list(rng.choice(rows, min(2, len(rows)), replace=False))
    mv = np.concatenate([NB.nni_moves(N, ch, root), full[np.array(pick)]])
    ref_v = _oracle_all(N, ch, root, bl, model, st, w, mv)
    for form, tol in ((F64, 1e-10), (F32, 1e-6)):
        inc, _, _ = _scorers(ctx, model, st, w, N, form, s_batch=3)
        _say(f"incremental few targets {kind} form {form}: {len(mv)} moves")
        got = inc.score(ch, root, bl, mv).cpu().numpy()
        err = np.abs(got - ref_v) / np.abs(ref_v)
        assert err.max() <= tol, (kind, form, mv[err.argmax()], got[err.argmax()], ref_v[err.argmax()])


def test_incremental_three_taxa(ctx):
    from phylocaml_amd import spr, synth

    N, P = 3, 64
    ch = np.array([1, 2, 0, 3], np.int32)  # node 3 = (1, 2), root 4 = (0, 3)
    root = 4
    bl = np.array([0.1, 0.2, 0.3, 0.05, 0.0])
    model = _model(3)
    st, w = _data(synth.Tree(N, ch, root, bl), model, P, 3)
    mv = spr.moves(N, ch, root)
    assert len(mv) > 0
    ref_v = _oracle_all(N, ch, root, bl, model, st, w, mv)
    inc, _, _ = _scorers(ctx, model, st, w, N, F64)
    got = inc.score(ch, root, bl, mv).cpu().numpy()
    assert np.all(np.abs(got - ref_v) <= 1e-10 * np.abs(ref_v))


@pytest.mark.parametrize("form,tol", [(F64, 1e-10), (F32, 2e-6)])
def test_incremental_matches_full_scorer_every_move(ctx, form, tol):
    """Every SPR move of a 64-taxon tree over 3,000 patterns, in shuffled
    order with repeats, 7 pruned subtrees per batch, against the per-candidate
    scorer in FP64."""
    from phylocaml_amd import spr, synth

    N, P = 64, 3000
    tree = synth.yule(N, 21, 0.08)
    model = _model(21)
    st, w = _data(tree, model, P, 21)
    mv = spr.moves(N, tree.children, tree.root)
    rng = np.random.default_rng(0)
    mv2 = np.concatenate([mv, mv[rng.choice(len(mv), 500)]])
    mv2 = mv2[rng.permutation(len(mv2))]
    inc, _, _ = _scorers(ctx, model, st, w, N, form, s_batch=7)
    _, full64, _ = _scorers(ctx, model, st, w, N, F64, chunk=1024)
    _say(f"incremental vs full 64 x 3000 form {form}: {len(mv2)} moves")
    got = inc.score(tree.children, tree.root, tree.blen, mv2).cpu().numpy()
    ref_v = full64.score(tree.children, tree.root, tree.blen, mv2).cpu().numpy()
    err = np.abs(got - ref_v) / np.abs(ref_v)
    assert err.max() <= tol, (mv2[err.argmax()], got[err.argmax()], ref_v[err.argmax()])
    # a subset, and a second call on the same workspace
    sub = mv2[::37]
    got2 = inc.score(tree.children, tree.root, tree.blen, sub).cpu().numpy()
    assert np.array_equal(got2, got[::37])


def test_incremental_invalid_moves(ctx):
    from phylocaml_amd import _lib, spr, synth

    N, P = 10, 100
    tree = synth.yule(N, 2, 0.1)
    model = _model(2)
    st, w = _data(tree, model, P, 2)
    inc, _, _ = _scorers(ctx, model, st, w, N, F64)
    mv = spr.moves(N, tree.children, tree.root)
    p = np.zeros(2 * N - 1, np.int64)
    p[tree.children] = np.repeat(np.arange(N, 2 * N - 1), 2)
    s = int(tree.children[0])
    bad = [(tree.root, 0), (s, s), (s, int(p[s])), (-1, 0), (0, 2 * N - 1)]
    for b in bad:
        with pytest.raises(_lib.PhyError):
            inc.score(tree.children, tree.root, tree.blen,
                      np.concatenate([mv[:3], np.array([b], np.int32)]))
    # the scorer still works after the errors
    got = inc.score(tree.children, tree.root, tree.blen, mv[:5]).cpu().numpy()
    assert np.all(np.isfinite(got))


def test_plan_runs_equal_the_one_shot_scores_and_replay_from_a_graph():
    """A prepared plan (phy_spr_plan_*) scores like the one-shot call bit for
    bit; captured in a HIP graph, replays after the branch lengths change in
    place equal eager runs at the new lengths, with eager work between."""
    from phylocaml_amd import device as D, likelihood as LK, spr, synth

    N, P = 40, 1500
    tree = synth.yule(N, 31, 0.08)
    model = _model(31)
    st, w = _data(tree, model, P, 31)
    mv = spr.moves(N, tree.children, tree.root)
    s = torch.cuda.Stream()
    ctx = D.Context(0, stream=s.cuda_stream)
    with torch.cuda.stream(s):
        dm = LK.DeviceModel(ctx, model)
        tips = D.to_device(synth.pack_dna(st))
        wd = D.to_device(w)
        for form in (F64, F32):
            one = spr.IncrementalScorer(ctx, dm, N, tips, D.words(P), wd, P, form=form, s_batch=6)
            ref0 = one.score(tree.children, tree.root, tree.blen, mv).cpu().numpy()
            plan = spr.SprPlan(ctx, dm, N, tips, D.words(P), wd, P, tree.children, tree.root, mv,
                               form=form, s_batch=6)
            bl = D.to_device(np.asarray(tree.blen, np.float64))
            plan.run(bl)
            ctx.sync()
            assert np.array_equal(plan.lnl.cpu().numpy(), ref0), form
            g = plan.capture(bl)
            rng = np.random.default_rng(form)
            for rep in range(3):
                new = np.asarray(tree.blen) * rng.uniform(0.5, 1.5, 2 * N - 1)
                new[tree.root] = 0.0
                bl.copy_(torch.from_numpy(new))
                torch.cuda.synchronize()
                g.replay()
                torch.cuda.synchronize()
                got = plan.lnl.cpu().numpy().copy()
                want = one.score(tree.children, tree.root, new, mv).cpu().numpy()  # eager
                print(f"form {form} replay {rep}: max |diff| {np.max(np.abs(got - want)):.3g}",
                      flush=True)
                assert np.array_equal(got, want), (form, rep)
            plan.close()
