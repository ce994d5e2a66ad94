"""Incremental TBR scoring (phy_lk_tbr_scores_incr, phy_tbr_plan_create) on
the GPU.

Per pruned subtree, one repair of the base tree's partials, one pre-order
pass over the subtree giving the root vector of every needed rerooting and
one pre-order pass over the pruned tree score every (target, rerooting)
pair, keeping the attachment (v = -1) included.  Checked against the FP64
oracle evaluating each candidate tree (phy_tbr_apply) in full -- FP64
within 1e-10 relative, FP32 within 1e-6 -- and against the per-candidate
scorer (phy_lk_tbr_scores) on every move of a larger tree; rate-class
counts 1, 2, 4 and 8, zero rate classes, caterpillars (the pruned tree a
single leaf at the root: only keeping moves), rerootings on the edges just
below s (accepted by phy_tbr_apply, not enumerated by phy_tbr_moves),
subtrees with more than 32 rerootings (several slots), moves in any order
with repeats, invalid moves, and a prepared plan replayed from a HIP
graph."""
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
    from phylocaml_amd import device as D, likelihood as LK, neighborhood as NB, synth

    P = st.shape[1]
    dm = LK.DeviceModel(ctx, model)
    tips = D.to_device(synth.pack_dna(st))
    wd = D.to_device(w)
    inc = NB.Scorer(ctx, dm, N, tips, D.words(P), wd, P, form=form, s_batch=s_batch,
                    incremental=True)
    full = NB.Scorer(ctx, dm, N, tips, D.words(P), wd, P, form=form, chunk=chunk,
                     incremental=False)
    return inc, full


def _oracle_all(N, ch, root, bl, model, st, w, mv):
    from phylocaml_amd import neighborhood as NB

    tips_m = (1 << st).astype(np.uint32)
    out = np.empty(len(mv))
    for k in range(len(mv)):
        c2, r2, b2 = NB.apply(N, ch, root, bl, mv[k])
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


def _child_reroots(N, ch, root, max_moves=40, seed=0):
    """TBR triples (s, w, v) with w a child of s: the rerooting on the edge
    just below s (w's half of the edge above it, the rest joined to s's
    other child), with v = -1 and with every SPR target of s."""
    from phylocaml_amd import spr

    spr_mv = spr.moves(N, ch, root)
    out = []
    for s in range(N, 2 * N - 1):
        if s == root:
            continue
        for w in ch[2 * (s - N): 2 * (s - N) + 2]:
            out.append((s, int(w), -1))
            out += [(s, int(w), int(v)) for v in spr_mv[spr_mv[:, 0] == s][:, 1]]
    out = np.array(out, np.int32).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    return out[rng.choice(len(out), min(max_moves, len(out)), replace=False)]


@pytest.mark.parametrize("form,tol", [(F64, 1e-10), (F32, 1e-6)])
def test_tbr_incremental_matches_oracle(ctx, form, tol):
    from phylocaml_amd import neighborhood as NB, synth

    N, P = 12, 257
    tree = synth.yule(N, 5, 0.08)
    model = _model(5)
    st, w = _data(tree, model, P, 5)
    mv = np.concatenate([NB.tbr_moves(N, tree.children, tree.root),
                         _child_reroots(N, tree.children, tree.root)])
    inc, _ = _scorers(ctx, model, st, w, N, form, s_batch=5)
    _say(f"tbr incremental 12 x 257 form {form}: {len(mv)} moves")
    got = inc.score(tree.children, tree.root, tree.blen, mv).cpu().numpy()
    ref = _oracle_all(N, tree.children, tree.root, tree.blen, model, st, w, mv)
    err = np.abs(got - ref) / np.abs(ref)
    assert err.max() <= tol, (int(err.argmax()), mv[err.argmax()], got[err.argmax()], ref[err.argmax()])
    base, _ = O.lnl(N, tree.children, tree.root, tree.blen, model.priors, model.q, model.rates,
                    model.probs, (1 << st).astype(np.uint32), w)
    assert abs(float(inc.base_lnl.item()) - base) <= tol * abs(base)


@pytest.mark.parametrize("rates,ref,kind", [("constant", False, "random"), ("gamma2", False, "yule"),
                                            ("gamma8", False, "random"), ("gamma4", True, "yule"),
                                            ("custom0", True, "caterpillar"),
                                            ("gamma4", False, "caterpillar")])
def test_tbr_incremental_rate_classes_and_shapes(ctx, rates, ref, kind):
    """R = 1, 2, 4, 8, the reference's own Gamma categories (rate 0 first)
    and an invariant class, on random, Yule and caterpillar trees (whose
    root joins a leaf to the rest: pruning the rest leaves a single leaf,
    so its moves only keep the attachment)."""
    from phylocaml_amd import neighborhood as NB, synth

    N, P = 14, 600
    ch, root, bl = _tree(kind, N, 11)
    model = _model(11, rates, ref)
    st, w = _data(synth.Tree(N, ch, root, bl), model, P, 11)
    mv = NB.tbr_moves(N, ch, root, 3, 3)
    if kind == "caterpillar":
        big = int(ch[2 * (root - N)])
        assert big >= N and np.any((mv[:, 0] == big) & (mv[:, 2] == -1))
    ref_v = _oracle_all(N, ch, root, bl, model, st, w, mv)
    for form, tol in ((F64, 1e-10), (F32, 1e-6)):
        inc, _ = _scorers(ctx, model, st, w, N, form, s_batch=4)
        _say(f"tbr incremental {rates} {kind} form {form}: {len(mv)} moves")
        got = inc.score(ch, root, bl, mv).cpu().numpy()
        err = np.abs(got - ref_v) / np.abs(ref_v)
        assert np.all(np.isfinite(got))
        assert err.max() <= tol, (rates, kind, form, mv[err.argmax()], got[err.argmax()],
                                  ref_v[err.argmax()])


def test_tbr_incremental_three_and_four_taxa(ctx):
    from phylocaml_amd import neighborhood as NB, synth

    model = _model(3)
    for ch, root, bl in (
            (np.array([1, 2, 0, 3], np.int32), 4, np.array([0.1, 0.2, 0.3, 0.05, 0.0])),
            (np.array([0, 1, 2, 3, 4, 5], np.int32), 6,
             np.array([0.1, 0.2, 0.3, 0.15, 0.05, 0.07, 0.0]))):
        N = len(ch) // 2 + 1
        st, w = _data(synth.Tree(N, ch, root, bl), model, 64, 3)
        mv = NB.tbr_moves(N, ch, root)
        mv = np.concatenate([mv, _child_reroots(N, ch, root)]) if N > 3 else mv
        if len(mv) == 0:
            continue
        ref_v = _oracle_all(N, ch, root, bl, model, st, w, mv)
        inc, _ = _scorers(ctx, model, st, w, N, F64)
        _say(f"tbr incremental {N} taxa: {len(mv)} moves")
        got = inc.score(ch, root, bl, mv).cpu().numpy()
        assert np.all(np.abs(got - ref_v) <= 1e-10 * np.abs(ref_v)), (mv, got, ref_v)


@pytest.mark.parametrize("form,tol", [(F64, 1e-10), (F32, 2e-6)])
def test_tbr_incremental_matches_full_scorer(ctx, form, tol):
    """TBR moves of a 48-taxon tree over 2,000 patterns (join distance <= 4,
    rerootings unlimited: subtrees near the root need several slots of 32),
    in shuffled order with repeats, 7 slots per batch, against the
    per-candidate scorer in FP64."""
    from phylocaml_amd import neighborhood as NB, synth

    N, P = 48, 2000
    tree = synth.yule(N, 21, 0.08)
    model = _model(21)
    st, w = _data(tree, model, P, 21)
    mv = NB.tbr_moves(N, tree.children, tree.root, 4, -1)
    rng = np.random.default_rng(0)
    mv = mv[rng.choice(len(mv), min(len(mv), 6000), replace=False)]
    mv2 = np.concatenate([mv, mv[rng.choice(len(mv), 300)]])
    mv2 = mv2[rng.permutation(len(mv2))]
    counts = np.unique(mv2[mv2[:, 1] != mv2[:, 0]][:, :2], axis=0)
    assert np.max(np.bincount(counts[:, 0])) > 32  # a subtree spans several slots
    inc, _ = _scorers(ctx, model, st, w, N, form, s_batch=7)
    _, full64 = _scorers(ctx, model, st, w, N, F64, chunk=1024)
    _say(f"tbr incremental vs full 48 x 2000 form {form}: {len(mv2)} moves")
    got = inc.score(tree.children, tree.root, tree.blen, mv2).cpu().numpy()
    ref_v = full64.score(tree.children, tree.root, tree.blen, mv2).cpu().numpy()
    err = np.abs(got - ref_v) / np.abs(ref_v)
    assert err.max() <= tol, (mv2[err.argmax()], got[err.argmax()], ref_v[err.argmax()])
    sub = mv2[::37]
    got2 = inc.score(tree.children, tree.root, tree.blen, sub).cpu().numpy()
    assert np.allclose(got2, got[::37], rtol=tol, atol=0)


def test_tbr_incremental_invalid_moves(ctx):
    from phylocaml_amd import _lib, neighborhood as NB, synth

    N, P = 10, 100
    tree = synth.yule(N, 2, 0.1)
    model = _model(2)
    st, w = _data(tree, model, P, 2)
    inc, _ = _scorers(ctx, model, st, w, N, F64)
    ch, root = tree.children, tree.root
    mv = NB.tbr_moves(N, ch, root)
    par = np.full(2 * N - 1, -1)
    par[ch] = np.repeat(np.arange(N, 2 * N - 1), 2)
    s = next(v for v in range(N, 2 * N - 1) if v != root and par[v] != root)
    leaf = 0
    outside = next(v for v in range(N) if not _under(ch, N, s, v))
    bad = [(root, root, -1), (s, s, -1), (leaf, (leaf + 1) % N, -1), (s, outside, -1),
           (s, s, int(par[s])), (s, s, int(ch[2 * (s - N)])), (-1, 0, 0), (s, 2 * N - 1, -1),
           (s, s, -2)]
    for b in bad:
        with pytest.raises(_lib.PhyError):
            inc.score(ch, root, tree.blen, np.concatenate([mv[:3], np.array([b], np.int32)]))
    got = inc.score(ch, root, tree.blen, mv[:5]).cpu().numpy()
    assert np.all(np.isfinite(got))


def _under(ch, N, s, v):
    st = [s]
    while st:
        y = st.pop()
        if y == v:
            return True
        if y >= N:
            st += [int(ch[2 * (y - N)]), int(ch[2 * (y - N) + 1])]
    return False


def test_tbr_plan_runs_equal_the_one_shot_scores_and_replay_from_a_graph():
    """A prepared TBR plan scores like the one-shot call bit for bit; captured
    in a HIP graph, replays after the branch lengths change in place equal
    eager one-shot scores at the new lengths."""
    from phylocaml_amd import device as D, likelihood as LK, neighborhood as NB, spr, synth

    N, P = 30, 1200
    tree = synth.yule(N, 31, 0.08)
    model = _model(31)
    st, w = _data(tree, model, P, 31)
    mv = NB.tbr_moves(N, tree.children, tree.root, 3, 2)
    s = torch.cuda.Stream()
    ctx = D.Context(0, stream=s.cuda_stream)
    with torch.cuda.stream(s):
        dm = LK.DeviceModel(ctx, model)
        tips = D.to_device(synth.pack_dna(st))
        wd = D.to_device(w)
        for form in (F64, F32):
            one = NB.Scorer(ctx, dm, N, tips, D.words(P), wd, P, form=form, s_batch=6,
                            incremental=True)
            ref0 = one.score(tree.children, tree.root, tree.blen, mv).cpu().numpy()
            plan = spr.SprPlan(ctx, dm, N, tips, D.words(P), wd, P, tree.children, tree.root, mv,
                               form=form, s_batch=6)
            bl = D.to_device(np.asarray(tree.blen, np.float64))
            plan.run(bl)
            ctx.sync()
            assert np.array_equal(plan.lnl.cpu().numpy()[: len(mv)], ref0), form
            g = plan.capture(bl)
            rng = np.random.default_rng(form)
            for rep in range(3):
                new = np.asarray(tree.blen) * rng.uniform(0.5, 1.5, 2 * N - 1)
                new[tree.root] = 0.0
                bl.copy_(torch.from_numpy(new))
                torch.cuda.synchronize()
                g.replay()
                torch.cuda.synchronize()
                got = plan.lnl.cpu().numpy()[: len(mv)].copy()
                want = one.score(tree.children, tree.root, new, mv).cpu().numpy()
                _say(f"tbr plan form {form} replay {rep}: max |diff| {np.max(np.abs(got - want)):.3g}")
                assert np.array_equal(got, want), (form, rep)
            plan.close()
