"""GPU parity of the pruning kernels against the FP64 oracle.

Tolerances: FP64 forms match the oracle lnL within 1e-10 relative; the
FP32 forms (FP32 partials with int32 power-of-two exponents) within 1e-6
relative.  Derivatives are sums of per-pattern terms w_p g_p (first) and
w_p h_p (second), each term carrying a relative rounding error; they are
checked per branch against that branch's own term magnitudes
sum_p w_p |g_p| and sum_p w_p (|L''/L| + g_p^2) (from the oracle; h is a
difference, so its rounding error scales with its parts): FP64 within 1e-11
of them, FP32 within D32 = 1e-5."""
import os
import sys

import numpy as np
import pytest

from conftest import caterpillar, random_tree
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

F64, F32, FUSED = 0, 1, 2
D32 = 1e-5  # FP32 derivative tolerance, relative to the branch's term magnitudes
D64 = 1e-11


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from phylocaml_amd import device as D

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return D.Context(0)


def _model(rng, R):
    from phylocaml_amd import mlmodel as MM

    pi = rng.dirichlet(np.full(4, 5.0))
    ex = rng.dirichlet(np.ones(6))
    spec = MM.Spec(4, ("GTR", (ex / ex[-1])[:-1]), ("DiscreteGamma", R, 0.5),
                   ("Empirical", pi), MM.PHY_GAMMA_MEAN)
    if R == 1:
        spec.site_variation = ("Constant",)
    return MM.create(spec)


def _tips(rng, N, P):
    t = (1 << rng.integers(0, 4, (N, P))).astype(np.uint8)
    amb = rng.random((N, P)) < 0.1
    t[amb] |= (1 << rng.integers(0, 4, amb.sum())).astype(np.uint8)
    t[rng.random((N, P)) < 0.03] = 15
    return t


def _run(ctx, N, ch, root, bl, model, tips, w, form, derivs):
    from phylocaml_amd import device as D, likelihood as LK

    topo = D.Topology(N, ch, root, ctx)
    dm = LK.DeviceModel(ctx, model)
    P = tips.shape[1]
    ev = LK.Evaluator(ctx, topo, dm, P, form, derivs=derivs)
    tp = D.to_device(D.pack_nibbles(tips), ctx.device)
    ev.run(tp, D.words(P), D.to_device(w, ctx.device), D.to_device(bl, ctx.device),
           site_lnl=True)
    ctx.sync()
    out = {"lnl": float(ev.lnl.cpu()[0]), "site": ev.site_lnl.cpu().numpy()}
    if derivs:
        out["d1"], out["d2"] = ev.d1.cpu().numpy(), ev.d2.cpu().numpy()
    out["ev"] = ev
    return out


def _oracle(N, ch, root, bl, model, tips, w, derivs, scales=False):
    return O.lnl(N, ch, root, bl, model.priors, model.q, model.rates, model.probs,
                 tips.astype(np.uint32), w, derivs=derivs, threads=min(16, os.cpu_count() or 1),
                 scales=scales)


def _check_derivs(got, od1, od2, a1, a2, mask, tol, what):
    """|d - oracle| <= tol * (sum_p w_p |term_p|) on every non-root branch;
    prints the largest ratio so a run shows the margin."""
    e1 = np.abs(got["d1"] - od1)[mask] / np.maximum(a1[mask], 1e-300)
    e2 = np.abs(got["d2"] - od2)[mask] / np.maximum(a2[mask], 1e-300)
    _say(f"{what}: max d1 err / scale {e1.max():.3e}, d2 {e2.max():.3e} (tol {tol:g})")
    assert e1.max() <= tol, (what, "d1", float(e1.max()))
    assert e2.max() <= tol, (what, "d2", float(e2.max()))


@pytest.mark.parametrize("N", [2, 3, 9, 64, 300])
@pytest.mark.parametrize("R", [1, 4, 8])
def test_lnl_and_derivatives_match_oracle(ctx, rng, N, R):
    for kind in ("random", "caterpillar"):
        _say(f"prune N={N} R={R} {kind}")
        ch, root = random_tree(rng, N) if kind == "random" else caterpillar(N)
        P = int(rng.integers(1, 700))
        bl = rng.exponential(0.1, 2 * N - 1) + 1e-3
        bl[root] = 0.0
        model = _model(rng, R)
        tips = _tips(rng, N, P)
        w = rng.integers(1, 6, P).astype(np.uint32)
        ov, osl, od1, od2, a1, a2 = _oracle(N, ch, root, bl, model, tips, w, True, True)
        g = _run(ctx, N, ch, root, bl, model, tips, w, F64, True)
        assert abs(g["lnl"] - ov) <= 1e-10 * abs(ov), (g["lnl"], ov)
        assert np.allclose(g["site"], osl, rtol=1e-11, atol=1e-11)
        mask = np.arange(2 * N - 1) != root
        _check_derivs(g, od1, od2, a1, a2, mask, D64, f"FP64 N={N} R={R}")
        g32 = _run(ctx, N, ch, root, bl, model, tips, w, F32, True)
        assert abs(g32["lnl"] - ov) <= 1e-6 * abs(ov), (g32["lnl"], ov)
        _check_derivs(g32, od1, od2, a1, a2, mask, D32, f"FP32 N={N} R={R}")
        fu32 = _run(ctx, N, ch, root, bl, model, tips, w, F32 | FUSED, False)
        assert fu32["lnl"] == g32["lnl"]  # same arithmetic, operands from LDS
        fu64 = _run(ctx, N, ch, root, bl, model, tips, w, F64 | FUSED, False)
        assert fu64["lnl"] == g["lnl"]


@pytest.mark.parametrize("rpw,reg", [("1", "0"), ("1", "1"), ("1", "2"), ("2", "0")])
@pytest.mark.parametrize("R", [2, 4, 8])
def test_rate_classes_per_wave(ctx, rng, monkeypatch, R, rpw, reg):
    # single precision evaluates one or two rate classes per wave, and the
    # fused stack keeps 0, 1 or 2 of its levels in registers: every layout,
    # stored and fused, gives the oracle's lnL and site lnL
    monkeypatch.setenv("PHY_PRUNE_RPW", rpw)
    monkeypatch.setenv("PHY_PRUNE_REG", reg)
    N = 150
    _say(f"prune rate classes per wave R={R} rpw={rpw} register levels={reg}")
    ch, root = random_tree(rng, N)
    P = 1000
    bl = rng.exponential(0.1, 2 * N - 1) + 1e-3
    bl[root] = 0.0
    model = _model(rng, R)
    tips = _tips(rng, N, P)
    w = rng.integers(1, 6, P).astype(np.uint32)
    ov, osl, od1, od2, a1, a2 = _oracle(N, ch, root, bl, model, tips, w, True, True)
    mask = np.arange(2 * N - 1) != root
    g32 = _run(ctx, N, ch, root, bl, model, tips, w, F32, True)
    assert abs(g32["lnl"] - ov) <= 1e-6 * abs(ov), (g32["lnl"], ov)
    assert np.allclose(g32["site"], osl, rtol=1e-5, atol=1e-5)
    _check_derivs(g32, od1, od2, a1, a2, mask, D32, f"FP32 R={R} rpw={rpw}")
    fu32 = _run(ctx, N, ch, root, bl, model, tips, w, F32 | FUSED, False)
    assert fu32["lnl"] == g32["lnl"]
    assert np.array_equal(fu32["site"], g32["site"])
    g64 = _run(ctx, N, ch, root, bl, model, tips, w, F64, False)
    fu64 = _run(ctx, N, ch, root, bl, model, tips, w, F64 | FUSED, False)
    assert abs(fu64["lnl"] - ov) <= 1e-10 * abs(ov)
    assert fu64["lnl"] == g64["lnl"]
    assert np.array_equal(fu64["site"], g64["site"])


def test_stored_partials_match_oracle_pieces(ctx, rng):
    # the partial of a cherry is (P_a t_a) * (P_b t_b): check one node directly
    from phylocaml_amd import device as D

    N = 5
    ch, root = caterpillar(N)
    bl = rng.exponential(0.2, 2 * N - 1)
    model = _model(rng, 4)
    tips = _tips(rng, N, 200)
    w = np.ones(200, np.uint32)
    g = _run(ctx, N, ch, root, bl, model, tips, w, F64, False)
    part, ex = g["ev"].partial(N)  # node N = (leaf 0, leaf 1)
    part = part.cpu().numpy()[:, :200]
    ex = ex.cpu().numpy()[:, :200]
    for r in range(4):
        Pa = O.compose(model.q, model.rates[r] * bl[0])
        Pb = O.compose(model.q, model.rates[r] * bl[1])
        ta = (tips[0][None, :] >> np.arange(4)[:, None]) & 1
        tb = (tips[1][None, :] >> np.arange(4)[:, None]) & 1
        exp = (Pa @ ta) * (Pb @ tb)
        got = part[4 * r: 4 * r + 4] * np.ldexp(1.0, ex[r])[None, :]
        assert np.allclose(got, exp, rtol=1e-12, atol=1e-15 * np.abs(exp).max())


def test_node_api_matches_tree(ctx, rng):
    from phylocaml_amd import likelihood as LK

    N = 3
    ch = np.array([0, 1, 3, 2], np.int32)  # node 3 = (0,1), root 4 = (3,2)
    root = 4
    bl = np.array([0.1, 0.2, 0.3, 0.15, 0.0])
    model = _model(rng, 4)
    P = 333
    tips = _tips(rng, N, P)
    w = rng.integers(1, 4, P).astype(np.uint32)
    dm = LK.DeviceModel(ctx, model)
    for form in (F64, F32):
        l0 = LK.leaf(ctx, dm, tips[0], bl[0], w, form)
        l1 = LK.leaf(ctx, dm, tips[1], bl[1], w, form)
        l2 = LK.leaf(ctx, dm, tips[2], bl[2], w, form)
        n3 = LK.median_2(ctx, dm, None, l0, l1, blen=bl[3])
        rt = LK.median_2(ctx, dm, None, n3, l2)
        v = LK.root_lnl(ctx, dm, rt)
        ov, _ = _oracle(N, ch, root, bl, model, tips, w, False)
        tol = 1e-10 if form == F64 else 1e-6
        assert abs(v - ov) <= tol * abs(ov)
        assert LK.root_cost(ctx, dm, rt) == -v


def test_512_taxa_200k_patterns_full_size(ctx):
    """512-taxon Yule tree, 200,000 unique GTR+G4 patterns (the pruning
    benchmark workload): FP64 lnL and derivatives vs the oracle, FP32-scaled
    lnL error vs FP64."""
    from phylocaml_amd import device as D, likelihood as LK, mlmodel as MM, synth

    N, P = 512, 200_000
    _say("512 taxa x 200,000 patterns: generating")
    tree = synth.yule(N, 2, 0.05)
    pi, co = synth.gtr_params(4, 5.0, 2)
    spec = MM.Spec(4, ("GTR", co), ("DiscreteGamma", 4, 0.5), ("Empirical", pi))
    model = MM.create(spec)
    st = synth.simulate(tree, model.priors, model.q, model.rates, P, 2, unique=True)
    w = synth.poisson_weights(P, 3.0, 2)
    tips = np.where(st < 4, 1 << np.minimum(st, 3), 15).astype(np.uint8)
    _say("512 taxa x 200,000 patterns: GPU")
    g64 = _run(ctx, N, tree.children, tree.root, tree.blen, model, tips, w, F64, True)
    g32s = _run(ctx, N, tree.children, tree.root, tree.blen, model, tips, w, F32, True)
    g32 = _run(ctx, N, tree.children, tree.root, tree.blen, model, tips, w, F32 | FUSED, False)
    _say("512 taxa x 200,000 patterns: oracle")
    ov, osl, od1, od2, a1, a2 = _oracle(N, tree.children, tree.root, tree.blen, model, tips, w,
                                        True, True)
    _say(f"lnL oracle {ov!r} fp64 {g64['lnl']!r} fp32 {g32['lnl']!r}")
    assert abs(g64["lnl"] - ov) <= 1e-10 * abs(ov)
    assert abs(g32["lnl"] - ov) <= 1e-6 * abs(ov)
    assert g32s["lnl"] == g32["lnl"]
    mask = np.arange(2 * N - 1) != tree.root
    _check_derivs(g64, od1, od2, a1, a2, mask, D64, "512 x 200k FP64")
    _check_derivs(g32s, od1, od2, a1, a2, mask, D32, "512 x 200k FP32")


def _ref_gamma_workload(kind):
    """512-taxon Yule tree, 200,000 patterns simulated under the reference's
    own Gamma categories (quantile(i/4) of Gamma(0.5, scale 0.5): rate_0 = 0,
    lib/mlModel.ml:93-99) or a custom distribution with an invariant class
    (rate 0, probability 0.2)."""
    from phylocaml_amd import mlmodel as MM, synth

    N, P = 512, 200_000
    tree = synth.yule(N, 5, 0.05)
    pi, co = synth.gtr_params(4, 5.0, 5)
    if kind == "ref":
        sv = ("DiscreteGamma", 4, 0.5)
    else:
        sv = ("DiscreteCustom", [(0.0, 0.2), (0.5, 0.3), (1.2, 0.3), (2.5, 0.2)])
    model = MM.create(MM.Spec(4, ("GTR", co), sv, ("Empirical", pi), MM.PHY_GAMMA_REF))
    assert model.rates[0] == 0.0
    # simulate under positive rates (every pattern then varies somewhere)
    sim_rates = np.where(model.rates > 0, model.rates, 1.0)
    st = synth.simulate(tree, model.priors, model.q, sim_rates, P, 5, unique=True)
    w = synth.poisson_weights(P, 3.0, 5)
    tips = np.where(st < 4, 1 << np.minimum(st, 3), 15).astype(np.uint8)
    return N, tree, model, tips, w


@pytest.mark.parametrize("kind", ["ref", "custom"])
def test_zero_rate_class_512_taxa_200k_patterns(ctx, kind):
    """A rate class of rate 0 (P = I) leaves an all-zero class partial at
    every non-constant pattern; the other classes keep their own exponents
    (down to about -245 at this depth), so the classes of a pattern must be
    aligned over the non-zero ones only.  Every form against the oracle:
    lnL, and d1/d2 from the stored forms."""
    N, tree, model, tips, w = _ref_gamma_workload(kind)
    _say(f"zero-rate class ({kind}) 512 x 200,000: GPU")
    args = (N, tree.children, tree.root, tree.blen, model, tips, w)
    g64 = _run(ctx, *args, F64, True)
    g32 = _run(ctx, *args, F32, True)
    f64 = _run(ctx, *args, F64 | FUSED, False)
    f32 = _run(ctx, *args, F32 | FUSED, False)
    _say(f"zero-rate class ({kind}) 512 x 200,000: oracle")
    ov, osl, od1, od2, a1, a2 = _oracle(*args, True, True)
    _say(f"lnL oracle {ov!r} fp64 {g64['lnl']!r} fp32 {g32['lnl']!r}")
    assert np.all(np.isfinite(osl))
    for g, tol in ((g64, 1e-10), (f64, 1e-10), (g32, 1e-6), (f32, 1e-6)):
        assert np.isfinite(g["lnl"]) and abs(g["lnl"] - ov) <= tol * abs(ov), (g["lnl"], ov)
    assert np.allclose(g64["site"], osl, rtol=1e-10, atol=1e-9)
    assert np.allclose(g32["site"], osl, rtol=1e-5, atol=1e-4)
    mask = np.arange(2 * N - 1) != tree.root
    _check_derivs(g64, od1, od2, a1, a2, mask, D64, f"zero-rate {kind} FP64")
    _check_derivs(g32, od1, od2, a1, a2, mask, D32, f"zero-rate {kind} FP32")


def test_argument_errors(ctx, rng):
    from phylocaml_amd import _lib, device as D, likelihood as LK

    N = 4
    ch, root = caterpillar(N)
    topo = D.Topology(N, ch, root, ctx)
    dm = LK.DeviceModel(ctx, _model(rng, 4))
    with pytest.raises(_lib.PhyError):
        LK.Evaluator(ctx, topo, dm, 10, F64 | FUSED, derivs=True)
    ev = LK.Evaluator(ctx, topo, dm, 10, F64)
    buf = torch.zeros(64, dtype=torch.uint64, device="cuda:0")
    rc = _lib.lib.phy_lk_tree(ctx.handle, topo.handle, dm.handle, F64, 10, D.dptr(buf), 1,
                              D.dptr(buf), D.dptr(buf), D.dptr(ev.lnl), None, None, None,
                              D.dptr(ev.work), 16)
    assert rc == 1  # workspace too small


def test_unsupported_rate_class_count_is_refused(ctx, rng):
    # three rate classes: the model is valid (1..8), the pruning passes take
    # 1, 2, 4 or 8 and refuse the evaluation with PHY_ERANGE before launching
    from phylocaml_amd import _lib, device as D, likelihood as LK

    N, P = 5, 10
    ch, root = caterpillar(N)
    topo = D.Topology(N, ch, root, ctx)
    dm = LK.DeviceModel(ctx, _model(rng, 3))
    tips = D.to_device(D.pack_nibbles(_tips(rng, N, P)), ctx.device)
    w = D.to_device(np.ones(P, np.uint32), ctx.device)
    bl = D.to_device(np.full(2 * N - 1, 0.1), ctx.device)
    with pytest.raises(_lib.PhyError) as e:
        ev = LK.Evaluator(ctx, topo, dm, P, F64)
        ev.run(tips, D.words(P), w, bl)
    assert e.value.code == 5


@pytest.mark.parametrize("parts", ["2", "8", "16"])
@pytest.mark.parametrize("R", [1, 2, 4, 8])
def test_split_form_is_bit_identical(ctx, rng, monkeypatch, R, parts):
    # the split form (parts side by side, then the top program reading the
    # parts' root partials) performs the whole-tree pass's arithmetic in the
    # same order: its lnL and site lnL equal the whole-tree fused form's bit
    # for bit, in both precisions, and the oracle's within tolerance
    monkeypatch.setenv("PHY_PRUNE_PARTS", parts)
    N = 1100
    _say(f"prune split form R={R} parts<={parts} N={N}")
    ch, root = random_tree(rng, N)
    P = 777
    bl = rng.exponential(0.1, 2 * N - 1) + 1e-3
    bl[root] = 0.0
    model = _model(rng, R)
    tips = _tips(rng, N, P)
    w = rng.integers(1, 6, P).astype(np.uint32)
    ov = _oracle(N, ch, root, bl, model, tips, w, False)[0]
    for form, tol in ((F32 | FUSED, 1e-6), (F64 | FUSED, 1e-10)):
        monkeypatch.setenv("PHY_PRUNE_SPLIT", "0")
        whole = _run(ctx, N, ch, root, bl, model, tips, w, form, False)
        monkeypatch.setenv("PHY_PRUNE_SPLIT", "1")
        split = _run(ctx, N, ch, root, bl, model, tips, w, form, False)
        assert split["lnl"] == whole["lnl"], (form, split["lnl"], whole["lnl"])
        assert np.array_equal(split["site"], whole["site"])
        assert abs(split["lnl"] - ov) <= tol * abs(ov), (form, split["lnl"], ov)


@pytest.mark.parametrize("P", [1, 63, 65, 300, 4099])
def test_split_form_small_and_ragged_pattern_counts(ctx, rng, monkeypatch, P):
    # pattern counts below, at and past a 64-pattern tile and the 256-pattern
    # pitch: the split form's padded grids and XCD-ordered tiles cover every
    # tile once, so it still equals the whole-tree form bit for bit
    N = 400
    _say(f"prune split form ragged P={P}")
    ch, root = random_tree(rng, N)
    bl = rng.exponential(0.1, 2 * N - 1) + 1e-3
    bl[root] = 0.0
    model = _model(rng, 4)
    tips = _tips(rng, N, P)
    w = rng.integers(1, 6, P).astype(np.uint32)
    ov = _oracle(N, ch, root, bl, model, tips, w, False)[0]
    for form, tol in ((F32 | FUSED, 1e-6), (F64 | FUSED, 1e-10)):
        monkeypatch.setenv("PHY_PRUNE_SPLIT", "0")
        whole = _run(ctx, N, ch, root, bl, model, tips, w, form, False)
        monkeypatch.setenv("PHY_PRUNE_SPLIT", "1")
        split = _run(ctx, N, ch, root, bl, model, tips, w, form, False)
        assert split["lnl"] == whole["lnl"], (form, split["lnl"], whole["lnl"])
        assert np.array_equal(split["site"], whole["site"])
        assert abs(split["lnl"] - ov) <= tol * abs(ov), (form, split["lnl"], ov)


@pytest.mark.parametrize("split", ["0", "1"])
@pytest.mark.parametrize("R", [1, 2, 4, 8])
def test_stored_form_stack_in_lds_matches_hbm_stack(ctx, rng, monkeypatch, R, split):
    # the stored form with its operand stack in LDS (whole tree, or split:
    # the parts side by side writing every partial, then the top reading the
    # part roots from their rows) writes the same partials, exponents, lnL
    # and derivatives as the form that reads children back from HBM, bit for
    # bit
    monkeypatch.setenv("PHY_PRUNE_SPLIT", split)
    N = 300
    P = 1000
    _say(f"prune stored form, LDS stack vs HBM, R={R} split={split}")
    ch, root = random_tree(rng, N)
    bl = rng.exponential(0.1, 2 * N - 1) + 1e-3
    bl[root] = 0.0
    model = _model(rng, R)
    tips = _tips(rng, N, P)
    w = rng.integers(1, 6, P).astype(np.uint32)
    for form in (F32, F64):
        got = {}
        for flag in ("0", "1"):
            monkeypatch.setenv("PHY_PRUNE_STORED_LDS", flag)
            monkeypatch.setenv("PHY_PRUNE_RPW", "1")
            got[flag] = _run(ctx, N, ch, root, bl, model, tips, w, form, True)
        a, b = got["0"], got["1"]
        assert a["lnl"] == b["lnl"]
        assert np.array_equal(a["site"], b["site"])
        assert np.array_equal(a["d1"], b["d1"]) and np.array_equal(a["d2"], b["d2"])
        for v in (N, N + N // 2, 2 * N - 2):
            pa, ea = a["ev"].partial(v)
            pb, eb = b["ev"].partial(v)
            assert torch.equal(pa[:, :P].cpu(), pb[:, :P].cpu())
            assert torch.equal(ea[:, :P].cpu(), eb[:, :P].cpu())
