"""The character bookkeeping of NodeData.S (cardinal, get_codes, mem,
recode, filter_codes, filter_codes_comp, compare; lib/nodeData.ml:9-16, as
Node.Make1D calls them at lib/node.ml:93-145) on the Fitch and likelihood
values.  Pinned by what the operations must preserve: filtering commutes
with the per-site Fitch median and its cost equals the oracle's over the
kept sites; the lnL of the kept and of the dropped patterns add up to the
lnL of all of them; compare is a total order that agrees with a host
lexicographic comparison."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from phylocaml_amd import device as D

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return D.Context(0)


def test_fitch_values_filter_recode_compare(ctx, rng):
    from phylocaml_amd import nonadditive as NA

    print("nodedata fitch", flush=True)
    S = 1001
    a, b = rng.integers(1, 16, (2, S)).astype(np.uint8)
    w = rng.integers(1, 6, S).astype(np.uint32)
    codes = [int(c) for c in rng.permutation(5000)[:S]]
    x = NA.of_sets(ctx, a, w, codes)
    y = NA.of_sets(ctx, b, w, codes)
    assert NA.cardinal(x) == S and NA.get_codes(x) == set(codes)
    assert NA.mem(None, x) and NA.mem([codes[3], -1], x) and not NA.mem([-1], x)
    keep = set(codes[i] for i in rng.choice(S, 300, replace=False))
    sel = [i for i, c in enumerate(codes) if c in keep]
    fx, fy = NA.filter_codes(ctx, keep, x), NA.filter_codes(ctx, keep, y)
    assert NA.to_sets(fx).tolist() == a[sel].tolist() and NA.cardinal(fx) == len(sel)
    assert [codes[i] for i in sel] == fx.codes
    assert fx.weights.view(torch.int32).cpu().numpy().tolist() == w[sel].astype(np.int32).tolist()
    comp = NA.filter_codes_comp(ctx, keep, x)
    assert NA.cardinal(comp) + NA.cardinal(fx) == S and not (NA.get_codes(comp) & keep)
    assert NA.filter_codes(ctx, [-1], x) is None and NA.filter_codes_comp(ctx, codes, x) is None
    # an unweighted block, and an index order that is not sorted
    u = NA.of_sets(ctx, a, None, codes)
    rev = list(reversed(sel))
    fu = NA._select(ctx, u, rev)
    assert NA.to_sets(fu).tolist() == a[rev].tolist() and fu.weights is None
    # an index out of range is refused before anything reaches the GPU
    import ctypes as C

    from phylocaml_amd import device as D
    from phylocaml_amd._lib import lib, ptr

    bad = np.array([0, S], np.int64)
    out = torch.empty((1,), dtype=torch.uint64, device=x.states.device)
    assert lib.phy_fitch_select(ctx.handle, D.dptr(x.states), S, None, ptr(bad, C.c_int64), 2,
                                D.dptr(out), None) == 1  # PHY_EINVAL
    # filtering commutes with the per-site median; the cost is over the kept sites
    m_f = NA.median_2(ctx, None, None, fx, fy)
    m = NA.median_2(ctx, None, None, x, y)
    assert NA.to_sets(m_f).tolist() == NA.to_sets(NA.filter_codes(ctx, keep, m)).tolist()
    om, oc = O.fitch_median2(a[sel], b[sel], w[sel])
    assert NA.to_sets(m_f).tolist() == np.asarray(om).tolist() and m_f.cost == oc
    # recode keeps the states
    r = NA.recode(lambda c: c + 10_000, x)
    assert NA.get_codes(r) == {c + 10_000 for c in codes}
    assert NA.to_sets(r).tolist() == a.tolist()
    # compare: a total order agreeing with (codes, sets site by site)
    vals = [NA.of_sets(ctx, v, w, codes) for v in (a, b, np.where(np.arange(S) == 500, 15, a))]
    vals.append(NA.of_sets(ctx, a, w, [c + 1 for c in codes]))
    keys = [(v.codes, NA.to_sets(v).tolist()) for v in vals]
    for i, u in enumerate(vals):
        assert NA.compare(u, u) == 0
        for j, v in enumerate(vals):
            want = (keys[i] > keys[j]) - (keys[i] < keys[j])
            assert NA.compare(u, v) == want and NA.compare(v, u) == -want


def test_likelihood_values_filter_patterns(ctx, rng):
    from phylocaml_amd import likelihood as LK, mlmodel as MM

    print("nodedata likelihood", flush=True)
    P = 700
    pi = rng.dirichlet(np.full(4, 5.0))
    ex = rng.dirichlet(np.ones(6))
    model = MM.create(MM.Spec(4, ("GTR", (ex / ex[-1])[:-1]), ("DiscreteGamma", 4, 0.5),
                              ("Empirical", pi), MM.PHY_GAMMA_MEAN))
    dm = LK.DeviceModel(ctx, model)
    ta, tb = (1 << rng.integers(0, 4, (2, P))).astype(np.uint8)
    w = rng.integers(1, 6, P).astype(np.uint32)
    x = LK.leaf(ctx, dm, ta, 0.07, w)
    y = LK.leaf(ctx, dm, tb, 0.21, w)
    assert LK.cardinal(x) == P and LK.get_codes(x) == set(range(P))
    keep = set(int(i) for i in rng.choice(P, 250, replace=False))
    parts = []
    for f in (LK.filter_codes, LK.filter_codes_comp):
        fx, fy = f(keep, x), f(keep, y)
        root = LK.median_2(ctx, dm, None, fx, fy)
        parts.append(LK.root_lnl(ctx, dm, root))
    top = LK.median_2(ctx, dm, None, x, y)
    assert top.lnl is None
    c = LK.cost(top)  # evaluated on first use, through the model that made top
    full = LK.root_lnl(ctx, dm, top)
    assert c == -full == LK.cost(top) == LK.root_cost(ctx, dm, top)
    assert abs(parts[0] + parts[1] - full) <= 1e-11 * abs(full)
    # the same as the oracle's two-taxon tree on the kept patterns
    sel = sorted(keep)
    ov, _ = O.lnl(2, np.array([0, 1], np.int32), 2, np.array([0.07, 0.21, 0.0]), model.priors,
                  model.q, model.rates, model.probs, np.stack([ta[sel], tb[sel]]).astype(np.uint32),
                  w[sel])
    assert abs(parts[0] - ov) <= 1e-10 * abs(ov)
    r = LK.recode(lambda c: -c, x)
    assert LK.get_codes(r) == {-c for c in range(P)}
    assert LK.compare(x, x) == 0 and LK.compare(x, y) == -LK.compare(y, x) != 0
