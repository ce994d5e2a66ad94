"""phy_lk_tree_noalloc (the allocation-free fused lnL) through the C ABI:
two calls in a row, the second on a smaller tree and pattern set, each equal
bit for bit to phy_lk_tree on the same inputs (same kernels and launch
shapes over a workspace of the caller's), in FP64 and FP32; a workspace
beyond the fixed 8 MiB scratch, and the stored form, are refused."""
import ctypes as C
import sys

import numpy as np
import pytest

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


def _inputs(ctx, seed, N, P):
    from phylocaml_amd import device as D, likelihood as LK, mlmodel as MM, synth

    tree = synth.yule(N, seed, 0.1)
    pi, co = synth.gtr_params(4, 5.0, seed)
    model = MM.create(MM.Spec(4, ("GTR", co), ("DiscreteGamma", 4, 0.5), ("Empirical", pi)))
    st = synth.simulate(tree, model.priors, model.q, model.rates, P, seed)
    w = synth.poisson_weights(P, 3.0, seed)
    topo = D.Topology(N, tree.children, tree.root, ctx)
    dm = LK.DeviceModel(ctx, model)
    return {"topo": topo, "dm": dm, "tips": D.to_device(synth.pack_dna(st), ctx.device),
            "stride": D.words(P), "w": D.to_device(w, ctx.device),
            "bl": D.to_device(tree.blen, ctx.device), "P": P}


def _existing(ctx, x, form):
    from phylocaml_amd import _lib, device as D

    L = _lib.lib
    nb = C.c_size_t()
    _lib.check(L.phy_lk_tree_work_size(x["topo"].handle, x["dm"].handle, x["P"], form, 0,
                                       C.byref(nb)))
    work = torch.empty((nb.value,), dtype=torch.uint8, device=f"cuda:{ctx.device}")
    lnl = torch.zeros(1, dtype=torch.float64, device=f"cuda:{ctx.device}")
    site = torch.zeros(x["P"], dtype=torch.float64, device=f"cuda:{ctx.device}")
    ctx.check(L.phy_lk_tree(ctx.handle, x["topo"].handle, x["dm"].handle, form, x["P"],
                            D.dptr(x["tips"]), x["stride"], D.dptr(x["w"]), D.dptr(x["bl"]),
                            D.dptr(lnl), D.dptr(site), None, None, D.dptr(work), nb.value),
              "lk_tree")
    ctx.sync()
    return float(lnl.cpu()[0]), site.cpu().numpy()


def _noalloc(ctx, x, form, site=True):
    from phylocaml_amd import _lib, device as D

    lnl = torch.zeros(1, dtype=torch.float64, device=f"cuda:{ctx.device}")
    sl = torch.zeros(x["P"], dtype=torch.float64, device=f"cuda:{ctx.device}")
    rc = _lib.lib.phy_lk_tree_noalloc(ctx.handle, x["topo"].handle, x["dm"].handle, form, x["P"],
                                      D.dptr(x["tips"]), x["stride"], D.dptr(x["w"]),
                                      D.dptr(x["bl"]), D.dptr(lnl), D.dptr(sl) if site else None)
    if rc:
        return rc, None, None
    ctx.sync()
    return 0, float(lnl.cpu()[0]), sl.cpu().numpy()


@pytest.mark.parametrize("form", [2, 3])  # PHY_FUSED (FP64), PHY_FUSED | PHY_F32
def test_noalloc_equals_the_workspace_entry_point(ctx, form):
    big = _inputs(ctx, 3, 300, 20_003)
    small = _inputs(ctx, 4, 61, 777)
    for name, x in (("300 taxa x 20,003 patterns", big), ("61 taxa x 777 patterns", small)):
        _say(f"lk_tree_noalloc form={form}: {name}")
        rc, v, s = _noalloc(ctx, x, form)
        assert rc == 0
        ref, rs = _existing(ctx, x, form)
        assert v == ref, (v, ref)
        assert np.array_equal(s, rs)
    # a third call without site lnL, back on the larger input
    rc, v, _ = _noalloc(ctx, big, form, site=False)
    assert rc == 0 and v == _existing(ctx, big, form)[0]


def test_noalloc_refuses_what_it_cannot_hold(ctx):
    from phylocaml_amd import _lib, device as D, likelihood as LK, mlmodel as MM

    # 5,000-taxon caterpillar, FP64, 8 rate classes: 9,999 x 8 P matrices of
    # 128 bytes = 10.2 MB of workspace > 8 MiB
    N, P = 5000, 64
    ch = np.zeros(2 * (N - 1), np.int32)
    prev = 0
    for k in range(N - 1):
        ch[2 * k], ch[2 * k + 1] = prev, k + 1
        prev = N + k
    topo = D.Topology(N, ch, 2 * N - 2, ctx)
    model = MM.create(MM.Spec(4, ("GTR", np.ones(5)), ("DiscreteGamma", 8, 0.5), ("Equal",)))
    dm = LK.DeviceModel(ctx, model)
    tips = D.to_device(np.zeros((N, D.words(P)), np.uint64) + np.uint64(0x1111111111111111),
                       ctx.device)
    w = D.to_device(np.ones(P, np.uint32), ctx.device)
    bl = D.to_device(np.full(2 * N - 1, 0.1), ctx.device)
    lnl = torch.zeros(1, dtype=torch.float64, device=f"cuda:{ctx.device}")
    args = (ctx.handle, topo.handle, dm.handle)
    rest = (P, D.dptr(tips), D.words(P), D.dptr(w), D.dptr(bl), D.dptr(lnl), None)
    assert _lib.lib.phy_lk_tree_noalloc(*args, 2, *rest) == 5  # PHY_ERANGE
    assert b"exceeds" in _lib.lib.phy_last_error(ctx.handle)
    assert _lib.lib.phy_lk_tree_noalloc(*args, 0, *rest) == 1  # stored form: PHY_EINVAL
