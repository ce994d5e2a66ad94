"""Whole-tree passes captured in a HIP graph (torch.cuda.CUDAGraph on a side
stream bound to the library's context) replay to the eager results, with
eager work and synchronisations between the replays.  The library's kernels
clear their accumulators with fill kernels, never hipMemsetAsync, so the
captured graphs hold kernel nodes only."""
import numpy as np
import pytest

from conftest import random_tree

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_pruning_graph_replays_match_eager(rng):
    from phylocaml_amd import device as D, likelihood as LK, mlmodel as MM

    _gpu()
    s = torch.cuda.Stream()
    ctx = D.Context(0, stream=s.cuda_stream)
    N, P = 200, 3000
    ch, root = random_tree(rng, N)
    bl = rng.exponential(0.1, 2 * N - 1) + 1e-3
    bl[root] = 0.0
    pi = rng.dirichlet(np.full(4, 5.0))
    ex = rng.dirichlet(np.ones(6))
    model = MM.create(MM.Spec(4, ("GTR", (ex / ex[-1])[:-1]), ("DiscreteGamma", 4, 0.5),
                              ("Empirical", pi), MM.PHY_GAMMA_MEAN))
    tips_u8 = (1 << rng.integers(0, 4, (N, P))).astype(np.uint8)
    w = rng.integers(1, 5, P).astype(np.uint32)
    with torch.cuda.stream(s):
        topo = D.Topology(N, ch, root, ctx)
        dm = LK.DeviceModel(ctx, model)
        tips = D.to_device(D.pack_nibbles(tips_u8))
        wd, bld = D.to_device(w), D.to_device(bl)
        for form, derivs in ((3, False), (1, True), (0, True)):
            ev = LK.Evaluator(ctx, topo, dm, P, form, derivs=derivs)
            ev.run(tips, D.words(P), wd, bld)
            ctx.sync()
            ref = (ev.lnl.cpu().numpy().copy(),
                   None if not derivs else ev.d1.cpu().numpy().copy())
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                ev.run(tips, D.words(P), wd, bld)
            for k in range(4):
                ev.lnl.zero_()
                if derivs:
                    ev.d1.zero_()
                g.replay()
                torch.cuda.synchronize()
                assert np.array_equal(ev.lnl.cpu().numpy(), ref[0]), (form, k)
                if derivs:
                    assert np.array_equal(ev.d1.cpu().numpy(), ref[1]), (form, k)
                # eager work between the replays
                ev.run(tips, D.words(P), wd, bld)
                ctx.sync()


def test_fitch_graph_replays_match_eager(rng):
    from phylocaml_amd import device as D, nonadditive as NA

    _gpu()
    s = torch.cuda.Stream()
    ctx = D.Context(0, stream=s.cuda_stream)
    N, S = 300, 20000  # split into parts: two launches per pass
    ch, root = random_tree(rng, N)
    leaves = rng.integers(1, 16, (N, S)).astype(np.uint8)
    with torch.cuda.stream(s):
        topo = D.Topology(N, ch, root, ctx)
        ft = NA.alloc_tree(ctx, topo, S, D.pack_nibbles(leaves), with_final=True)
        NA.downpass(ctx, topo, ft)  # eager first: allocates the split scratch
        NA.uppass(ctx, topo, ft)
        ctx.sync()
        ref_sets = ft.sets.cpu().numpy().copy()
        ref_fin = ft.final.cpu().numpy().copy()
        ref_tot = int(ft.total.cpu()[0])
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            NA.downpass(ctx, topo, ft)
            NA.uppass(ctx, topo, ft)
        for k in range(4):
            ft.sets[N:].zero_()
            ft.final.zero_()
            g.replay()
            torch.cuda.synchronize()
            assert int(ft.total.cpu()[0]) == ref_tot, k
            assert np.array_equal(ft.sets.cpu().numpy(), ref_sets), k
            assert np.array_equal(ft.final.cpu().numpy(), ref_fin), k
            NA.downpass(ctx, topo, ft)
            ctx.sync()
