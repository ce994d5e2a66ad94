"""Branch-length Newton with its step control on the device
(phy_newton_propose, phy_lk_tree with derivatives, phy_newton_accept).

The eager device loop follows the host state machine of optimize_branches
(Newton proposal, or halving after a proposal that lowered lnL) bit for
bit: the same derivative kernels and the same FP64 step arithmetic.  A HIP
graph of k iterations replays to k eager iterations, with eager work and
synchronisations between the replays; run to convergence, the device loop
reaches the host optimiser's lnL."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _case(ctx, N=40, P=2000, seed=3):
    from phylocaml_amd import device as D, likelihood as LK, mlmodel as MM, synth

    tree = synth.yule(N, seed, 0.08)
    pi, co = synth.gtr_params(4, 5.0, seed)
    model = MM.create(MM.Spec(4, ("GTR", co), ("DiscreteGamma", 4, 0.5), ("Empirical", pi)))
    st = synth.simulate(tree, model.priors, model.q, model.rates, P, seed)
    w = synth.poisson_weights(P, 3.0, seed)
    topo = D.Topology(N, tree.children, tree.root, ctx)
    dm = LK.DeviceModel(ctx, model)
    tips = D.to_device(synth.pack_dna(st))
    wd = D.to_device(w)
    start = np.full(2 * N - 1, 0.1)
    start[tree.root] = 0.0
    return tree, model, topo, dm, tips, D.words(P), wd, P, start


def test_device_newton_follows_the_host_state_machine():
    from phylocaml_amd import device as D, optimize as OPT

    _gpu()
    ctx = D.Context(0)
    tree, model, topo, dm, tips, stride, wd, P, start = _case(ctx)
    N = tree.n_leaves
    dn = OPT.DeviceNewton(ctx, topo, dm, tips, stride, wd, P, start)
    obj = OPT.DeviceObjective(ctx, topo, tips, stride, wd, P)
    t = start.copy()
    lnl, d1, d2 = obj.derivs(model, t)
    t0, l0, _, _ = dn.result()
    assert np.array_equal(t0, t) and l0 == lnl
    newton, tn = True, t.copy()
    n_acc = n_half = 0
    for it in range(12):
        tn = OPT.newton_update(t, d1, d2, tree.children, N, tree.root) if newton else 0.5 * (t + tn)
        ln_n, d1n, d2n = obj.derivs(model, tn)
        if ln_n >= lnl:
            t, lnl, d1, d2, newton = tn, ln_n, d1n, d2n, True
            n_acc += 1
        else:
            newton = False
            n_half += 1
        dn.step()
        tg, lg, acc, half = dn.result()
        print(f"iteration {it}: lnL {lg!r} accepted {acc} halvings {half}", flush=True)
        assert np.array_equal(tg, t), it
        assert lg == lnl, (it, lg, lnl)
        assert (acc, half) == (n_acc, n_half)
    assert lnl > l0


def test_newton_graph_replays_equal_eager_iterations():
    from phylocaml_amd import device as D, optimize as OPT

    _gpu()
    s = torch.cuda.Stream()
    ctx = D.Context(0, stream=s.cuda_stream)
    with torch.cuda.stream(s):
        case = _case(ctx, N=64, P=3000, seed=5)
        _, _, topo, dm, tips, stride, wd, P, start = case
        eager = OPT.DeviceNewton(ctx, topo, dm, tips, stride, wd, P, start)
        graphed = OPT.DeviceNewton(ctx, topo, dm, tips, stride, wd, P, start)
        g = graphed.capture(4)
        # the capture recorded, it did not run
        assert np.array_equal(graphed.result()[0], eager.result()[0])
        for rep in range(3):
            g.replay()
            torch.cuda.synchronize()
            eager.run(4)  # eager work between the replays
            te, le, ae, he = eager.result()
            tg, lg, ag, hg = graphed.result()
            print(f"replay {rep}: lnL {lg!r} accepted {ag} halvings {hg}", flush=True)
            assert np.array_equal(tg, te) and lg == le and (ag, hg) == (ae, he), rep


def test_device_newton_reaches_the_host_optimum():
    from phylocaml_amd import device as D, optimize as OPT

    _gpu()
    ctx = D.Context(0)
    tree, model, topo, dm, tips, stride, wd, P, start = _case(ctx, N=48, P=2500, seed=9)
    obj = OPT.DeviceObjective(ctx, topo, tips, stride, wd, P)
    host = OPT.optimize_branches(obj, model, start, tree.children, tree.n_leaves, tree.root,
                                 max_iter=40)
    dn = OPT.DeviceNewton(ctx, topo, dm, tips, stride, wd, P, start)
    dn.run(40)
    _, lg, acc, half = dn.result()
    print(f"host {host.lnl!r} ({host.derivative_passes} passes), device {lg!r} "
          f"({acc} accepted, {half} halvings)", flush=True)
    assert lg >= host.lnl - 1e-9 * abs(host.lnl)
