"""The OCaml stubs (ocaml/phylo_hip_stubs.c) run on the GPU: stub_driver
calls every external of ocaml/nonAdditive_c.ml and ocaml/likelihood_c.ml
through a mock OCaml runtime (tests/mock_caml), building the modules'
records as OCaml lays them out, and its results are checked against the
oracle on the same seeded inputs: Fitch costs and sets bit for bit, FP64
lnL within 1e-10, derivatives within 1e-9 of their term magnitudes, the
composed P(t) within 1e-12."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "mock_caml", "stub_driver")
NF, SF, NL, SL, SEED = 57, 1003, 23, 1500, 7


def _fnv(a):
    h = 1469598103934665603
    for x in np.asarray(a, np.int64).tolist():
        h ^= x
        h = (h * 1099511628211) % (1 << 64)
    return h


@pytest.fixture(scope="module")
def out():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.access(DRIVER, os.X_OK):
        pytest.fail("tests/mock_caml/stub_driver is not built (run __graft_entry__.build())")
    print(f"stub_driver {NF} {SF} {NL} {SL} {SEED}", flush=True)
    sys.stdout.flush()
    r = subprocess.run([DRIVER, str(NF), str(SF), str(NL), str(SL), str(SEED)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    res = {}
    for line in r.stdout.splitlines():
        k, *v = line.split()
        res[k] = v
    assert "done" in res
    return res


def test_fitch_stubs_match_the_oracle(out):
    from phylocaml_amd import device as D, synth

    tree = synth.yule(NF, SEED, 0.1)
    leaves = D.unpack_nibbles(synth.fitch_leaves(NF, SF, SEED), SF)
    w = (1 + np.arange(SF) % 3).astype(np.uint32)
    sets, site_cost, node_cost, total = O.fitch_downpass(NF, tree.children, tree.root, leaves, w)
    fin = O.fitch_uppass(NF, tree.children, tree.root, sets)
    assert float(out["fitch_nodes_total"][0]) == total
    assert float(out["fitch_tree_total"][0]) == total
    assert int(out["fitch_site_cost_hash"][0]) == _fnv(site_cost)
    assert int(out["fitch_root_hash"][0]) == _fnv(sets[tree.root])
    assert float(out["fitch_prev_cost"][0]) == node_cost[tree.root - NF]
    hp = hf = 0
    for v in range(NF, 2 * NF - 1):
        hp = (hp * 31 + _fnv(sets[v])) % (1 << 64)
        hf = (hf * 31 + _fnv(fin[v])) % (1 << 64)
    assert int(out["fitch_prelim_hash"][0]) == hp
    assert int(out["fitch_final_hash"][0]) == hf
    a, b = tree.children[2 * (tree.root - NF)], tree.children[2 * (tree.root - NF) + 1]
    dist = float((w * ((sets[a] & sets[b]) == 0)).sum())
    assert float(out["fitch_distance1"][0]) == dist
    even = np.arange(0, SF, 2)
    sd = float((w[even] * ((leaves[0, even] & leaves[1, even]) == 0)).sum())
    assert float(out["fitch_select_distance1"][0]) == sd
    assert out["fitch_exception"] == ["2"]  # Invalid_argument on a site-count mismatch


def test_likelihood_stubs_match_the_oracle(out):
    from phylocaml_amd import mlmodel as MM, synth

    tree = synth.yule(NL, SEED, 0.1)
    pi, co = synth.gtr_params(4, 5.0, SEED)
    Q = MM.gtr_q(pi, co)
    rates = MM.gamma_rates(0.5, 0.5, 4, MM.PHY_GAMMA_MEAN)
    probs = np.full(4, 0.25)
    st = synth.simulate(tree, pi, Q, rates, SL, SEED)
    pats, w = synth.compress(st)
    P = pats.shape[1]
    assert int(out["lk_patterns"][0]) == P
    tips = (1 << pats).astype(np.uint32)
    ch, root, bl = tree.children, tree.root, tree.blen
    ov, _, od1, od2, a1, a2 = O.lnl(NL, ch, root, bl, pi, Q, rates, probs, tips, w,
                                    derivs=True, scales=True, threads=8)
    rel = lambda x: abs(float(x) - ov) / abs(ov)  # noqa: E731
    assert rel(out["lk_nodes_lnl"][0]) <= 1e-10
    assert rel(out["lk_tree_lnl"][0]) <= 1e-10
    assert rel(out["lk_tree_lnl_f32"][0]) <= 1e-6
    a, b = ch[2 * (root - NL)], ch[2 * (root - NL) + 1]
    f, g, h = (float(x) for x in out["lk_edge_root"])
    assert abs(f - ov) <= 1e-10 * abs(ov)
    assert abs(g - od1[a]) <= 1e-9 * a1[a] and abs(h - od2[a]) <= 1e-9 * a2[a]
    if a >= NL:
        assert rel(out["lk_median3"][0]) <= 1e-10
    d1 = np.array([float(x) for x in out["lk_tree_d1"]])
    d2 = np.array([float(x) for x in out["lk_tree_d2"]])
    mask = np.arange(2 * NL - 1) != root
    assert np.all(np.abs(d1 - od1)[mask] <= 1e-9 * a1[mask])
    assert np.all(np.abs(d2 - od2)[mask] <= 1e-9 * a2[mask])
    # adjust_3 of the root edge: the oracle's lnL at the new length agrees
    # and its derivative there vanishes
    t, lnl = (float(x) for x in out["lk_adjust3"])
    bl2 = bl.copy()
    bl2[a], bl2[b] = t, 0.0
    o2 = O.lnl(NL, ch, root, bl2, pi, Q, rates, probs, tips, w, derivs=True, scales=True,
               threads=8)
    assert abs(lnl - o2[0]) <= 1e-10 * abs(o2[0]) and o2[0] >= ov
    assert abs(o2[2][a]) <= 1e-6 * o2[4][a]
    # the MlModel stubs: diagonalize + compose = exp(Q t)
    Pm = np.array([float(x) for x in out["compose_0.3"]]).reshape(4, 4)
    assert np.allclose(Pm, O.compose(Q, 0.3), rtol=0, atol=1e-12)
    assert out["blocking_depth"] == ["0"]
