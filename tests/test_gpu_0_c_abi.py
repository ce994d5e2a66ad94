"""The C ABI driven from plain C (tests/c_abi/abi_demo.c, no Python in that
process): context, topology, Fitch down- and up-pass, likelihood model and
the whole-tree likelihood in three forms with branch derivatives, the call
sequence of the OCaml stubs.  Its outputs must equal the oracle's: Fitch
sets and costs bit for bit, FP64 lnL within 1e-10 (derivatives 1e-8), FP32
within 1e-6, the fused FP32 form identical to the stored one.

The module sorts first among the GPU tests: the harness is started as a child
process while this process has not yet initialised the GPU."""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

from conftest import random_tree

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "c_abi", "abi_demo")


def _case(tmp_path, seed, N, S, P):
    from phylocaml_amd import mlmodel as MM

    rng = np.random.default_rng(seed)
    ch, root = random_tree(rng, N)
    leaves = rng.integers(1, 16, (N, S)).astype(np.uint8)
    fw = rng.integers(1, 5, S).astype(np.uint32)
    pi = rng.dirichlet(np.full(4, 5.0))
    ex = rng.dirichlet(np.ones(6))
    model = MM.create(MM.Spec(4, ("GTR", (ex / ex[-1])[:-1]), ("DiscreteGamma", 4, 0.5),
                              ("Empirical", pi), MM.PHY_GAMMA_MEAN))
    tips = (1 << rng.integers(0, 4, (N, P))).astype(np.uint8)
    tips[rng.random((N, P)) < 0.05] = 15
    w = rng.integers(1, 6, P).astype(np.uint32)
    bl = rng.exponential(0.1, 2 * N - 1) + 1e-3
    bl[root] = 0.0
    R = len(model.rates)
    blob = b"".join([
        struct.pack("<5q", N, root, S, P, R),
        np.ascontiguousarray(ch, "<i4").tobytes(), leaves.tobytes(), fw.astype("<u4").tobytes(),
        np.asarray(model.priors, "<f8").tobytes(), np.asarray(model.q, "<f8").tobytes(),
        np.asarray(model.rates, "<f8").tobytes(), np.asarray(model.probs, "<f8").tobytes(),
        tips.tobytes(), w.astype("<u4").tobytes(), bl.astype("<f8").tobytes()])
    fin = tmp_path / f"in{seed}.bin"
    fin.write_bytes(blob)
    return dict(N=N, S=S, P=P, ch=ch, root=root, leaves=leaves, fw=fw, model=model, tips=tips,
                w=w, bl=bl, fin=fin, fout=tmp_path / f"out{seed}.bin")


def _read(c):
    N, S = c["N"], c["S"]
    M = 2 * N - 1
    b = c["fout"].read_bytes()
    off = 0

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(b, dtype=dtype, count=n, offset=off)
        off += a.nbytes
        return a

    out = {"total": int(take("<u8", 1)[0]), "site_cost": take("<u4", S),
           "prelim": take(np.uint8, (N - 1) * S).reshape(N - 1, S),
           "final": take(np.uint8, M * S).reshape(M, S), "lnl": take("<f8", 3),
           "d1": take("<f8", M), "d2": take("<f8", M)}
    assert off == len(b)
    return out


@pytest.mark.parametrize("seed,N,S,P", [(1, 2, 1, 1), (2, 60, 1000, 300), (3, 200, 777, 513)])
def test_plain_c_consumer_matches_oracle(tmp_path, seed, N, S, P):
    if torch.cuda.device_count() == 0:
        pytest.skip("no GPU")
    if not os.path.exists(EXE):
        pytest.skip("tests/c_abi/abi_demo not built (make -C tests/c_abi)")
    if torch.cuda.is_initialized():
        pytest.skip("start this module before anything in the process initialises the GPU")
    c = _case(tmp_path, seed, N, S, P)
    print(f"abi_demo seed={seed} N={N} S={S} P={P}", flush=True)
    r = subprocess.run([EXE, str(c["fin"]), str(c["fout"])], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    g = _read(c)
    osets, osc, _, otot = O.fitch_downpass(N, c["ch"], c["root"], c["leaves"], c["fw"])
    ofin = O.fitch_uppass(N, c["ch"], c["root"], osets)
    assert g["total"] == otot
    assert np.array_equal(g["site_cost"], osc)
    assert np.array_equal(g["prelim"], osets[N:])
    assert np.array_equal(g["final"], ofin)
    m = c["model"]
    ov, _, od1, od2 = O.lnl(N, c["ch"], c["root"], c["bl"], m.priors, m.q, m.rates, m.probs,
                            c["tips"].astype(np.uint32), c["w"], derivs=True)
    l64, l32, l32f = (float(x) for x in g["lnl"])
    assert abs(l64 - ov) <= 1e-10 * abs(ov), (l64, ov)
    assert abs(l32 - ov) <= 1e-6 * abs(ov), (l32, ov)
    assert l32f == l32
    scale = float(c["w"].sum())
    mask = np.arange(2 * N - 1) != c["root"]
    assert np.all(np.abs(g["d1"] - od1)[mask] <= 1e-8 * (np.abs(od1[mask]) + scale))
    assert np.all(np.abs(g["d2"] - od2)[mask] <= 1e-8 * (np.abs(od2[mask]) + scale))


AA_EXE = os.path.join(HERE, "c_abi", "aa_demo")


def test_plain_c_batched_20_state_matches_oracle(tmp_path):
    # the batched 20-state pruning driven from C (tests/c_abi/aa_demo.c):
    # per-tree lnL within the BF16 tolerance of the FP64 oracle, and a second
    # run of the same batch bit-identical to the first
    from phylocaml_amd import aa

    if torch.cuda.device_count() == 0:
        pytest.skip("no GPU")
    if not os.path.exists(AA_EXE):
        pytest.skip("tests/c_abi/aa_demo not built (make -C tests/c_abi)")
    if torch.cuda.is_initialized():
        pytest.skip("start this module before anything in the process initialises the GPU")
    rng = np.random.default_rng(17)
    shapes = [(2, 40), (33, 900), (70, 317), (129, 200), (17, 1)]
    trees = [aa.synth_tree(4000 + k, N, P, n_sim=150) for k, (N, P) in enumerate(shapes)]
    trees[2].tips[:, ::6] = 255
    R = len(trees[0].rates)
    cat = np.concatenate
    blob = b"".join([
        struct.pack("<3q", len(trees), R, 4),
        np.array([t.n_leaves for t in trees], "<i4").tobytes(),
        np.array([t.n_patterns for t in trees], "<i8").tobytes(),
        np.array([t.root for t in trees], "<i4").tobytes(),
        cat([np.asarray(t.children, "<i4") for t in trees]).tobytes(),
        cat([np.asarray(t.blen, "<f8") for t in trees]).tobytes(),
        cat([np.asarray(t.pi, "<f8") for t in trees]).tobytes(),
        cat([np.asarray(t.q, "<f8").ravel() for t in trees]).tobytes(),
        cat([np.asarray(t.rates, "<f8") for t in trees]).tobytes(),
        cat([np.asarray(t.probs, "<f8") for t in trees]).tobytes(),
        cat([np.ascontiguousarray(t.tips, np.uint8).ravel() for t in trees]).tobytes(),
        cat([np.asarray(t.weights, "<u4") for t in trees]).tobytes()])
    fin, fout = tmp_path / "aa_in.bin", tmp_path / "aa_out.bin"
    fin.write_bytes(blob)
    print(f"aa_demo trees={len(trees)} seed=17", flush=True)
    r = subprocess.run([AA_EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    b = fout.read_bytes()
    T = len(trees)
    pn = int(np.frombuffer(b, "<i8", 1)[0])
    lnl = np.frombuffer(b, "<f8", 2 * T, 8).reshape(2, T)
    assert pn == sum(t.n_patterns * (t.n_leaves - 1) for t in trees)
    assert np.array_equal(lnl[0], lnl[1])
    for k, t in enumerate(trees):
        ov, _ = O.lnl(t.n_leaves, t.children, t.root, t.blen, t.pi, t.q, t.rates, t.probs,
                      aa.tips_bitmask(t.tips), t.weights)
        assert abs(lnl[0, k] - ov) <= 2e-3 * abs(ov), (k, lnl[0, k], ov)


SK_EXE = os.path.join(HERE, "c_abi", "sankoff_demo")


def test_plain_c_sankoff_matches_oracle(tmp_path):
    # the whole-tree Sankoff pass driven from C (tests/c_abi/sankoff_demo.c):
    # per-site costs and the weighted total bit for bit
    if torch.cuda.device_count() == 0:
        pytest.skip("no GPU")
    if not os.path.exists(SK_EXE):
        pytest.skip("tests/c_abi/sankoff_demo not built (make -C tests/c_abi)")
    if torch.cuda.is_initialized():
        pytest.skip("start this module before anything in the process initialises the GPU")
    rng = np.random.default_rng(23)
    N, S, K = 150, 777, 5
    ch, root = random_tree(rng, N)
    m = rng.integers(1, 20, (K, K))
    tcm = (np.triu(m, 1) + np.tril(m, -1) * 2).astype(np.int32)  # asymmetric, zero diagonal
    leaves = (1 << rng.integers(0, K, (N, S))).astype(np.uint8)
    leaves[rng.random((N, S)) < 0.05] |= 3
    leaves[rng.random((N, S)) < 0.01] = 0  # missing: every state
    w = rng.integers(1, 6, S).astype(np.uint32)
    fin, fout = tmp_path / "sk_in.bin", tmp_path / "sk_out.bin"
    fin.write_bytes(struct.pack("<4q", N, root, S, K) + np.ascontiguousarray(ch, "<i4").tobytes()
                    + tcm.astype("<i4").tobytes() + leaves.tobytes() + w.astype("<u4").tobytes())
    print(f"sankoff_demo N={N} S={S} K={K}", flush=True)
    r = subprocess.run([SK_EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    b = fout.read_bytes()
    total = int(np.frombuffer(b, "<u8", 1)[0])
    sc = np.frombuffer(b, "<u4", S, 8)
    osc, otot = O.sankoff(N, ch, root, tcm, leaves, w)
    assert total == otot and np.array_equal(sc, osc)
